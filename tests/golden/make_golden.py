"""Writes tests/golden/golden.npz: high-precision (mpmath, 40 digits) values
that pin the FP64 oracle independently of its own recurrences.

  legendre : P_l^m(x) with the Condon-Shortley phase, l <= 12, 7 points
  harmonic : Y_l^m(theta, phi), l <= 10, 5 directions
  wigner   : d^l_{m1,m2}(beta) by Wigner's explicit sum, l <= 12, 4 angles
  direct   : potential and gradient of 48 seeded charges (direct sum)

The reference ships no numeric fixtures (its tests compare against closed
forms at run time), so these values come from the mathematics alone.
Run: python tests/golden/make_golden.py
"""
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 40
HERE = os.path.dirname(os.path.abspath(__file__))


def wigner_sum(l, m1, m2, beta):
    c, s = mp.cos(beta / 2), mp.sin(beta / 2)
    tot = mp.mpf(0)
    for k in range(max(0, m2 - m1), min(l - m1, l + m2) + 1):
        num = mp.sqrt(mp.factorial(l + m1) * mp.factorial(l - m1) *
                      mp.factorial(l + m2) * mp.factorial(l - m2))
        den = (mp.factorial(k) * mp.factorial(l - m1 - k) *
               mp.factorial(l + m2 - k) * mp.factorial(m1 - m2 + k))
        tot += (-1) ** (m1 - m2 + k) * num / den * \
            c ** (2 * l - m1 + m2 - 2 * k) * s ** (m1 - m2 + 2 * k)
    return tot


def legendre_rodrigues(l, m, x):
    """(-1)^m (1-x^2)^(m/2) d^m/dx^m P_l(x), P_l from Rodrigues' formula with
    exact integer polynomial coefficients."""
    from fractions import Fraction
    from math import comb, factorial
    # (x^2 - 1)^l coefficients, index = power of x
    c = [0] * (2 * l + 1)
    for k in range(l + 1):
        c[2 * k] = comb(l, k) * (-1) ** (l - k)
    for _ in range(l + m):  # differentiate l + m times
        c = [i * c[i] for i in range(1, len(c))] or [0]
    scale = Fraction(1, 2 ** l * factorial(l))
    poly = mp.mpf(0)
    for i, ci in enumerate(c):
        if ci:
            poly += mp.mpf(ci) * mp.mpf(scale.numerator) / \
                mp.mpf(scale.denominator) * x ** i
    return (-1) ** m * (1 - x * x) ** (mp.mpf(m) / 2) * poly


def main():
    out = {}
    xs = [-0.95, -0.5, -0.1, 0.0, 0.3, 0.77, 0.999]
    L = 12
    leg = np.zeros((len(xs), L + 1, L + 1))
    for i, x in enumerate(xs):
        for l in range(L + 1):
            for m in range(l + 1):
                leg[i, l, m] = float(legendre_rodrigues(l, m, mp.mpf(x)))
    out["legendre_x"] = np.array(xs)
    out["legendre"] = leg
    dirs = [(0.3, 0.2), (1.1, -2.0), (2.9, 4.0), (np.pi / 2, 0.0), (0.0, 0.0)]
    LY = 10
    Y = np.zeros((len(dirs), LY + 1, LY + 1), dtype=np.complex128)
    for i, (t, p) in enumerate(dirs):
        for l in range(LY + 1):
            for m in range(l + 1):
                Y[i, l, m] = complex(mp.spherharm(l, m, mp.mpf(t), mp.mpf(p)))
    out["harmonic_dirs"] = np.array(dirs)
    out["harmonic"] = Y
    betas = [0.0, 0.61547970867038734, np.pi / 2, -0.84106867056793033]
    LW = 12
    d = np.zeros((len(betas), LW + 1, 2 * LW + 1, 2 * LW + 1))
    for i, b in enumerate(betas):
        for l in range(LW + 1):
            for m1 in range(-l, l + 1):
                for m2 in range(-l, l + 1):
                    d[i, l, m1 + LW, m2 + LW] = float(
                        wigner_sum(l, m1, m2, mp.mpf(b)))
    out["wigner_beta"] = np.array(betas)
    out["wigner"] = d
    rng = np.random.default_rng(20240601)
    n = 48
    pos = rng.uniform(-1, 1, (n, 3))
    q = rng.uniform(-1, 1, n)
    pot = np.zeros(n)
    grad = np.zeros((n, 3))
    for i in range(n):
        s = mp.mpf(0)
        g = [mp.mpf(0)] * 3
        for j in range(n):
            if i == j:
                continue
            dv = [mp.mpf(pos[i, k]) - mp.mpf(pos[j, k]) for k in range(3)]
            r = mp.sqrt(sum(v * v for v in dv))
            s += mp.mpf(q[j]) / r
            for k in range(3):
                g[k] -= mp.mpf(q[j]) * dv[k] / r ** 3
        pot[i] = float(s)
        grad[i] = [float(v) for v in g]
    out["direct_pos"] = pos
    out["direct_q"] = q
    out["direct_pot"] = pot
    out["direct_grad"] = grad
    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)


if __name__ == "__main__":
    main()
