"""Seeded synthetic particle systems of the benchmark configurations.

Counter-based SplitMix64 streams (seed, stream id, counter) give the same
FP64 inputs on every machine.  Distributions:
  uniform  : x ~ U[0,1)^3, masses 1 - U in (0,1]            (1e4, p=8, q=20)
  plummer  : Plummer sphere, scale radius a, truncated at rmax, equal masses
             1/N, by the inverse-CDF form of tests/newtonpot3.c:548-559
  signed   : x ~ U[0,1)^3, charges 2U - 1 in [-1, 1)
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed, stream, start, n):
    """n outputs of SplitMix64 at counters start .. start+n-1."""
    with np.errstate(over="ignore"):
        base = np.uint64((seed * 0x100000001B3 + stream * 0x9E3779B1) &
                         0xFFFFFFFFFFFFFFFF)
        ctr = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = base + ctr * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
        return z ^ (z >> np.uint64(31))


def uniform(seed, stream, n, start=0):
    """U[0,1) doubles with 53 random bits."""
    return (splitmix64(seed, stream, start, n) >> np.uint64(11)).astype(
        np.float64) * (1.0 / 9007199254740992.0)


def uniform_cube(n, seed):
    x, y, z = (uniform(seed, s, n) for s in (0, 1, 2))
    m = 1.0 - uniform(seed, 3, n)
    return x, y, z, m


def signed_cube(n, seed):
    x, y, z = (uniform(seed, s, n) for s in (0, 1, 2))
    q = 2.0 * uniform(seed, 3, n) - 1.0
    return x, y, z, q


def plummer(n, seed, a=0.1, rmax=1.0):
    """Plummer sphere by rejection of r > rmax, equal masses 1/n."""
    out = []
    got, start = 0, 0
    while got < n:
        k = max(1024, int((n - got) * 1.2))
        u1 = uniform(seed, 0, k, start)
        u2 = uniform(seed, 1, k, start)
        u3 = uniform(seed, 2, k, start)
        start += k
        u1 = np.where(u1 > 0, u1, 1e-300)
        with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
            r = a / np.sqrt(u1 ** (-2.0 / 3.0) - 1.0)
        ok = np.isfinite(r) & (r <= rmax)
        r, u2, u3 = r[ok], u2[ok], u3[ok]
        z = r * (1.0 - 2.0 * u2)
        t = np.sqrt(np.maximum(r * r - z * z, 0.0))
        ph = 2.0 * np.pi * u3
        out.append(np.stack([t * np.cos(ph), t * np.sin(ph), z]))
        got += r.size
    p = np.concatenate(out, axis=1)[:, :n]
    m = np.full(n, 1.0 / n)
    return p[0].copy(), p[1].copy(), p[2].copy(), m


def sample_targets(n, k, seed):
    """k distinct indices in [0, n) (Floyd's algorithm on a SplitMix64
    stream), sorted."""
    k = min(k, n)
    u = uniform(seed, 7, k)
    chosen = set()
    for j, t in enumerate(range(n - k, n)):
        r = int(u[j] * (t + 1))
        chosen.add(t if r in chosen else r)
    return np.array(sorted(chosen), dtype=np.int64)


CONFIGS = {
    # name: (generator, n, order, leafcap, lb, ub, seed)
    "uniform_1e4_p8_q20": ("uniform", 10_000, 8, 20, (0, 0, 0), (1, 1, 1),
                           0xA2A0 + 1),
    "plummer_2p20_p10_q64": ("plummer", 1 << 20, 10, 64, (-1, -1, -1),
                             (1, 1, 1), 0xA2A0 + 2),
    "signed_4m_q64": ("signed", 4_194_304, 10, 64, (0, 0, 0), (1, 1, 1),
                      0xA2A0 + 3),
}


def make(kind, n, seed):
    if kind == "uniform":
        return uniform_cube(n, seed)
    if kind == "plummer":
        return plummer(n, seed)
    if kind == "signed":
        return signed_cube(n, seed)
    raise ValueError(kind)


def batch_systems(nsys, seed):
    """Independent systems of the batched configuration: sizes
    round(10^(3 + 2u)) (log-uniform in [1e3, 1e5]), each a uniform cube or
    a Plummer sphere with probability 1/2, masses in (0, 1].  Returns the
    concatenated (x, y, z, m), the offsets and the per-system root bounds."""
    u = uniform(seed, 20, 2 * nsys).reshape(nsys, 2)
    xs, ys, zs, ms, lbs, ubs, offs = [], [], [], [], [], [], [0]
    for k in range(nsys):
        n = int(round(10 ** (3 + 2 * u[k, 0])))
        if u[k, 1] < 0.5:
            x, y, z, _ = uniform_cube(n, seed * 1000 + k)
            lbs.append((0, 0, 0))
            ubs.append((1, 1, 1))
        else:
            x, y, z, _ = plummer(n, seed * 1000 + k)
            lbs.append((-1, -1, -1))
            ubs.append((1, 1, 1))
        m = 1.0 - uniform(seed * 1000 + k, 9, n)
        for a, b in ((xs, x), (ys, y), (zs, z), (ms, m)):
            a.append(b)
        offs.append(offs[-1] + n)
    cat = [np.concatenate(a) for a in (xs, ys, zs, ms)]
    return cat, offs, np.array(lbs, float), np.array(ubs, float)


def serving_requests(n_requests, seed, nmax=2_000_000, orders=(6, 8, 12)):
    """Requests of the serving configuration: N = round(10^(3 + u log10
    2000)) (log-uniform in [1e3, 2e6]), p drawn from `orders`, half
    potential-only, uniform cube or Plummer (the first N points of seeded
    2e6-point pools, so each request is an i.i.d. sample), masses in (0, 1],
    leaf 64.  Returns a list of (x, y, z, m, lb, ub, order, leafcap,
    want_grad) and the Poisson inter-arrival gaps for rate 1 (divide by the
    rate)."""
    pools = {"cube": uniform_cube(nmax, seed), "plummer": plummer(nmax, seed)}
    mass = 1.0 - uniform(seed, 9, nmax)
    u = uniform(seed, 10, 5 * n_requests).reshape(n_requests, 5)
    reqs = []
    for k in range(n_requests):
        n = min(nmax, int(round(10 ** (3 + u[k, 0] * np.log10(2000)))))
        order = orders[min(len(orders) - 1, int(u[k, 1] * len(orders)))]
        kind = "cube" if u[k, 2] < 0.5 else "plummer"
        want_grad = bool(u[k, 3] >= 0.5)
        x, y, z, _ = pools[kind]
        lb, ub = ((0, 0, 0), (1, 1, 1)) if kind == "cube" else \
            ((-1, -1, -1), (1, 1, 1))
        reqs.append((x[:n], y[:n], z[:n], mass[:n], lb, ub, order, 64,
                     want_grad))
    gaps = -np.log(1.0 - u[:, 4])
    return reqs, gaps
