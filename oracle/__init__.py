"""ctypes bindings of the FP64 CPU oracle (oracle/aran_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The engine (aran_amd) never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
SRC = os.path.join(HERE, "aran_oracle.c")


def build():
    """Compile the oracle with gcc (no FMA contraction, portable -O2)."""
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-std=gnu99", "-ffp-contract=off",
           "-o", LIB_PATH + ".tmp", SRC, "-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB_PATH + ".tmp", LIB_PATH)


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class OrNode(C.Structure):
    _fields_ = [("depth", C.c_int), ("ix", C.c_int), ("iy", C.c_int),
                ("iz", C.c_int), ("pbegin", C.c_int), ("pend", C.c_int),
                ("parent", C.c_int), ("isleaf", C.c_int),
                ("child_begin", C.c_int), ("nchild", C.c_int),
                ("center", C.c_double * 3)]


NODE_DTYPE = np.dtype([("depth", "i4"), ("ix", "i4"), ("iy", "i4"),
                       ("iz", "i4"), ("pbegin", "i4"), ("pend", "i4"),
                       ("parent", "i4"), ("isleaf", "i4"),
                       ("child_begin", "i4"), ("nchild", "i4"),
                       ("center", "f8", 3)])
assert NODE_DTYPE.itemsize == C.sizeof(OrNode)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        d, i = C.c_double, C.c_int
        L.or_alpha.restype = d
        L.or_alpha.argtypes = [i, i]
        L.or_beta.restype = d
        L.or_beta.argtypes = [i]
        L.or_legendre.argtypes = [i, d, d, _dp]
        L.or_legendre_special.argtypes = [i, d, d, _dp, _dp]
        L.or_harmonics.argtypes = [i, d, d, d, d, _dp]
        L.or_wigner_size.restype = C.c_long
        L.or_wigner_size.argtypes = [i]
        L.or_wigner_d.argtypes = [i, d, _dp]
        L.or_wigner_D.argtypes = [i, d, d, d, _dp]
        L.or_wigner_factorial.restype = d
        L.or_wigner_factorial.argtypes = [i, i, i, d]
        L.or_rotate.argtypes = [i, _dp, d, d, d, _dp]
        L.or_rotate_inverse.argtypes = [i, _dp, d, d, d, _dp]
        L.or_p2m.argtypes = [i, _dp, d, _dp, _dp]
        L.or_m2l_rotate.argtypes = [i, _dp, _dp, i, _dp, _dp]
        L.or_m2m_rotate.argtypes = [i, _dp, _dp, i, _dp, _dp]
        L.or_l2l_rotate.argtypes = [i, _dp, _dp, i, _dp, _dp]
        L.or_m2l_flavour.argtypes = [i, i, _dp, _dp, i, _dp, _dp]
        L.or_m2m_flavour.argtypes = [i, i, _dp, _dp, i, _dp, _dp]
        L.or_l2l_flavour.argtypes = [i, i, _dp, _dp, i, _dp, _dp]
        L.or_local_evaluate.restype = d
        L.or_local_evaluate.argtypes = [i, _dp, _dp, _dp]
        L.or_local_gradient.argtypes = [i, _dp, _dp, _dp, _dp]
        L.or_multipole_evaluate.restype = d
        L.or_multipole_evaluate.argtypes = [i, _dp, _dp, _dp]
        L.or_multipole_gradient.argtypes = [i, _dp, _dp, _dp, _dp]
        L.or_root_box.argtypes = [_dp, _dp, C.c_long, _dp, _dp, _dp, _dp,
                                  C.POINTER(d)]
        L.or_keys.argtypes = [C.c_long, _dp, _dp, _dp, _dp, d, _up]
        L.or_tree_build.restype = C.c_long
        L.or_tree_build.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, i,
                                    _lp, C.c_void_p, C.c_long]
        L.or_traverse.argtypes = [C.c_void_p, C.c_long, _lp, _lp, C.c_long,
                                  C.POINTER(C.c_long), _lp, _lp, C.c_long,
                                  C.POINTER(C.c_long)]
        L.or_p2l.argtypes = [i, _dp, d, _dp, _dp]
        _cl = C.POINTER(C.c_long)
        L.or_traverse_full.argtypes = [C.c_void_p, C.c_long, C.c_long, _lp,
                                       _lp, C.c_long, _cl, _lp, _lp,
                                       C.c_long, _cl, _lp, _lp, C.c_long,
                                       _cl]
        L.or_fmm_solve_full.restype = i
        L.or_fmm_solve_full.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp,
                                        _dp, i, i, C.c_long, _dp, _dp, _lp]
        L.or_fmm_solve.restype = i
        L.or_fmm_solve.argtypes = [C.c_long, _dp, _dp, _dp, _dp, _dp, _dp, i,
                                   i, _dp, _dp, _lp]
        L.or_direct.argtypes = [C.c_long, _dp, _dp, _dp, _dp, C.c_long, _lp,
                                _dp, _dp]
        _lib = L
    return _lib


def tri(n):
    return n * (n + 1) // 2


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def cplx_view(a):
    return a.view(np.complex128)


def legendre(L, x):
    P = np.zeros(tri(L + 1))
    lib().or_legendre(L, float(x), float(np.sqrt(1 - x * x)), P)
    return P


def legendre_special(L, x):
    P = legendre(L, x)
    S = np.zeros(tri(L + 1))
    lib().or_legendre_special(L, float(x), float(np.sqrt(1 - x * x)), P, S)
    return S


def harmonics(L, theta, phi):
    Y = np.zeros(2 * tri(L + 1))
    lib().or_harmonics(L, np.cos(theta), np.sin(theta), np.cos(phi),
                       np.sin(phi), Y)
    return cplx_view(Y)


def wigner_d(L, beta):
    d = np.zeros(lib().or_wigner_size(L))
    lib().or_wigner_d(L, float(beta), d)
    return d


def wigner_index(l, mp, m):
    return l * (l + 1) * (4 * l - 1) // 6 + mp * (2 * l + 1) + l + m


def wigner_factorial(l, m1, m2, beta):
    return lib().or_wigner_factorial(l, m1, m2, float(beta))


def rotate(deg, src, alpha, beta, gamma, inverse=False):
    s = _c(np.asarray(src, dtype=np.complex128).view(np.float64))
    dst = np.zeros(2 * tri(deg + 1))
    f = lib().or_rotate_inverse if inverse else lib().or_rotate
    f(deg, s, alpha, beta, gamma, dst)
    return cplx_view(dst)


def p2m(nm, pos, q, center, M=None):
    if M is None:
        M = np.zeros(tri(nm), dtype=np.complex128)
    Mf = M.view(np.float64)
    lib().or_p2m(nm, _c(pos), float(q), _c(center), Mf)
    return M


# flavour: 0 rotate / z-translate / rotate back (the operator of record),
# 1 the O(p^4) "normal" formulas, 2 K. Kylin's recurrences
def m2l(nm, M, xs, pl, xd, L=None, flavour=0):
    if L is None:
        L = np.zeros(tri(pl + 1), dtype=np.complex128)
    if flavour == 0:
        lib().or_m2l_rotate(nm, _c(M.view(np.float64)), _c(xs), pl,
                            L.view(np.float64), _c(xd))
    else:
        lib().or_m2l_flavour(flavour, nm, _c(M.view(np.float64)), _c(xs), pl,
                             L.view(np.float64), _c(xd))
    return L


def m2m(nm, M, xs, nmd, xd, Md=None, flavour=0):
    if Md is None:
        Md = np.zeros(tri(nmd), dtype=np.complex128)
    if flavour == 0:
        lib().or_m2m_rotate(nm, _c(M.view(np.float64)), _c(xs), nmd,
                            Md.view(np.float64), _c(xd))
    else:
        lib().or_m2m_flavour(flavour, nm, _c(M.view(np.float64)), _c(xs), nmd,
                             Md.view(np.float64), _c(xd))
    return Md


def l2l(pl, Ls, xs, pld, xd, Ld=None, flavour=0):
    if Ld is None:
        Ld = np.zeros(tri(pld + 1), dtype=np.complex128)
    if flavour == 0:
        lib().or_l2l_rotate(pl, _c(Ls.view(np.float64)), _c(xs), pld,
                            Ld.view(np.float64), _c(xd))
    else:
        lib().or_l2l_flavour(flavour, pl, _c(Ls.view(np.float64)), _c(xs),
                             pld, Ld.view(np.float64), _c(xd))
    return Ld


def local_eval(pl, L, center, pos):
    return lib().or_local_evaluate(pl, _c(L.view(np.float64)), _c(center),
                                   _c(pos))


def local_grad(pl, L, center, pos):
    g = np.zeros(3)
    lib().or_local_gradient(pl, _c(L.view(np.float64)), _c(center), _c(pos),
                            g)
    return g


def p2l(pl, pos, q, center, L=None):
    if L is None:
        L = np.zeros(tri(pl + 1), dtype=np.complex128)
    lib().or_p2l(pl, _c(pos), float(q), _c(center), L.view(np.float64))
    return L


def multipole_eval(nm, M, center, pos):
    return lib().or_multipole_evaluate(nm, _c(M.view(np.float64)),
                                       _c(center), _c(pos))


def multipole_grad(nm, M, center, pos):
    g = np.zeros(3)
    lib().or_multipole_gradient(nm, _c(M.view(np.float64)), _c(center),
                                _c(pos), g)
    return g


def root_box(lb, ub, x, y, z):
    lo = np.zeros(3)
    w = C.c_double()
    lib().or_root_box(_c(lb), _c(ub), len(x), _c(x), _c(y), _c(z), lo,
                      C.byref(w))
    return lo, w.value


def keys(x, y, z, lo, width):
    k = np.zeros(len(x), dtype=np.uint64)
    lib().or_keys(len(x), _c(x), _c(y), _c(z), _c(lo), float(width), k)
    return k


def tree(x, y, z, lb, ub, leafcap):
    n = len(x)
    max_nodes = 64 + 8 * (n // (leafcap + 1) + 1) * 22
    nodes = np.zeros(max_nodes, dtype=NODE_DTYPE)
    perm = np.zeros(n, dtype=np.int64)
    nn = lib().or_tree_build(n, _c(x), _c(y), _c(z), _c(lb), _c(ub),
                             int(leafcap), perm, nodes.ctypes.data, max_nodes)
    if nn < 0:
        raise RuntimeError("oracle tree build overflow")
    return perm, nodes[:nn].copy()


def traverse(nodes, max_far=None, max_near=None):
    nn = len(nodes)
    max_far = max_far or 400 * nn + 16
    max_near = max_near or 64 * nn + 16
    fs = np.zeros(max_far, dtype=np.int64)
    fd = np.zeros(max_far, dtype=np.int64)
    na = np.zeros(max_near, dtype=np.int64)
    nb = np.zeros(max_near, dtype=np.int64)
    nf, nn_ = C.c_long(), C.c_long()
    nodes = np.ascontiguousarray(nodes)
    lib().or_traverse(nodes.ctypes.data, nn, fs, fd, max_far, C.byref(nf),
                      na, nb, max_near, C.byref(nn_))
    if nf.value > max_far or nn_.value > max_near:
        return traverse(nodes, 2 * nf.value + 16, 2 * nn_.value + 16)
    return (fs[:nf.value], fd[:nf.value]), (na[:nn_.value], nb[:nn_.value])


def traverse_semifar(nodes, thr, max_far=None, max_near=None, max_semi=None):
    """Traversal with semifar pairs (leaf, deeper node) for a threshold
    thr >= 0: returns (far_src, far_dst), (near_a, near_b),
    (semifar_leaf, semifar_node)."""
    nn = len(nodes)
    max_far = max_far or 400 * nn + 16
    max_near = max_near or 64 * nn + 16
    max_semi = max_semi or 64 * nn + 16
    bufs = [np.zeros(k, dtype=np.int64) for k in
            (max_far, max_far, max_near, max_near, max_semi, max_semi)]
    nf, nn_, ns = C.c_long(), C.c_long(), C.c_long()
    nodes = np.ascontiguousarray(nodes)
    lib().or_traverse_full(nodes.ctypes.data, nn, int(thr), bufs[0], bufs[1],
                           max_far, C.byref(nf), bufs[2], bufs[3], max_near,
                           C.byref(nn_), bufs[4], bufs[5], max_semi,
                           C.byref(ns))
    if nf.value > max_far or nn_.value > max_near or ns.value > max_semi:
        return traverse_semifar(nodes, thr, 2 * nf.value + 16,
                                2 * nn_.value + 16, 2 * ns.value + 16)
    return ((bufs[0][:nf.value], bufs[1][:nf.value]),
            (bufs[2][:nn_.value], bufs[3][:nn_.value]),
            (bufs[4][:ns.value], bufs[5][:ns.value]))


def fmm_solve(x, y, z, q, lb, ub, order, leafcap, semifar=None):
    """Serial FP64 FMM.  semifar: None (no semifar pairs) or a particle-count
    threshold >= 0.  Returns pot, grad (n x 3) and the counters zero, p2p,
    p2m, m2m, m2l, l2l, l2p (and p2l, m2p when semifar is given)."""
    n = len(x)
    pot = np.zeros(n)
    grad = np.zeros(3 * n)
    stats = np.zeros(9, dtype=np.int64)
    thr = -1 if semifar is None else int(semifar)
    rc = lib().or_fmm_solve_full(n, _c(x), _c(y), _c(z), _c(q), _c(lb),
                                 _c(ub), int(order), int(leafcap), thr, pot,
                                 grad, stats)
    if rc != 0:
        raise RuntimeError("oracle solve failed rc=%d" % rc)
    return pot, grad.reshape(n, 3), (stats[:7] if semifar is None else stats)


def direct(x, y, z, q, targets):
    targets = np.ascontiguousarray(targets, dtype=np.int64)
    nt = len(targets)
    pot = np.zeros(nt)
    grad = np.zeros(3 * nt)
    lib().or_direct(len(x), _c(x), _c(y), _c(z), _c(q), nt, targets, pot,
                    grad)
    return pot, grad.reshape(nt, 3)
