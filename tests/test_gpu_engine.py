"""GPU parity of the HIP engine against the FP64 oracle, through the C-ABI.

Bars (stated per test): integer tree / list data bit-exact; FP64 potentials
and gradients within 1e-10 relative of the oracle FMM (same tree, same
operators, different summation order and rotation formulation); FMM error
against the FP64 direct sum at the level the expansion order gives.
"""
import ctypes as C

import numpy as np
import pytest
import torch  # noqa: F401  (imported at collection: first import is slow)

import oracle as O
from aran_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dev(torch, *arrs):
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
            for a in arrs]


def _fmm(torch):
    from aran_amd.fmm import Fmm3d
    return Fmm3d()


def _relerr_pot(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def _relerr_grad(a, b):
    # per particle |g - g_ref| / |g_ref|, the reference's field metric
    # (tests/newtonfield3.c:673-680), with a floor at 1e-8 of the rms field
    nb = np.linalg.norm(b, axis=1)
    floor = 1e-8 * np.sqrt(np.mean(nb ** 2))
    return np.max(np.linalg.norm(a - b, axis=1) / np.maximum(nb, floor))


# ---------------------------------------------------------------- operators

def _new_dev(L, p):
    return L.aran_development3d_new(0, p)


def _arr(ptr, n):
    return np.ctypeslib.as_array(ptr, shape=(n,))


@pytest.mark.parametrize("p", [1, 4, 8, 10, 15])
def test_single_operators_match_oracle(torch_cuda, p):
    from aran_amd import solver as S
    L = S.lib()
    rng = np.random.default_rng(p)
    mc, lc = O.tri(p), O.tri(p + 1)
    src_c = np.array([0.11, -0.07, 0.2])
    dst_c = np.array([1.3, 0.9, -1.7])
    d_src = _new_dev(L, p)
    d_dst = _new_dev(L, p)
    ni_s = S.node_info(src_c)
    ni_d = S.node_info(dst_c)
    # P2M of a few particles
    Mref = np.zeros(mc, dtype=np.complex128)
    for k in range(5):
        pos = src_c + rng.uniform(-0.3, 0.3, 3)
        q = rng.uniform(0.1, 1.0)
        L.aran_development3d_p2m(C.byref(S.VsgVector3d(*pos)), q,
                                 C.byref(ni_s), d_src)
        O.p2m(p, pos, q, src_c, Mref)
    M = S.multipole(d_src)
    assert np.max(np.abs(M - Mref)) <= 1e-13 * np.max(np.abs(Mref))
    # M2L to a far centre
    L.aran_development3d_m2l_rotate(C.byref(ni_s), d_src, C.byref(ni_d), d_dst)
    Lref = O.m2l(p, Mref, src_c, p, dst_c)
    Lg = S.local(d_dst)
    assert np.max(np.abs(Lg - Lref)) <= 1e-12 * np.max(np.abs(Lref))
    # M2M to a parent centre
    par_c = src_c + np.array([0.25, -0.25, 0.25])
    ni_p = S.node_info(par_c)
    d_par = _new_dev(L, p)
    L.aran_development3d_m2m_rotate(C.byref(ni_s), d_src, C.byref(ni_p), d_par)
    Mpref = O.m2m(p, Mref, src_c, p, par_c)
    Mp = S.multipole(d_par)
    assert np.max(np.abs(Mp - Mpref)) <= 1e-12 * np.max(np.abs(Mpref))
    # L2L to a child centre (z below: exercises the z flip)
    ch_c = dst_c + np.array([-0.1, 0.15, -0.2])
    ni_c = S.node_info(ch_c)
    d_ch = _new_dev(L, p)
    L.aran_development3d_l2l_rotate(C.byref(ni_d), d_dst, C.byref(ni_c), d_ch)
    Lcref = O.l2l(p, Lref, dst_c, p, ch_c)
    Lc = S.local(d_ch)
    assert np.max(np.abs(Lc - Lcref)) <= 1e-12 * np.max(np.abs(Lcref))
    # L2P potential and gradient
    for k in range(4):
        x = ch_c + rng.uniform(-0.2, 0.2, 3)
        v = L.aran_development3d_local_evaluate(C.byref(ni_c), d_ch,
                                                C.byref(S.VsgVector3d(*x)))
        g = S.VsgVector3d()
        L.aran_development3d_local_gradient_evaluate(
            C.byref(ni_c), d_ch, C.byref(S.VsgVector3d(*x)), C.byref(g))
        vref = O.local_eval(p, Lcref, ch_c, x)
        gref = O.local_grad(p, Lcref, ch_c, x)
        assert abs(v.re - vref) <= 1e-12 * abs(vref)
        assert np.linalg.norm(np.array([g.x, g.y, g.z]) - gref) <= \
            1e-11 * np.linalg.norm(gref)
    for d in (d_src, d_dst, d_par, d_ch):
        L.aran_development3d_free(d)


def test_m2l_kat_against_newton_potential(torch_cuda):
    """tests/m2l3.c:211-340 on the GPU operator: multipole of a unit charge at
    (0.1,0.2,0.3), M2L to 7 vectors, local checked against 1/|x-p| at
    1e-7, at order 15 (the reference's order 26 runs in
    test_gpu_reference_gates.py, in every flavour)."""
    from aran_amd import solver as S
    L = S.lib()
    p = 15
    src = np.array([0.1, 0.2, 0.3])
    d_src = _new_dev(L, p)
    ni0 = S.node_info([0., 0., 0.])
    L.aran_development3d_p2m(C.byref(S.VsgVector3d(*src)), 1.0, C.byref(ni0),
                             d_src)
    rng = np.random.default_rng(3)
    for t in [(5, 5, 5), (5, 5, -5), (0, 0, -5), (0, 0, 5), (5, 0, 5),
              (0, 5, 5), (-5, -5, 5)]:
        t = np.array(t, dtype=float)
        d_dst = _new_dev(L, p)
        ni_t = S.node_info(t)
        L.aran_development3d_m2l_rotate(C.byref(ni0), d_src, C.byref(ni_t),
                                        d_dst)
        for _ in range(50):
            v = rng.normal(size=3)
            x = t + v / np.linalg.norm(v) * rng.uniform(0, 1)
            res = L.aran_development3d_local_evaluate(
                C.byref(ni_t), d_dst, C.byref(S.VsgVector3d(*x))).re
            ref = 1.0 / np.linalg.norm(x - src)
            assert abs(res - ref) <= 1e-7 * abs(ref), (t, x)
        L.aran_development3d_free(d_dst)
    L.aran_development3d_free(d_src)


# ---------------------------------------------------------------- tree

def _tree_case(kind, n, seed):
    if kind == "uniform":
        return W.uniform_cube(n, seed), (0, 0, 0), (1, 1, 1)
    if kind == "plummer":
        return W.plummer(n, seed), (-1, -1, -1), (1, 1, 1)
    if kind == "outside":
        x, y, z, m = W.uniform_cube(n, seed)
        return (3 * x - 1, y, z, m), (0, 0, 0), (1, 1, 1)
    raise ValueError(kind)


@pytest.mark.parametrize("kind,n,q", [("uniform", 5000, 20),
                                      ("plummer", 20000, 32),
                                      ("outside", 3000, 8),
                                      ("uniform", 300, 1)])
def test_tree_and_lists_bit_exact(torch_cuda, kind, n, q):
    torch = torch_cuda
    (x, y, z, m), lb, ub = _tree_case(kind, n, 11)
    f = _fmm(torch)
    dx, dy, dz, dm = _dev(torch, x, y, z, m)
    f.solve(dx, dy, dz, dm, lb, ub, 4, q)
    perm, nodes, centers, lo, width = f.export_tree()
    operm, onodes = O.tree(x, y, z, np.array(lb, float), np.array(ub, float), q)
    olo, owidth = O.root_box(np.array(lb, float), np.array(ub, float), x, y, z)
    assert width == owidth and np.array_equal(lo, olo)
    assert np.array_equal(perm, operm)
    ref = np.stack([onodes[k] for k in ("depth", "ix", "iy", "iz", "pbegin",
                                        "pend", "parent", "isleaf",
                                        "child_begin", "nchild")], axis=1)
    assert nodes.shape == ref.shape
    assert np.array_equal(nodes, ref)
    assert np.array_equal(centers, onodes["center"])
    # far list == ordered far pairs of the oracle traversal
    lists = f.export_lists()
    (fs, fd), (na, nb) = O.traverse(onodes)
    gpu_far = set()
    fo = lists["far_ofs"]
    for t in range(len(fo) - 1):
        for e in range(fo[t], fo[t + 1]):
            gpu_far.add((int(lists["far_src"][e]), t))
    assert gpu_far == set(zip(fs.tolist(), fd.tolist()))
    assert len(gpu_far) == len(fs)
    # near list expanded to ordered leaf pairs == oracle near pairs
    leaf_of = np.full(n, -1)
    for i in np.nonzero(onodes["isleaf"])[0]:
        leaf_of[onodes["pbegin"][i]:onodes["pend"][i]] = i
    gpu_near = set()
    no = lists["near_ofs"]
    for t, A in enumerate(lists["leaves"]):
        for e in range(no[t], no[t + 1]):
            b, en = lists["near_begin"][e], lists["near_end"][e]
            for B in np.unique(leaf_of[b:en]):
                key = (int(A), int(B))
                assert key not in gpu_near, "leaf pair listed twice"
                gpu_near.add(key)
    ref_near = set()
    for a, b in zip(na.tolist(), nb.tolist()):
        ref_near.add((a, b))
        ref_near.add((b, a))
    assert gpu_near == ref_near


@pytest.mark.parametrize("kind,n,q", [("uniform", 200_000, 8),
                                      ("plummer", 300_000, 16)])
def test_large_tree_bit_exact(torch_cuda, kind, n, q):
    """Trees of 10^4 - 10^5 nodes: level launches of many workgroups and
    node / octant scans across workgroups, against the oracle's tree."""
    torch = torch_cuda
    (x, y, z, m), lb, ub = _tree_case(kind, n, 23)
    f = _fmm(torch)
    dx, dy, dz, dm = _dev(torch, x, y, z, m)
    f.solve(dx, dy, dz, dm, lb, ub, 4, q)
    perm, nodes, centers, lo, width = f.export_tree()
    assert nodes.shape[0] > 20_000
    operm, onodes = O.tree(x, y, z, np.array(lb, float), np.array(ub, float), q)
    assert np.array_equal(perm, operm)
    ref = np.stack([onodes[k] for k in ("depth", "ix", "iy", "iz", "pbegin",
                                        "pend", "parent", "isleaf",
                                        "child_begin", "nchild")], axis=1)
    assert nodes.shape == ref.shape
    assert np.array_equal(nodes, ref)
    assert np.array_equal(centers, onodes["center"])
    # the per-leaf near ranges cover, for every leaf, each particle of its
    # near leaves once: counts against the oracle's near pairs
    lists = f.export_lists()
    (fs, fd), (na, nb) = O.traverse(onodes)
    size = onodes["pend"] - onodes["pbegin"]
    want = np.zeros(len(size), np.int64)
    two = na != nb  # (a, a): a leaf with itself, listed once
    np.add.at(want, na, size[nb])
    np.add.at(want, nb[two], size[na[two]])
    leaves = np.asarray(lists["leaves"])
    no = np.asarray(lists["near_ofs"])
    span = np.asarray(lists["near_end"]) - np.asarray(lists["near_begin"])
    got = np.add.reduceat(span, no[:-1]) if len(span) else np.zeros(0)
    got = np.where(no[1:] > no[:-1], got, 0)
    assert np.array_equal(got, want[leaves])
    fo = np.asarray(lists["far_ofs"])
    assert fo[-1] == len(fs)


def test_near_lists_one_launch_and_rerun(torch_cuda):
    """A context that has solved before builds its near lists in one launch
    sized from the last solve's ranges per leaf.  A small, shallow problem
    first sets that size far too low for the clustered one after it, whose
    lists then outgrow their arrays and are rebuilt by a second run; a third
    solve takes the one-launch path within its size.  Every result and list
    equals a fresh context's (two-pass lists) bitwise."""
    torch = torch_cuda
    (xs, ys, zs, ms), lbs, ubs = _tree_case("uniform", 2000, 5)
    (x, y, z, m), lb, ub = _tree_case("plummer", 200_000, 6)
    dev = _dev(torch, x, y, z, m)
    fresh = _fmm(torch)
    want = [a.cpu().numpy() for a in fresh.solve(*dev, lb, ub, 6, 8)]
    want_lists = fresh.export_lists()
    want_st = fresh.stats()
    f = _fmm(torch)
    f.solve(*_dev(torch, xs, ys, zs, ms), lbs, ubs, 6, 64)
    for _ in range(2):  # rerun after overflow, then within the size
        got = [a.cpu().numpy() for a in f.solve(*dev, lb, ub, 6, 8)]
        assert np.array_equal(got[0], want[0])
        assert np.array_equal(got[1], want[1])
        assert f.stats() == want_st
        lists = f.export_lists()
        for k in ("leaves", "near_ofs", "near_begin", "near_end"):
            assert np.array_equal(lists[k], want_lists[k]), k


# ---------------------------------------------------------------- solve

@pytest.mark.parametrize("kind,n,p,q", [("uniform", 10_000, 8, 20),
                                        ("plummer", 8000, 10, 64),
                                        ("outside", 2000, 6, 16),
                                        ("uniform", 500, 15, 10),
                                        ("uniform", 400, 1, 5)])
def test_solve_matches_oracle(torch_cuda, kind, n, p, q):
    """Whole FP64 solve vs the oracle FMM: max relative 1e-10 (potential,
    pointwise) and 1e-10 (gradient, pointwise with a 1e-8 rms floor);
    counters identical."""
    torch = torch_cuda
    (x, y, z, m), lb, ub = _tree_case(kind, n, 5)
    f = _fmm(torch)
    dx, dy, dz, dm = _dev(torch, x, y, z, m)
    pot, grad = f.solve(dx, dy, dz, dm, lb, ub, p, q)
    pot = pot.cpu().numpy()
    grad = grad.cpu().numpy()
    opot, ograd, ostats = O.fmm_solve(x, y, z, m, np.array(lb, float),
                                      np.array(ub, float), p, q)
    assert _relerr_pot(pot, opot) <= 1e-10
    assert _relerr_grad(grad, ograd) <= 1e-10
    st = f.stats()
    for i, k in enumerate(("zero", "p2p", "p2m", "m2m", "m2l", "l2l", "l2p")):
        assert st[k] == ostats[i], k


def test_solve_error_against_direct_sum(torch_cuda):
    """N=10^4 uniform, p=8, q=20 (the reference-sized case): FMM error vs
    the GPU FP64 direct sum on 4096 targets; the GPU direct sum itself vs
    the oracle's compensated direct sum at 1e-13."""
    torch = torch_cuda
    x, y, z, m = W.uniform_cube(10_000, 0xA2A0 + 1)
    f = _fmm(torch)
    dx, dy, dz, dm = _dev(torch, x, y, z, m)
    pot, grad = f.solve(dx, dy, dz, dm, (0, 0, 0), (1, 1, 1), 8, 20)
    tg = W.sample_targets(len(x), 4096, 0xA2A0 + 1)
    dpot, dgrad = f.direct(dx, dy, dz, dm, torch.from_numpy(tg))
    dpot, dgrad = dpot.cpu().numpy(), dgrad.cpu().numpy()
    opot, ograd = O.direct(x, y, z, m, tg[:64])
    assert _relerr_pot(dpot[:64], opot) <= 1e-13
    assert np.max(np.abs(dgrad[:64] - ograd)) <= 1e-12 * np.max(np.abs(ograd))
    p_ = pot.cpu().numpy()[tg]
    g_ = grad.cpu().numpy()[tg]
    l2p = np.linalg.norm(p_ - dpot) / np.linalg.norm(dpot)
    l2g = np.linalg.norm(g_ - dgrad) / np.linalg.norm(dgrad)
    assert l2p <= 5e-6 and _relerr_pot(p_, dpot) <= 1e-3
    assert l2g <= 1e-3


def test_edge_cases(torch_cuda):
    torch = torch_cuda
    f = _fmm(torch)
    e = torch.empty(0, dtype=torch.float64, device="cuda")
    f.solve(e, e, e, e, (0, 0, 0), (1, 1, 1), 8, 20)
    assert f.stats()["nodes"] == 0
    # one particle: zero field
    one = torch.tensor([0.5], dtype=torch.float64, device="cuda")
    pot, grad = f.solve(one, one, one, one, (0, 0, 0), (1, 1, 1), 8, 20)
    assert pot.item() == 0.0 and torch.all(grad == 0)
    # two particles
    xs = torch.tensor([0.1, 0.9], dtype=torch.float64, device="cuda")
    ys = torch.tensor([0.2, 0.3], dtype=torch.float64, device="cuda")
    zs = torch.tensor([0.5, 0.5], dtype=torch.float64, device="cuda")
    qs = torch.tensor([1.0, 2.0], dtype=torch.float64, device="cuda")
    pot, grad = f.solve(xs, ys, zs, qs, (0, 0, 0), (1, 1, 1), 8, 1)
    r = np.hypot(0.8, 0.1)
    assert np.allclose(pot.cpu().numpy(), [2.0 / r, 1.0 / r], rtol=1e-14)
    # invalid order (above 30; above 15 outside the plain form) and
    # non-finite input are rejected
    with pytest.raises(RuntimeError):
        f.solve(xs, ys, zs, qs, (0, 0, 0), (1, 1, 1), 31, 1)
    with pytest.raises(RuntimeError):
        f.solve(xs, ys, zs, qs, (0, 0, 0), (1, 1, 1), 0, 1)
    f.set_precision("mixed")
    with pytest.raises(RuntimeError):
        f.solve(xs, ys, zs, qs, (0, 0, 0), (1, 1, 1), 16, 1)
    f.set_precision("plain")
    bad = torch.tensor([0.1, float("nan")], dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):
        f.solve(bad, ys, zs, qs, (0, 0, 0), (1, 1, 1), 8, 1)


@pytest.mark.parametrize("form", ["mixed", "bf16x3"])
def test_fp32_near_field_coincident_coordinates(torch_cuda, form):
    """Two distinct particles 1e-12 apart share their FP32 leaf-relative
    coordinates, so the FP32 near field sees r2 = 0 for the pair; the
    targets whose FP32 sums are not finite are summed again in FP64, so the
    result stays finite and equal to the FP64 direct sum (the pair's term,
    1e12, dominates their potentials)."""
    torch = torch_cuda
    f = _fmm(torch)
    f.set_precision(form)
    rng = np.random.default_rng(11)
    n = 300
    x, y, z = rng.random((3, n)) * 0.9 + 0.05
    q = 1.0 - rng.random(n)
    x[1], y[1], z[1] = x[0] + 1e-12, y[0], z[0]
    dev = [torch.from_numpy(a).cuda() for a in (x, y, z, q)]
    pot, grad = f.solve(*dev, (0, 0, 0), (1, 1, 1), 6, 16)
    pot, grad = pot.cpu().numpy(), grad.cpu().numpy()
    assert np.all(np.isfinite(pot)) and np.all(np.isfinite(grad))
    dp, dg = f.direct(*dev, torch.arange(n))
    dp = dp.cpu().numpy()
    assert np.max(np.abs(pot[:2] - dp[:2]) / np.abs(dp[:2])) <= 1e-9
    assert np.linalg.norm(pot - dp) / np.linalg.norm(dp) <= 1e-6


def test_solver_api_newtonpot3_style(torch_cuda):
    """AranSolver3d used as tests/newtonpot3.c / newtonfield3.c use it, with
    the library callbacks and the declared particle layout; results match
    the oracle at 1e-10 and counters match."""
    from aran_amd import solver as S
    L = S.lib()
    n, p, q = 2400, 10, 100
    x, y, z, m = W.uniform_cube(n, 77)
    x, y, z = 2 * 0.999 * x - 0.999, 2 * 0.999 * y - 0.999, 2 * 0.999 * z - 0.999
    pts = (S.PointAccumBoth * n)()
    for i in range(n):
        pts[i].vector = S.VsgVector3d(x[i], y[i], z[i])
        pts[i].density = m[i]
    lb, ub = S.VsgVector3d(-1, -1, -1), S.VsgVector3d(1, 1, 1)
    tree = L.vsg_prtree3d_new_full(C.byref(lb), C.byref(ub),
                                   S.fn("vsg_vector3d_vector3d_locfunc"),
                                   S.fn("vsg_vector3d_dist"), None, q)
    assert tree
    dev = L.aran_development3d_new(0, p)
    solver = L.aran_solver3d_new(tree, L.aran_development3d_get_type(),
                                 C.cast(dev, C.c_void_p),
                                 S.fn("aran_development3d_set_zero"))
    L.aran_solver3d_set_functions(
        solver, S.fn("aran_newton3d_p2p_potfield"), S.fn("aran_newton3d_p2m"),
        S.fn("aran_development3d_m2m_rotate"),
        S.fn("aran_development3d_m2l_rotate"),
        S.fn("aran_development3d_l2l_rotate"),
        S.fn("aran_newton3d_l2p_potfield"))
    L.aran_solver3d_set_particle_layout(solver, S.PointAccumBoth.density.offset,
                                        S.PointAccumBoth.pot.offset, 0,
                                        S.PointAccumBoth.field.offset)
    for i in range(n):
        L.aran_solver3d_insert_point(solver, C.addressof(pts[i]))
    assert L.aran_solver3d_point_count(solver) == n
    L.aran_solver3d_solve(solver)
    assert L.aran_solver3d_status(solver) == 0
    st = S.get_stats(solver)
    pot = np.array([pts[i].pot for i in range(n)])
    grad = np.array([[pts[i].field.x, pts[i].field.y, pts[i].field.z]
                     for i in range(n)])
    opot, ograd, ostats = O.fmm_solve(x, y, z, m, -np.ones(3), np.ones(3), p, q)
    assert _relerr_pot(pot, opot) <= 1e-10
    assert _relerr_grad(grad, ograd) <= 1e-10
    for i, k in enumerate(("zero", "p2p", "p2m", "m2m", "m2l", "l2l", "l2p")):
        assert st[k] == ostats[i], k
    # the reference's own end-to-end gate: 1e-3 pointwise vs direct sum
    tg = np.arange(0, n, 3)
    dpot, dgrad = O.direct(x, y, z, m, tg)
    assert _relerr_pot(pot[tg], dpot) <= 1e-3
    # an unrecognised callback set is refused, results untouched
    L.aran_solver3d_set_functions(solver, S.fn("aran_newton3d_p2p_pot"),
                                  S.fn("aran_newton3d_p2m"),
                                  S.fn("aran_development3d_m2m_rotate"),
                                  S.fn("aran_development3d_m2l_rotate"),
                                  S.fn("aran_development3d_l2l_rotate"),
                                  S.fn("aran_newton3d_l2p_field"))
    before = pot.copy()
    L.aran_solver3d_solve(solver)
    assert L.aran_solver3d_status(solver) != 0
    assert np.array_equal(before, np.array([pts[i].pot for i in range(n)]))
    L.aran_solver3d_free(solver)


@pytest.mark.parametrize("kind,n,p,q", [("uniform", 10_000, 8, 20),
                                        ("plummer", 20_000, 10, 64),
                                        ("uniform", 3000, 14, 32),
                                        ("plummer", 5000, 6, 16),
                                        ("outside", 2000, 3, 8)])
def test_mixed_form_close_to_plain(torch_cuda, kind, n, p, q):
    """Mixed form (FP32 M2L on scaled expansions, FP32 P2P) against the plain
    FP64 form on the same tree: potential L2 relative difference <= 5e-7 and
    pointwise <= 1e-5, gradient L2 <= 5e-6; counters identical."""
    torch = torch_cuda
    (x, y, z, m), lb, ub = _tree_case(kind, n, 9)
    f = _fmm(torch)
    dev = _dev(torch, x, y, z, m)
    f.set_precision("plain")
    p0, g0 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    s0 = f.stats()
    f.set_precision("mixed")
    p1, g1 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    assert f.stats() == s0
    assert np.linalg.norm(p1 - p0) / np.linalg.norm(p0) <= 5e-7
    assert _relerr_pot(p1, p0) <= 1e-5
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) <= 5e-6
    # deterministic: a second mixed solve is bitwise identical
    p2, g2 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    assert np.array_equal(p1, p2) and np.array_equal(g1, g2)


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_batch_equals_individual_solves(torch_cuda, precision):
    """Independent systems solved in one launch sequence give, for each
    system, exactly (bitwise) the result of solving it alone; counters add
    up.  Includes an empty system and systems outside their given bounds."""
    torch = torch_cuda
    rng = np.random.default_rng(3)
    systems = []
    for k in range(7):
        n = [900, 0, 3000, 1, 2500, 40, 1800][k]
        if k % 2 == 0:
            x, y, z, m = W.uniform_cube(n, 100 + k)
            lb, ub = (0, 0, 0), (1, 1, 1)
        else:
            x, y, z, m = W.plummer(n, 100 + k) if n else (np.zeros(0),) * 4
            lb, ub = (-1, -1, -1), (1, 1, 1)
        if k == 4:
            x = x * 3 - 1  # grows the root cube
        systems.append((x, y, z, m, lb, ub))
    p, q = 8, 24
    f = _fmm(torch)
    f.set_precision(precision)
    offs = np.cumsum([0] + [len(s[0]) for s in systems])
    cat = [np.concatenate([s[i] for s in systems]) for i in range(4)]
    dev = _dev(torch, *cat)
    pb, gb = [t.cpu().numpy() for t in f.solve_batch(
        offs, *dev, [s[4] for s in systems], [s[5] for s in systems], p, q)]
    sb = f.stats()
    tot = {k: 0 for k in ("zero", "p2p", "p2m", "m2m", "m2l", "l2l", "l2p")}
    for k, (x, y, z, m, lb, ub) in enumerate(systems):
        if len(x) == 0:
            continue
        d = _dev(torch, x, y, z, m)
        p1, g1 = [t.cpu().numpy() for t in f.solve(*d, lb, ub, p, q)]
        sl = slice(offs[k], offs[k + 1])
        assert np.array_equal(pb[sl], p1), k
        assert np.array_equal(gb[sl], g1), k
        s1 = f.stats()
        for key in tot:
            tot[key] += s1[key]
    tot["zero"] += 1  # the empty system's root
    for key in tot:
        assert sb[key] == tot[key], key


def test_served_results_equal_standalone(torch_cuda):
    """Requests of mixed sizes, orders and outputs submitted together are
    served in batches; every served result is bitwise the standalone solve
    of that request (mixed form: these batches are small enough that the
    batch and the request alone choose the same far-field split); a request
    over the budget is refused."""
    torch = torch_cuda
    from aran_amd.server import EBUDGET, WORKER_BYTES, Server, footprint
    reqs = []
    for k in range(12):
        n = [1200, 5000, 800, 20000, 3000, 1, 9000, 2500, 700, 15000, 4000,
             6000][k]
        order = [6, 8][k % 2]
        if k % 3 == 0:
            x, y, z, m = W.plummer(n, 500 + k)
            lb, ub = (-1, -1, -1), (1, 1, 1)
        else:
            x, y, z, m = W.uniform_cube(n, 500 + k)
            lb, ub = (0, 0, 0), (1, 1, 1)
        reqs.append((x, y, z, m, lb, ub, order, 32, k % 2 == 0))
    srv = Server(budget_bytes=footprint(40000, 8, 32) * 2 + 2 * WORKER_BYTES, nworkers=2,
                 precision="mixed", max_batch=8)
    srv.pause()
    ids = [srv.submit(*r[:4], r[4], r[5], r[6], r[7], want_grad=r[8])
           for r in reqs]
    srv.pause(False)
    assert all(i >= 0 for i in ids)
    assert srv.submit(*W.uniform_cube(10, 1), (0, 0, 0), (1, 1, 1), 8, 32) >= 0
    big = srv.submit(*W.uniform_cube(10, 1), (0, 0, 0), (1, 1, 1), 8, 32)
    srv.wait(big - 1)
    srv.wait(big)
    out = [srv.wait(i) for i in ids]
    st = srv.stats()
    assert st["served"] == len(reqs) + 2 and st["batches"] < st["served"]
    srv.close()
    f = _fmm(torch)
    f.set_precision("mixed")
    for r, (pot, grad, times) in zip(reqs, out):
        d = _dev(torch, *r[:4])
        p1, g1 = [t.cpu().numpy() for t in f.solve(*d, r[4], r[5], r[6], r[7])]
        assert np.array_equal(pot, p1)
        if r[8]:
            assert np.array_equal(grad, g1)
        else:
            assert grad is None
        assert times["latency"] >= times["service"] > 0
    small = Server(budget_bytes=footprint(1000, 8, 32) * 2 + 2 * WORKER_BYTES, nworkers=2)
    rid = small.submit(*W.uniform_cube(5000, 3), (0, 0, 0), (1, 1, 1), 8, 32)
    assert rid == EBUDGET
    small.close()


def test_prepared_server(torch_cuda):
    """prepare() sizes every worker for 30000 particles and builds the tables
    of orders 6 and 8; requests served afterwards are still bitwise their
    standalone solves, and prepare() is refused while a request is queued.
    Plain form: its served results are bitwise whatever batches the requests
    land in (the other forms' far-field split follows the batch size)."""
    torch = torch_cuda
    from aran_amd.server import Server
    srv = Server(budget_bytes=4e9, nworkers=2, precision="plain", max_batch=4)
    srv.prepare(30000, (6, 8), 32)
    reqs = []
    for k, n in enumerate([25000, 300, 7000, 1]):
        x, y, z, m = W.uniform_cube(n, 700 + k)
        reqs.append((x, y, z, m, (0, 0, 0), (1, 1, 1), (6, 8)[k % 2], 32))
    ids = [srv.submit(*r[:4], r[4], r[5], r[6], r[7]) for r in reqs]
    out = [srv.wait(i) for i in ids]
    srv.pause()
    rid = srv.submit(*W.uniform_cube(50, 9), (0, 0, 0), (1, 1, 1), 8, 32)
    with pytest.raises(RuntimeError):
        srv.prepare(1000, (8,), 32)
    srv.pause(False)
    srv.wait(rid)
    srv.close()
    f = _fmm(torch)
    f.set_precision("plain")
    for r, (pot, grad, _) in zip(reqs, out):
        d = _dev(torch, *r[:4])
        p1, g1 = [t.cpu().numpy() for t in f.solve(*d, r[4], r[5], r[6], r[7])]
        assert np.array_equal(pot, p1) and np.array_equal(grad, g1)


@pytest.mark.parametrize("form", ["plain", "bf16x3"])
def test_order_switches_reuse_tables(torch_cuda, form):
    """A context keeps the tables of recent orders: solves that cycle over
    five orders (more than it caches, so some are rebuilt after eviction)
    and then alternate over two (swapped in from the cache) give, for every
    repeat of an order, bitwise the result of that order's first solve."""
    torch = torch_cuda
    d = _dev(torch, *W.uniform_cube(3000, 808))
    f = _fmm(torch)
    f.set_precision(form)
    first = {}
    for p in [6, 8, 10, 12, 14, 6, 8, 10, 12, 14, 6, 8, 6, 8, 14]:
        pot, grad = [t.cpu().numpy() for t in f.solve(*d, (0, 0, 0),
                                                       (1, 1, 1), p, 32)]
        if p not in first:
            first[p] = (pot, grad)
        else:
            assert np.array_equal(pot, first[p][0]), p
            assert np.array_equal(grad, first[p][1]), p


@pytest.mark.parametrize("form,code", [("plain", 0), ("bf16x3", 2)])
def test_solver_api_structs_match_device_solve(torch_cuda, form, code):
    """150,000 particle structs (so that the gather from and the scatter to
    them run on several host threads), solved through the AranSolver3d
    pointer API in the given form: every potential and field is bitwise the
    device-array solve of the same system; an unknown form is refused."""
    torch = torch_cuda
    from aran_amd import solver as S
    L = S.lib()
    n, p, q = 150_000, 8, 64
    x, y, z, m = W.plummer(n, 91)
    dt = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("density", "f8"),
                   ("pot", "f8"), ("fx", "f8"), ("fy", "f8"), ("fz", "f8")])
    assert dt.itemsize == C.sizeof(S.PointAccumBoth)
    pts = np.zeros(n, dtype=dt)
    pts["x"], pts["y"], pts["z"], pts["density"] = x, y, z, m
    lb, ub = S.VsgVector3d(-1, -1, -1), S.VsgVector3d(1, 1, 1)
    tree = L.vsg_prtree3d_new_full(C.byref(lb), C.byref(ub),
                                   S.fn("vsg_vector3d_vector3d_locfunc"),
                                   S.fn("vsg_vector3d_dist"), None, q)
    dev = L.aran_development3d_new(0, p)
    solver = L.aran_solver3d_new(tree, L.aran_development3d_get_type(),
                                 C.cast(dev, C.c_void_p),
                                 S.fn("aran_development3d_set_zero"))
    L.aran_solver3d_set_functions(
        solver, S.fn("aran_newton3d_p2p_potfield"), S.fn("aran_newton3d_p2m"),
        S.fn("aran_development3d_m2m_rotate"),
        S.fn("aran_development3d_m2l_rotate"),
        S.fn("aran_development3d_l2l_rotate"),
        S.fn("aran_newton3d_l2p_potfield"))
    L.aran_solver3d_set_particle_layout(solver, S.PointAccumBoth.density.offset,
                                        S.PointAccumBoth.pot.offset, 0,
                                        S.PointAccumBoth.field.offset)
    assert L.aran_solver3d_set_precision(solver, 7) == -1
    assert L.aran_solver3d_set_precision(solver, code) == 0
    base = pts.ctypes.data
    for i in range(n):
        L.aran_solver3d_insert_point(solver, base + i * dt.itemsize)
    L.aran_solver3d_solve(solver)
    assert L.aran_solver3d_status(solver) == 0
    L.aran_solver3d_free(solver)
    f = _fmm(torch)
    f.set_precision(form)
    d = _dev(torch, x, y, z, m)
    p1, g1 = [t.cpu().numpy() for t in f.solve(*d, (-1, -1, -1), (1, 1, 1),
                                               p, q)]
    assert np.array_equal(pts["pot"], p1)
    assert np.array_equal(np.stack([pts["fx"], pts["fy"], pts["fz"]], 1), g1)
