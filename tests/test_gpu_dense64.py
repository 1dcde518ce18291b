"""The plain form's far field as dense FP64 products on MFMA (orders <= 15,
rotate flavour) against the per-pair rotate kernel (ARAN_PLAIN_FAR=rotate)
and against the oracle.

Both kernels apply the same operator in FP64; they differ in the order of
the sums only, so potentials agree to 1e-12 relative and gradients to 1e-12
of the largest gradient norm.  Every order takes its own instantiation (row
splits from p = 11, four waves per tile from p = 11, three-step operator
groups at p = 15), so each is run."""
import numpy as np
import pytest
import torch  # noqa: F401

import oracle as O
from aran_amd import workloads as W

pytestmark = pytest.mark.gpu


def _solve(f, dev, lb, ub, p, q):
    pot, grad = f.solve(*dev, lb, ub, p, q)
    st = f.stats()
    st["m2l_tiles"] = f.far_plan()["tiles"]
    return pot.cpu().numpy(), grad.cpu().numpy(), st


@pytest.mark.parametrize("p", list(range(1, 16)))
def test_dense64_equals_rotate_kernel(torch_cuda, monkeypatch, p):
    from aran_amd.fmm import Fmm3d
    torch = torch_cuda
    n, q = 12_000, 24
    x, y, z, m = W.plummer(n, 0xA2A0 + 60 + p)
    dev = [torch.from_numpy(a).cuda() for a in (x, y, z, m)]
    lb, ub = (-1., -1., -1.), (1., 1., 1.)
    f = Fmm3d()
    dp, dg, ds = _solve(f, dev, lb, ub, p, q)
    assert ds["m2l_tiles"] > 0  # the dense kernel ran
    monkeypatch.setenv("ARAN_PLAIN_FAR", "rotate")
    g = Fmm3d()
    rp, rg, rs = _solve(g, dev, lb, ub, p, q)
    assert rs["m2l_tiles"] == 0
    for k in ("zero", "p2p", "p2m", "m2m", "m2l", "l2l", "l2p"):
        assert ds[k] == rs[k], k
    assert np.max(np.abs(dp - rp) / np.abs(rp)) <= 1e-12
    gmax = np.max(np.linalg.norm(rg, axis=1))
    assert np.max(np.linalg.norm(dg - rg, axis=1)) <= 1e-12 * gmax


@pytest.mark.parametrize("p", [4, 10, 15])
def test_dense64_matches_oracle(torch_cuda, p):
    """Uniform cube N = 6,000, leaf 16 (depth >= 3): the dense far field
    against the oracle at 1e-10 with identical counters."""
    from aran_amd.fmm import Fmm3d
    torch = torch_cuda
    n, q = 6000, 16
    x, y, z, m = W.uniform_cube(n, 0xA2A0 + 70 + p)
    dev = [torch.from_numpy(a).cuda() for a in (x, y, z, m)]
    f = Fmm3d()
    pot, grad, st = _solve(f, dev, (0, 0, 0), (1, 1, 1), p, q)
    assert st["m2l_tiles"] > 0 and st["depth"] >= 3
    opot, ograd, ost = O.fmm_solve(x, y, z, m, np.zeros(3), np.ones(3), p, q)
    assert np.max(np.abs(pot - opot) / np.abs(opot)) <= 1e-10
    nb = np.linalg.norm(ograd, axis=1)
    assert np.max(np.linalg.norm(grad - ograd, axis=1) / nb) <= 1e-10
    for i, k in enumerate(("zero", "p2p", "p2m", "m2m", "m2l", "l2l", "l2p")):
        assert st[k] == ost[i], k


def test_dense64_batch_equals_standalone(torch_cuda):
    """Eight systems of mixed sizes in one batched plain solve: each equals
    its standalone solve bit for bit (fixed offset ranges per order)."""
    from aran_amd.fmm import Fmm3d
    torch = torch_cuda
    cat, offs, lbs, ubs = W.batch_systems(8, 0xA2A0 + 80)
    dev = [torch.from_numpy(a).cuda() for a in cat]
    f = Fmm3d()
    pot, grad = f.solve_batch(offs, *dev, lbs, ubs, 8, 32)
    pot, grad = pot.cpu().numpy(), grad.cpu().numpy()
    g = Fmm3d()
    for s in range(8):
        a, b = offs[s], offs[s + 1]
        sp, sg = g.solve(*[t[a:b] for t in dev], lbs[s], ubs[s], 8, 32)
        assert np.array_equal(sp.cpu().numpy(), pot[a:b]), s
        assert np.array_equal(sg.cpu().numpy(), grad[a:b]), s
