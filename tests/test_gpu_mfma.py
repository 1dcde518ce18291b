"""GPU tests of the MFMA far field (dense BF16 M2L operators, 3 or 6
products) against the plain FP64 form on the same tree, through the C-ABI.

Bars: BF16x6 (6 products, FP32-level) potential L2 relative difference to
the plain form <= 5e-7, pointwise <= 1e-5, gradient L2 <= 5e-6 (the bar of
the FP32 VALU form); BF16x3 (3 products) potential L2 <= 3e-6, gradient L2
<= 3e-5; counters identical to the plain form; repeated solves bitwise
identical.
"""
import numpy as np
import pytest
import torch  # noqa: F401

from aran_amd import workloads as W

pytestmark = pytest.mark.gpu


def _dev(torch, *arrs):
    return [torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()
            for a in arrs]


def _case(kind, n, seed):
    if kind == "plummer":
        return W.plummer(n, seed), (-1, -1, -1), (1, 1, 1)
    if kind == "signed":
        return W.signed_cube(n, seed), (0, 0, 0), (1, 1, 1)
    return W.uniform_cube(n, seed), (0, 0, 0), (1, 1, 1)


BARS = {"bf16x6": (5e-7, 1e-5, 5e-6), "bf16x3": (3e-6, 1e-4, 3e-5)}


@pytest.mark.parametrize("form", ["bf16x6", "bf16x3"])
@pytest.mark.parametrize("kind,n,p,q", [("uniform", 10_000, 8, 20),
                                        ("plummer", 20_000, 10, 64),
                                        ("uniform", 6000, 12, 32),
                                        ("plummer", 5000, 15, 32),
                                        ("uniform", 4000, 4, 8),
                                        ("plummer", 3000, 1, 16)])
def test_mfma_form_close_to_plain(torch_cuda, form, kind, n, p, q):
    torch = torch_cuda
    from aran_amd.fmm import Fmm3d
    (x, y, z, m), lb, ub = _case(kind, n, 9)
    f = Fmm3d()
    dev = _dev(torch, x, y, z, m)
    f.set_precision("plain")
    p0, g0 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    s0 = f.stats()
    f.set_precision(form)
    p1, g1 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    assert f.stats() == s0
    l2, pw, gl2 = BARS[form]
    assert np.linalg.norm(p1 - p0) / np.linalg.norm(p0) <= l2
    assert np.max(np.abs(p1 - p0) / np.abs(p0)) <= pw
    assert np.linalg.norm(g1 - g0) / np.linalg.norm(g0) <= gl2
    p2, g2 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    assert np.array_equal(p1, p2) and np.array_equal(g1, g2)


def test_mfma_signed_charges_against_direct_sum(torch_cuda):
    """Signed charges in [-1, 1) (the precision-sweep distribution), N=2^16,
    p=10, leaf 64: BF16x6 potential L2 error against the FP64 direct sum on
    4096 targets within 1.5x of the plain form's."""
    torch = torch_cuda
    from aran_amd.fmm import Fmm3d
    n = 1 << 16
    x, y, z, qv = W.signed_cube(n, 31)
    dev = _dev(torch, x, y, z, qv)
    f = Fmm3d()
    tg = W.sample_targets(n, 4096, 31)
    dp, _ = f.direct(*dev, torch.from_numpy(tg))
    dp = dp.cpu().numpy()
    errs = {}
    for form in ("plain", "bf16x6"):
        f.set_precision(form)
        pot, _ = f.solve(*dev, (0, 0, 0), (1, 1, 1), 10, 64)
        pp = pot.cpu().numpy()[tg]
        errs[form] = np.linalg.norm(pp - dp) / np.linalg.norm(dp)
    assert errs["bf16x6"] <= 1.5 * errs["plain"] + 1e-9, errs


def test_mfma_batch_and_semifar(torch_cuda):
    """The MFMA form under a batched solve of independent systems (each
    system's result within the BF16x6 bar of its plain solve) and with
    semifar pairs (against the plain form with the same threshold)."""
    torch = torch_cuda
    from aran_amd.fmm import Fmm3d
    systems = [_case("plummer", 4000, 3), _case("uniform", 6000, 4),
               _case("plummer", 2500, 5)]
    offs = np.cumsum([0] + [len(s[0][0]) for s in systems])
    cat = [np.concatenate([s[0][i] for s in systems]) for i in range(4)]
    dev = _dev(torch, *cat)
    lbs = [s[1] for s in systems]
    ubs = [s[2] for s in systems]
    f = Fmm3d()
    out = {}
    for form in ("plain", "bf16x6"):
        f.set_precision(form)
        out[form] = [t.cpu().numpy() for t in
                     f.solve_batch(offs, *dev, lbs, ubs, 8, 32)]
    for k in range(len(systems)):
        sl = slice(offs[k], offs[k + 1])
        a, b = out["bf16x6"][0][sl], out["plain"][0][sl]
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 5e-7, k
    (x, y, z, m), lb, ub = _case("plummer", 20_000, 8)
    d1 = _dev(torch, x, y, z, m)
    res = {}
    for form in ("plain", "bf16x6"):
        f.set_precision(form)
        f.set_semifar(20)
        res[form] = f.solve(*d1, lb, ub, 10, 32)[0].cpu().numpy()
        assert f.stats()["p2l"] > 0
    f.set_semifar(None)
    a, b = res["bf16x6"], res["plain"]
    assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 5e-7


@pytest.mark.parametrize("kind,n,p,q", [("plummer", 1 << 17, 10, 64),
                                        ("uniform", 20_000, 8, 32),
                                        ("plummer", 8000, 6, 16),
                                        ("uniform", 6000, 12, 32),
                                        ("plummer", 5000, 15, 32),
                                        ("uniform", 3000, 2, 8),
                                        ("plummer", 3000, 1, 16)])
def test_one_wave_kernel_equals_eight_wave_kernel(torch_cuda, monkeypatch,
                                                  kind, n, p, q):
    """The three-product far field has two kernels that give every
    accumulator the same MFMA sequence: k_m2l_mfma1w (one wave per SIMD,
    256 accumulators; the default) and k_m2l_mfma (eight waves per
    workgroup; ARAN_MFMA_KERNEL=8w).  Their solves agree bit for bit."""
    torch = torch_cuda
    from aran_amd.fmm import Fmm3d
    (x, y, z, m), lb, ub = _case(kind, n, 17)
    dev = _dev(torch, x, y, z, m)
    f = Fmm3d()
    f.set_precision("bf16x3")
    monkeypatch.setenv("ARAN_MFMA_KERNEL", "8w")
    p8, g8 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    monkeypatch.delenv("ARAN_MFMA_KERNEL")
    p1, g1 = [t.cpu().numpy() for t in f.solve(*dev, lb, ub, p, q)]
    assert f.far_plan()["tiles"] > 0
    assert np.array_equal(p1, p8) and np.array_equal(g1, g8)


def test_one_wave_kernel_batch_equals_eight_wave_kernel(torch_cuda,
                                                         monkeypatch):
    """The same for a batch of independent systems of mixed sizes (tiles of
    several systems and levels in one launch, workgroups with missing
    tiles)."""
    torch = torch_cuda
    from aran_amd.fmm import Fmm3d
    systems = [_case("plummer", 4000, 3), _case("uniform", 6000, 4),
               _case("plummer", 2500, 5), _case("uniform", 700, 6),
               _case("plummer", 30_000, 7)]
    offs = np.cumsum([0] + [len(s[0][0]) for s in systems])
    cat = [np.concatenate([s[0][i] for s in systems]) for i in range(4)]
    dev = _dev(torch, *cat)
    lbs = [s[1] for s in systems]
    ubs = [s[2] for s in systems]
    f = Fmm3d()
    f.set_precision("bf16x3")
    monkeypatch.setenv("ARAN_MFMA_KERNEL", "8w")
    a = [t.cpu().numpy() for t in f.solve_batch(offs, *dev, lbs, ubs, 8, 32)]
    monkeypatch.delenv("ARAN_MFMA_KERNEL")
    b = [t.cpu().numpy() for t in f.solve_batch(offs, *dev, lbs, ubs, 8, 32)]
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
