"""aran_fmm3d_solve_noalloc (the no-allocation solve) against
aran_fmm3d_solve on the same context settings, through the C-ABI: bitwise
equal potentials and gradients on two calls in a row, the second with a
smaller input (the scratch keeps the first call's contents), in the plain
and mixed forms; inputs beyond its limits are refused with an error.  The
no-allocation solve runs the plain form's far field on the per-pair rotate
kernel (the dense FP64 operator tables would not fit its scratch), so the
allocating solve it is compared with is set to that kernel too
(ARAN_PLAIN_FAR=rotate)."""
import ctypes as C

import numpy as np
import pytest
import torch  # noqa: F401

from aran_amd import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(f, dev, lb, ub, p, q, pot, grad):
    from aran_amd._lib import lib
    lbv = (C.c_double * 3)(*lb)
    ubv = (C.c_double * 3)(*ub)
    return lib().aran_fmm3d_solve_noalloc(
        f._ctx, dev[0].numel(), *[C.c_void_p(t.data_ptr()) for t in dev],
        lbv, ubv, p, q, C.c_void_p(pot.data_ptr()),
        C.c_void_p(grad.data_ptr()))


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_noalloc_equals_allocating_solve(torch_cuda, form, monkeypatch):
    from aran_amd.fmm import Fmm3d
    torch = torch_cuda
    monkeypatch.setenv("ARAN_PLAIN_FAR", "rotate")
    f = Fmm3d()
    f.set_precision(form)
    for n, p, q, kind in ((20_000, 10, 32, "plummer"), (5000, 8, 20,
                                                       "uniform")):
        x, y, z, m = W.make(kind, n, 0xA2A0 + n)
        lb, ub = ((-1., -1., -1.), (1., 1., 1.)) if kind == "plummer" else \
            ((0., 0., 0.), (1., 1., 1.))
        dev = [torch.from_numpy(a).cuda() for a in (x, y, z, m)]
        ref_p, ref_g = f.solve(*dev, lb, ub, p, q)
        pot = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
        grad = torch.full((n, 3), np.nan, dtype=torch.float64, device="cuda")
        rc = _noalloc(f, dev, lb, ub, p, q, pot, grad)
        assert rc == 0
        torch.cuda.synchronize()
        assert torch.equal(pot, ref_p), (form, n)
        assert torch.equal(grad, ref_g), (form, n)


def test_noalloc_refuses_beyond_limits(torch_cuda, monkeypatch):
    from aran_amd._lib import lib
    from aran_amd.fmm import Fmm3d
    torch = torch_cuda
    monkeypatch.setenv("ARAN_PLAIN_FAR", "rotate")
    f = Fmm3d()
    n = 70_000
    x, y, z, m = W.uniform_cube(n, 5)
    dev = [torch.from_numpy(a).cuda() for a in (x, y, z, m)]
    pot = torch.empty(n, dtype=torch.float64, device="cuda")
    grad = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    assert _noalloc(f, dev, (0, 0, 0), (1, 1, 1), 8, 32, pot, grad) == -1
    assert b"limits" in lib().aran_last_error()
    dev = [t[:60_000] for t in dev]
    assert _noalloc(f, dev, (0, 0, 0), (1, 1, 1), 12, 32, pot, grad) == -1
    f.set_precision("bf16x3")
    assert _noalloc(f, dev, (0, 0, 0), (1, 1, 1), 8, 32, pot, grad) == -1
    # a tree too large for the scratch (leaf capacity 1) is refused, and the
    # next call still works
    f.set_precision("plain")
    assert _noalloc(f, dev, (0, 0, 0), (1, 1, 1), 8, 1, pot, grad) == -3
    assert b"scratch" in lib().aran_last_error()
    small = [t[:3000] for t in dev]
    ref_p, ref_g = f.solve(*small, (0, 0, 0), (1, 1, 1), 6, 16)
    assert _noalloc(f, small, (0, 0, 0), (1, 1, 1), 6, 16, pot, grad) == 0
    torch.cuda.synchronize()
    assert torch.equal(pot[:3000], ref_p) and torch.equal(grad[:3000], ref_g)
