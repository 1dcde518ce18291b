"""FMA particle files (tests/newtonfield3parallel.c:1020-1086, 1186-1212).

Host-only: the solver is created and filled but never solved, so these run
without a GPU.
"""
import ctypes as C

import numpy as np
import pytest

from aran_amd import solver as S


def _soa_solver(x, y, z, q, lb, ub):
    L = S.lib()
    lbv, ubv = S.VsgVector3d(*lb), S.VsgVector3d(*ub)
    tree = L.vsg_prtree3d_new_full(C.byref(lbv), C.byref(ubv), None, None,
                                   None, 8)
    dev = L.aran_development3d_new(0, 4)
    sol = L.aran_solver3d_new(tree, L.aran_development3d_get_type(), dev,
                              S.fn("aran_development3d_set_zero"))
    assert sol
    rc = L.aran_solver3d_insert_points_soa(sol, len(x), x.ctypes.data,
                                          y.ctypes.data, z.ctypes.data,
                                          q.ctypes.data)
    assert rc == 0
    return sol


def test_write_fma_header_and_lines(tmp_path):
    rng = np.random.default_rng(3)
    n = 257
    x, y, z = (rng.uniform(-1, 1, n) for _ in range(3))
    q = rng.uniform(-1, 1, n)
    lb, ub = (-1.0, -2.0, 0.5), (3.0, 1.0, 1.5)
    sol = _soa_solver(x, y, z, q, lb, ub)
    path = tmp_path / "out.fma"
    try:
        S.write_fma(sol, path)
    finally:
        S.lib().aran_solver3d_free(sol)
    lines = path.read_text().splitlines()
    assert len(lines) == n + 1
    head = lines[0].split()
    assert int(head[0]) == n
    # half-width is the largest half-extent of the bounds (here x: 2.0),
    # centre the bounds' centre
    assert float(head[1]) == 2.0
    assert [float(v) for v in head[2:]] == [1.0, -0.5, 1.0]
    body = np.array([[float(v) for v in ln.split()] for ln in lines[1:]])
    # %g keeps 6 significant digits
    np.testing.assert_allclose(body, np.stack([x, y, z, q], 1), rtol=1e-5,
                               atol=1e-6)
    assert lines[1].split()[0] == "%g" % x[0]


def test_read_points_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    n = 100
    pts = rng.uniform(-1, 1, (n, 4))
    path = tmp_path / "in.txt"
    path.write_text("".join("%d %.17g %.17g %.17g %.17g\n" % (i, *p)
                            for i, p in enumerate(pts)))
    x, y, z, q = S.read_fma_points(path)
    assert np.array_equal(np.stack([x, y, z, q], 1), pts)


def test_read_points_sizing_and_truncation(tmp_path):
    path = tmp_path / "in.txt"
    path.write_text("0 1 2 3 4\n1 5 6 7 8\n2 9 10 11 12\n")
    L = S.lib()
    f = S._open(path, b"r")
    try:
        a = [np.zeros(2) for _ in range(4)]
        n = L.aran_fma_read_points(f, 2, *[v.ctypes.data for v in a])
    finally:
        S._stdio().fclose(f)
    # the count covers every line; only max_points of them are stored
    assert n == 3
    assert list(a[0]) == [1.0, 5.0] and list(a[3]) == [4.0, 8.0]


def test_read_points_empty_and_malformed(tmp_path):
    empty = tmp_path / "empty.txt"
    empty.write_text("")
    x, y, z, q = S.read_fma_points(empty)
    assert len(x) == 0
    bad = tmp_path / "bad.txt"
    bad.write_text("0 1 2 3 4\n1 5 six 7 8\n")
    with pytest.raises(ValueError):
        S.read_fma_points(bad)


def test_written_file_reads_back_as_points(tmp_path):
    """A file written by write_fma, with the header dropped and an index
    column added, loads back to the same (rounded) particles."""
    rng = np.random.default_rng(9)
    n = 50
    x, y, z = (rng.uniform(-1, 1, n) for _ in range(3))
    q = np.ones(n)
    sol = _soa_solver(x, y, z, q, (-1, -1, -1), (1, 1, 1))
    out = tmp_path / "w.fma"
    try:
        S.write_fma(sol, out)
    finally:
        S.lib().aran_solver3d_free(sol)
    body = out.read_text().splitlines()[1:]
    inp = tmp_path / "r.txt"
    inp.write_text("".join("%d %s\n" % (i, ln) for i, ln in enumerate(body)))
    rx, ry, rz, rq = S.read_fma_points(inp)
    np.testing.assert_allclose(np.stack([rx, ry, rz, rq], 1),
                               np.stack([x, y, z, q], 1), rtol=1e-5,
                               atol=1e-6)
