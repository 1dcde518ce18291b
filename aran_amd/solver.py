"""Python mirror of the LibAran-compatible C API (include/aran/*.h).

Structs and functions are declared with ctypes exactly as the C headers
declare them, so a test written against this module exercises the same
entry points a C program of the reference (tests/newtonpot3.c,
tests/newtonfield3.c) would call.
"""
import ctypes as C

from ._lib import lib as _load


class VsgVector3d(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class VsgPRTreeKey3d(C.Structure):
    _fields_ = [("x", C.c_uint64), ("y", C.c_uint64), ("z", C.c_uint64),
                ("depth", C.c_uint8)]


class VsgPRTree3dNodeInfo(C.Structure):
    pass


VsgPRTree3dNodeInfo._fields_ = [
    ("lbound", VsgVector3d), ("ubound", VsgVector3d), ("center", VsgVector3d),
    ("point_count", C.c_uint), ("point_list", C.c_void_p),
    ("isleaf", C.c_int), ("depth", C.c_uint8),
    ("father_info", C.POINTER(VsgPRTree3dNodeInfo)),
    ("user_data", C.c_void_p), ("id", VsgPRTreeKey3d)]


class Complex128(C.Structure):
    """gcomplex128 returned by value (two doubles, as C99 double complex)"""
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class AranDevelopment3d(C.Structure):
    _fields_ = [("posdeg", C.c_uint8), ("negdeg", C.c_uint8),
                ("multipole", C.POINTER(C.c_double)),
                ("local", C.POINTER(C.c_double))]


class PointAccumPot(C.Structure):
    """struct _PointAccum of tests/newtonpot3.c:43-54"""
    _fields_ = [("vector", VsgVector3d), ("density", C.c_double),
                ("accum", C.c_double * 2), ("id", C.c_uint)]


class PointAccumField(C.Structure):
    """struct _PointAccum of tests/newtonfield3.c:46-55"""
    _fields_ = [("vector", VsgVector3d), ("density", C.c_double),
                ("field", VsgVector3d), ("id", C.c_uint)]


class PointAccumBoth(C.Structure):
    _fields_ = [("vector", VsgVector3d), ("density", C.c_double),
                ("pot", C.c_double), ("field", VsgVector3d)]


_declared = False


def lib():
    global _declared
    L = _load()
    if _declared:
        return L
    vp, d, u = C.c_void_p, C.c_double, C.c_uint
    dev_p = C.POINTER(AranDevelopment3d)
    ni_p = C.POINTER(VsgPRTree3dNodeInfo)
    v3_p = C.POINTER(VsgVector3d)
    L.vsg_prtree3d_new_full.restype = vp
    L.vsg_prtree3d_new_full.argtypes = [v3_p, v3_p, vp, vp, vp, u]
    L.aran_development3d_get_type.restype = C.c_size_t
    L.aran_development3d_new.restype = dev_p
    L.aran_development3d_new.argtypes = [C.c_uint8, C.c_uint8]
    L.aran_development3d_free.argtypes = [dev_p]
    L.aran_development3d_set_zero.argtypes = [dev_p]
    L.aran_development3d_p2m.argtypes = [v3_p, d, ni_p, dev_p]
    for f in ("m2m", "m2l", "l2l"):
        getattr(L, "aran_development3d_%s_rotate" % f).argtypes = [
            ni_p, dev_p, ni_p, dev_p]
    L.aran_development3d_local_evaluate.restype = Complex128
    L.aran_development3d_local_evaluate.argtypes = [ni_p, dev_p, v3_p]
    L.aran_development3d_local_gradient_evaluate.argtypes = [ni_p, dev_p, v3_p,
                                                             v3_p]
    L.aran_development3d_p2l.argtypes = [v3_p, d, ni_p, dev_p]
    L.aran_development3d_multipole_evaluate.restype = Complex128
    L.aran_development3d_multipole_evaluate.argtypes = [ni_p, dev_p, v3_p]
    L.aran_development3d_multipole_gradient_evaluate.argtypes = [
        ni_p, dev_p, v3_p, v3_p]
    L.aran_solver3d_new.restype = vp
    L.aran_solver3d_new.argtypes = [vp, C.c_size_t, vp, vp]
    L.aran_solver3d_free.argtypes = [vp]
    L.aran_solver3d_set_functions.argtypes = [vp] + [vp] * 6
    L.aran_solver3d_set_functions_full.argtypes = [vp] + [vp] * 8 + [u]
    L.aran_solver3d_reinit_stats.argtypes = [vp]
    L.aran_solver3d_set_particle_layout.argtypes = [vp, C.c_ssize_t,
                                                    C.c_ssize_t, C.c_int,
                                                    C.c_ssize_t]
    L.aran_solver3d_insert_point.argtypes = [vp, vp]
    L.aran_solver3d_remove_point.argtypes = [vp, vp]
    L.aran_solver3d_remove_point.restype = C.c_int
    L.aran_solver3d_point_count.restype = u
    L.aran_solver3d_point_count.argtypes = [vp]
    L.aran_solver3d_depth.restype = u
    L.aran_solver3d_depth.argtypes = [vp]
    L.aran_solver3d_solve.argtypes = [vp]
    L.aran_solver3d_status.restype = C.c_int
    L.aran_solver3d_status.argtypes = [vp]
    L.aran_solver3d_set_precision.restype = C.c_int
    L.aran_solver3d_set_precision.argtypes = [vp, C.c_int]
    L.aran_solver3d_get_stats.argtypes = [vp] + [C.POINTER(C.c_long)] * 11
    L.aran_solver3d_insert_points_soa.restype = C.c_int
    L.aran_solver3d_insert_points_soa.argtypes = [vp, C.c_long] + [vp] * 4
    L.aran_solver3d_get_results_soa.restype = C.c_int
    L.aran_solver3d_get_results_soa.argtypes = [vp, vp, vp]
    L.aran_solver3d_set_tolerance.argtypes = [vp, d]
    L.aran_solver3d_find_point.restype = vp
    L.aran_solver3d_find_point.argtypes = [vp, vp]
    L.aran_solver3d_write_fma.restype = C.c_int
    L.aran_solver3d_write_fma.argtypes = [vp, vp]
    L.aran_fma_read_points.restype = C.c_long
    L.aran_fma_read_points.argtypes = [vp, C.c_long] + [vp] * 4
    _declared = True
    return L


def fn(name):
    """Address of an exported function (for callback registration)."""
    return C.cast(getattr(lib(), name), C.c_void_p)


def node_info(center, depth=0):
    ni = VsgPRTree3dNodeInfo()
    ni.center = VsgVector3d(*center)
    ni.depth = depth
    ni.isleaf = 1
    return ni


def get_stats(solver):
    vals = [C.c_long() for _ in range(11)]
    lib().aran_solver3d_get_stats(solver, *[C.byref(v) for v in vals])
    keys = ("zero", "p2p", "p2m", "m2m", "m2l", "l2l", "l2p", "p2l", "m2p",
            "p2p_remote", "m2l_remote")
    return dict(zip(keys, [v.value for v in vals]))


_libc = None


def _stdio():
    global _libc
    if _libc is None:
        _libc = C.CDLL(None)
        _libc.fopen.restype = C.c_void_p
        _libc.fopen.argtypes = [C.c_char_p, C.c_char_p]
        _libc.fclose.restype = C.c_int
        _libc.fclose.argtypes = [C.c_void_p]
    return _libc


def _open(path, mode):
    f = _stdio().fopen(str(path).encode(), mode)
    if not f:
        raise OSError("cannot open %s" % path)
    return f


def write_fma(solver, path):
    """Write the solver's particles as an FMA file (aran_solver3d_write_fma)."""
    f = _open(path, b"w")
    try:
        rc = lib().aran_solver3d_write_fma(solver, f)
    finally:
        _stdio().fclose(f)
    if rc != 0:
        raise RuntimeError("aran_solver3d_write_fma failed (%d)" % rc)


def read_fma_points(path):
    """Read an "index x y z charge" particle file (aran_fma_read_points):
    returns x, y, z, charge as float64 numpy arrays."""
    import numpy as np
    f = _open(path, b"r")
    try:
        n = lib().aran_fma_read_points(f, 0, None, None, None, None)
    finally:
        _stdio().fclose(f)
    if n < 0:
        raise ValueError("malformed particle file %s" % path)
    out = [np.zeros(n) for _ in range(4)]
    f = _open(path, b"r")
    try:
        m = lib().aran_fma_read_points(f, n, *[a.ctypes.data for a in out])
    finally:
        _stdio().fclose(f)
    if m != n:
        raise ValueError("particle file %s changed while reading" % path)
    return tuple(out)
