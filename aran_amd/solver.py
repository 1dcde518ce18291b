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
    """struct _AranDevelopment3d (aran/arandevelopment3d.h:36-42): two
    opaque AranSphericalSeriesd pointers"""
    _fields_ = [("multipole", C.c_void_p), ("local", C.c_void_p)]


# VsgPRTree3dFunc / VsgPRTree3dNFIsleafFunc callbacks (include/vsg/vsgd.h)
PRTREE_FUNC = C.CFUNCTYPE(None, C.POINTER(VsgPRTree3dNodeInfo), C.c_void_p)
NF_ISLEAF_FUNC = C.CFUNCTYPE(C.c_int, C.POINTER(VsgPRTree3dNodeInfo),
                             C.c_void_p)
G_IN_ORDER, G_PRE_ORDER, G_POST_ORDER, G_LEVEL_ORDER = 0, 1, 2, 3


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
        for suffix in ("_rotate", "", "_kkylin"):
            getattr(L, "aran_development3d_%s%s" % (f, suffix)).argtypes = [
                ni_p, dev_p, ni_p, dev_p]
    L.aran_development3d_clone.restype = dev_p
    L.aran_development3d_clone.argtypes = [dev_p]
    L.aran_development3d_copy.argtypes = [dev_p, dev_p]
    L.aran_development3d_write.argtypes = [dev_p, vp]
    ser = vp
    L.aran_spherical_seriesd_new.restype = ser
    L.aran_spherical_seriesd_new.argtypes = [C.c_uint8, C.c_uint8]
    L.aran_spherical_seriesd_free.argtypes = [ser]
    L.aran_spherical_seriesd_clone.restype = ser
    L.aran_spherical_seriesd_clone.argtypes = [ser]
    L.aran_spherical_seriesd_copy.argtypes = [ser, ser]
    L.aran_spherical_seriesd_get_term.restype = C.POINTER(C.c_double)
    L.aran_spherical_seriesd_get_term.argtypes = [ser, C.c_int, C.c_int]
    L.aran_spherical_seriesd_get_posdeg.restype = C.c_uint8
    L.aran_spherical_seriesd_get_posdeg.argtypes = [ser]
    L.aran_spherical_seriesd_get_negdeg.restype = C.c_uint8
    L.aran_spherical_seriesd_get_negdeg.argtypes = [ser]
    L.aran_spherical_seriesd_set_zero.argtypes = [ser]
    L.aran_spherical_seriesd_write.argtypes = [ser, vp]
    L.aran_spherical_seriesd_evaluate.restype = Complex128
    L.aran_spherical_seriesd_evaluate.argtypes = [ser, v3_p]
    L.aran_spherical_seriesd_gradient_evaluate.argtypes = [ser, v3_p, v3_p]
    for f in ("translate", "to_local"):
        for suffix in ("", "_kkylin", "_rotate"):
            getattr(L, "aran_spherical_seriesd_%s%s" % (f, suffix)).argtypes = [
                ser, v3_p, ser, v3_p]
    L.aran_solver3d_traverse.argtypes = [vp, C.c_int, PRTREE_FUNC, vp]
    L.aran_solver3d_set_nf_isleaf.argtypes = [vp, NF_ISLEAF_FUNC, vp]
    L.aran_solver3d_set_children_order_hilbert.argtypes = [vp]
    L.aran_solver3d_set_children_order_default.argtypes = [vp]
    L.aran_solver3d_insert_point_local.argtypes = [vp, vp]
    L.vsg_vector3d_norm.restype = d
    L.vsg_vector3d_norm.argtypes = [v3_p]
    L.vsg_vector3d_dotp.restype = d
    L.vsg_vector3d_dotp.argtypes = [v3_p, v3_p]
    for f in ("add", "sub"):
        getattr(L, "vsg_vector3d_" + f).argtypes = [v3_p, v3_p, v3_p]
    L.vsg_vector3d_scalp.argtypes = [v3_p, d, v3_p]
    L.vsg_vector3d_set.argtypes = [v3_p, d, d, d]
    L.vsg_vector3d_to_spherical_internal.argtypes = [v3_p] + [
        C.POINTER(C.c_double)] * 5
    L.vsg_vector3d_from_spherical_internal.argtypes = [v3_p] + [d] * 5
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


def series_terms(series, positive):
    """The local (positive) or multipole (negative) terms of a series as a
    numpy complex128 view, m >= 0 packed by degree."""
    import numpy as np
    L = lib()
    if positive:
        n = L.aran_spherical_seriesd_get_posdeg(series) + 1
        ptr = L.aran_spherical_seriesd_get_term(series, 0, 0)
    else:
        n = L.aran_spherical_seriesd_get_negdeg(series)
        if n == 0:
            return np.zeros(0, dtype=np.complex128)
        ptr = L.aran_spherical_seriesd_get_term(series, -1, 0)
    cnt = n * (n + 1) // 2
    return np.ctypeslib.as_array(ptr, shape=(2 * cnt,)).view(np.complex128)


def multipole(dev):
    """Multipole terms (degrees < negdeg) of an AranDevelopment3d pointer."""
    return series_terms(dev.contents.multipole, False)


def local(dev):
    """Local terms (degrees <= posdeg) of an AranDevelopment3d pointer."""
    return series_terms(dev.contents.local, True)


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
