"""ctypes binding of libaran_amd.so.  Fails loudly when it is not built."""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libaran_amd.so")

_lib = None

vp, dp, lp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_long), C.c_int


def _declare(L):
    L.aran_last_error.restype = C.c_char_p
    L.aran_last_error.argtypes = []
    L.aran_init.restype = None
    L.aran_fmm3d_new.restype = vp
    L.aran_fmm3d_new.argtypes = [vp]
    L.aran_fmm3d_free.restype = None
    L.aran_fmm3d_free.argtypes = [vp]
    L.aran_fmm3d_stream.restype = vp
    L.aran_fmm3d_stream.argtypes = [vp]
    L.aran_fmm3d_solve.restype = ip
    L.aran_fmm3d_solve.argtypes = [vp, C.c_long, vp, vp, vp, vp, dp, dp, ip, ip,
                                   vp, vp]
    L.aran_fmm3d_solve_batch.restype = ip
    L.aran_fmm3d_solve_batch.argtypes = [vp, ip, C.POINTER(C.c_long), vp, vp,
                                         vp, vp, dp, dp, ip, ip, vp, vp]
    L.aran_fmm3d_get_stats.restype = ip
    L.aran_fmm3d_get_stats.argtypes = [vp, lp]
    L.aran_fmm3d_set_precision.restype = ip
    L.aran_fmm3d_set_precision.argtypes = [vp, ip]
    L.aran_fmm3d_set_timings.restype = ip
    L.aran_fmm3d_set_timings.argtypes = [vp, ip]
    L.aran_fmm3d_get_timings.restype = ip
    L.aran_fmm3d_get_timings.argtypes = [vp, C.POINTER(C.c_float)]
    L.aran_fmm3d_export_tree.restype = ip
    L.aran_fmm3d_export_tree.argtypes = [vp, vp, vp, vp, vp, vp]
    L.aran_fmm3d_export_lists.restype = ip
    L.aran_fmm3d_export_lists.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp]
    L.aran_fmm3d_get_kernel_timings.restype = ip
    L.aran_fmm3d_get_kernel_timings.argtypes = [vp, C.POINTER(C.c_float)]
    L.aran_fmm3d_get_far_plan.restype = ip
    L.aran_fmm3d_get_far_plan.argtypes = [vp, C.POINTER(C.c_int)]
    L.aran_fmm3d_set_autotune.restype = ip
    L.aran_fmm3d_set_autotune.argtypes = [vp, ip]
    L.aran_fmm3d_set_semifar.restype = ip
    L.aran_fmm3d_set_semifar.argtypes = [vp, C.c_long]
    L.aran_fmm3d_semifar_threshold_auto.restype = C.c_long
    L.aran_fmm3d_semifar_threshold_auto.argtypes = [ip]
    L.aran_fmm3d_export_semifar.restype = ip
    L.aran_fmm3d_export_semifar.argtypes = [vp, vp, vp, vp, vp]
    L.aran_fmm3d_direct.restype = ip
    L.aran_fmm3d_direct.argtypes = [vp, C.c_long, vp, vp, vp, vp, C.c_long, vp,
                                    vp, vp]
    L.aran_fmm3d_set_translations.restype = ip
    L.aran_fmm3d_set_translations.argtypes = [vp, ip, ip, ip]
    L.aran_fmm3d_solve_noalloc.restype = ip
    L.aran_fmm3d_solve_noalloc.argtypes = [vp, C.c_long, vp, vp, vp, vp, dp,
                                           dp, ip, ip, vp, vp]
    L.aran_fmm3d_export_expansions.restype = ip
    L.aran_fmm3d_export_expansions.argtypes = [vp, vp, vp]


def lib():
    """The loaded engine library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "aran_amd: HIP library %s is not built; run "
                "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def last_error():
    return lib().aran_last_error().decode()


def check(rc, what):
    if rc != 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, last_error()))
