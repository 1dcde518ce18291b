"""Python mirror of the serving entry points (include/aran/aranserver3d.h).

Requests are numpy FP64 host arrays; ctypes releases the GIL during each
call, so one thread can submit while another waits."""
import ctypes as C

import numpy as np

from ._lib import check, last_error, lib

PLAIN, MIXED = 0, 1
EBUDGET = -4
WORKER_BYTES = 512e6  # ARAN_SERVER3D_WORKER_BYTES
_declared = False


def _L():
    global _declared
    L = lib()
    if not _declared:
        vp, d, ip, lp = C.c_void_p, C.c_double, C.c_int, C.c_long
        L.aran_server3d_footprint.restype = d
        L.aran_server3d_footprint.argtypes = [lp, ip, ip]
        L.aran_server3d_new.restype = vp
        L.aran_server3d_new.argtypes = [d, ip, ip, ip]
        L.aran_server3d_free.restype = None
        L.aran_server3d_free.argtypes = [vp]
        L.aran_server3d_submit.restype = lp
        L.aran_server3d_submit.argtypes = [vp, lp, vp, vp, vp, vp,
                                           C.POINTER(d), C.POINTER(d), ip, ip,
                                           ip]
        L.aran_server3d_wait.restype = ip
        L.aran_server3d_wait.argtypes = [vp, lp, vp, vp, vp]
        L.aran_server3d_prepare.restype = ip
        L.aran_server3d_prepare.argtypes = [vp, lp, C.POINTER(ip), ip, ip]
        L.aran_server3d_pause.restype = ip
        L.aran_server3d_pause.argtypes = [vp, ip]
        L.aran_server3d_stats.restype = ip
        L.aran_server3d_stats.argtypes = [vp, vp]
        _declared = True
    return L


def footprint(n, order, leafcap):
    return _L().aran_server3d_footprint(int(n), int(order), int(leafcap))


class Server:
    def __init__(self, budget_bytes=200e9, nworkers=2, precision="mixed",
                 max_batch=64):
        code = {"plain": PLAIN, "mixed": MIXED, "bf16x3": 2,
                "bf16x6": 3}.get(precision, precision)
        self._s = _L().aran_server3d_new(float(budget_bytes), int(nworkers),
                                         int(code), int(max_batch))
        if not self._s:
            raise RuntimeError("aran_server3d_new failed: " + last_error())
        self._n = {}
        self._keep = {}

    def close(self):
        if self._s:
            _L().aran_server3d_free(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, x, y, z, q, lb, ub, order, leafcap, want_grad=True):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z, q)]
        n = arrs[0].size
        lbv = (C.c_double * 3)(*[float(v) for v in lb])
        ubv = (C.c_double * 3)(*[float(v) for v in ub])
        rid = _L().aran_server3d_submit(self._s, n, *[a.ctypes.data for a in arrs],
                                        lbv, ubv, int(order), int(leafcap),
                                        int(bool(want_grad)))
        if rid < 0:
            return rid
        self._n[rid] = (n, bool(want_grad))
        return rid

    def wait(self, rid):
        n, want_grad = self._n.pop(rid)
        pot = np.empty(n)
        grad = np.empty((n, 3)) if want_grad else None
        times = np.zeros(3)
        rc = _L().aran_server3d_wait(self._s, rid, pot.ctypes.data,
                                     grad.ctypes.data if want_grad else None,
                                     times.ctypes.data)
        check(rc, "aran_server3d_wait")
        return pot, grad, {"latency": times[0], "queue": times[1],
                           "service": times[2]}

    def prepare(self, n_max, orders, leafcap):
        """Reserve staging and workspace for batches of up to n_max particles
        and build the tables of each order, on every worker, before traffic."""
        ov = (C.c_int * max(1, len(orders)))(*[int(o) for o in orders])
        check(_L().aran_server3d_prepare(self._s, int(n_max), ov, len(orders),
                                         int(leafcap)), "aran_server3d_prepare")

    def pause(self, paused=True):
        check(_L().aran_server3d_pause(self._s, int(bool(paused))), "pause")

    def stats(self):
        v = np.zeros(4)
        check(_L().aran_server3d_stats(self._s, v.ctypes.data), "stats")
        return {"batches": int(v[0]), "served": int(v[1]),
                "max_batch": int(v[2]), "peak_bytes": float(v[3])}
