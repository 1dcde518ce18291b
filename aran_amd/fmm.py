"""Python mirror of the device-resident entry points (include/aran/aranfmm3d.h).

Inputs and outputs are torch float64 tensors on the GPU; the context
launches on torch's current stream, so no extra synchronisation is needed
around calls.
"""
import ctypes as C

import numpy as np

from ._lib import check, lib

STAT_NAMES = ("zero", "p2p", "p2m", "m2m", "m2l", "l2l", "l2p", "depth",
              "nodes", "leaves", "far_entries", "near_entries", "p2l", "m2p",
              "semifar_pairs")
TIMER_NAMES = ("tree", "lists", "up", "m2l", "down", "near", "total")
KERNEL_TIMER_NAMES = ("p2m", "m2m", "l2l", "l2p")


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class Fmm3d:
    """One FMM context (device workspace kept between solves)."""

    def __init__(self, stream=None):
        import torch
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        self._ctx = lib().aran_fmm3d_new(C.c_void_p(stream))
        if not self._ctx:
            raise RuntimeError("aran_fmm3d_new failed")
        self.n = 0

    def close(self):
        if self._ctx:
            lib().aran_fmm3d_free(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    PLAIN, MIXED = 0, 1

    def set_precision(self, precision):
        """"plain" (FP64 throughout), "mixed" (FP32 M2L and P2P), "bf16x3" /
        "bf16x6" (M2L on MFMA with BF16 operands in 3 / 6 products, FP32
        P2P)."""
        code = {"plain": 0, "mixed": 1, "bf16x3": 2, "bf16x6": 3}.get(
            precision, precision)
        check(lib().aran_fmm3d_set_precision(self._ctx, int(code)),
              "set_precision")

    def set_semifar(self, threshold):
        """Semifar P2L / M2P pairs for nodes holding more than `threshold`
        particles; None or a negative value disables them (the cost-model
        threshold of an order is semifar_threshold_auto(order))."""
        t = -1 if threshold is None else int(threshold)
        check(lib().aran_fmm3d_set_semifar(self._ctx, t), "set_semifar")

    @staticmethod
    def semifar_threshold_auto(order):
        return int(lib().aran_fmm3d_semifar_threshold_auto(int(order)))

    def set_translations(self, m2m="rotate", m2l="rotate", l2l="rotate"):
        """Translation flavours of M2M, M2L, L2L: "rotate" (the operators of
        record), "normal" (O(p^4) formulas) or "kylin" (plain form only)."""
        code = {"rotate": 0, "normal": 1, "kylin": 2}
        check(lib().aran_fmm3d_set_translations(
            self._ctx, code[m2m], code[m2l], code[l2l]), "set_translations")

    def export_expansions(self, order):
        """Multipoles and locals of the last solve (node order of
        export_tree) as complex128 arrays [nodes][terms]."""
        nn = self.stats()["nodes"]
        mc, lc = order * (order + 1) // 2, (order + 1) * (order + 2) // 2
        M = np.zeros((nn, mc), dtype=np.complex128)
        L = np.zeros((nn, lc), dtype=np.complex128)
        check(lib().aran_fmm3d_export_expansions(
            self._ctx, M.ctypes.data, L.ctypes.data), "export_expansions")
        return M, L

    def set_autotune(self, on=True):
        """Time the far-field workgroup schedules on first use of each (form,
        order, size bucket) and keep the fastest (results do not depend on
        the schedule)."""
        check(lib().aran_fmm3d_set_autotune(self._ctx, int(on)),
              "set_autotune")

    def set_timings(self, on=True):
        check(lib().aran_fmm3d_set_timings(self._ctx, int(on)), "set_timings")

    def solve(self, x, y, z, q, lb, ub, order, leafcap, pot=None, grad=None,
              want_pot=True, want_grad=True):
        import torch
        n = x.numel()
        for t in (x, y, z, q):
            assert t.dtype == torch.float64 and t.is_cuda and t.is_contiguous()
            assert t.numel() == n
        if want_pot and pot is None:
            pot = torch.empty(n, dtype=torch.float64, device=x.device)
        if want_grad and grad is None:
            grad = torch.empty((n, 3), dtype=torch.float64, device=x.device)
        lbv = (C.c_double * 3)(*[float(v) for v in lb])
        ubv = (C.c_double * 3)(*[float(v) for v in ub])
        rc = lib().aran_fmm3d_solve(self._ctx, n, _ptr(x), _ptr(y), _ptr(z),
                                    _ptr(q), lbv, ubv, int(order),
                                    int(leafcap), _ptr(pot), _ptr(grad))
        check(rc, "aran_fmm3d_solve")
        self.n = n
        return pot, grad

    def solve_batch(self, offsets, x, y, z, q, lbs, ubs, order, leafcap,
                    pot=None, grad=None):
        """Independent systems in one solve (offsets: nsys+1 host ints)."""
        import torch
        n = x.numel()
        nsys = len(offsets) - 1
        assert offsets[0] == 0 and offsets[-1] == n
        if pot is None:
            pot = torch.empty(n, dtype=torch.float64, device=x.device)
        if grad is None:
            grad = torch.empty((n, 3), dtype=torch.float64, device=x.device)
        off = (C.c_long * (nsys + 1))(*[int(v) for v in offsets])
        lbv = (C.c_double * (3 * nsys))(*[float(v) for v in
                                          np.asarray(lbs).ravel()])
        ubv = (C.c_double * (3 * nsys))(*[float(v) for v in
                                          np.asarray(ubs).ravel()])
        rc = lib().aran_fmm3d_solve_batch(self._ctx, nsys, off, _ptr(x),
                                          _ptr(y), _ptr(z), _ptr(q), lbv, ubv,
                                          int(order), int(leafcap), _ptr(pot),
                                          _ptr(grad))
        check(rc, "aran_fmm3d_solve_batch")
        self.n = n
        return pot, grad

    def far_plan(self):
        """Far field of the last solve: tiles (0: the per-pair kernel ran),
        offset ranges per tile and workgroup schedule."""
        v = (C.c_int * 3)()
        check(lib().aran_fmm3d_get_far_plan(self._ctx, v), "get_far_plan")
        return {"tiles": v[0], "ranges": v[1], "schedule": v[2]}

    def stats(self):
        v = (C.c_long * len(STAT_NAMES))()
        check(lib().aran_fmm3d_get_stats(self._ctx, v), "get_stats")
        return dict(zip(STAT_NAMES, list(v)))

    def timings(self):
        """Phase times of the last solve in ms (tree, lists, up, m2l, down,
        near, total) and the kernel groups inside them (p2m, m2m, l2l,
        l2p)."""
        v = (C.c_float * 7)()
        check(lib().aran_fmm3d_get_timings(self._ctx, v), "get_timings")
        k = (C.c_float * 4)()
        check(lib().aran_fmm3d_get_kernel_timings(self._ctx, k),
              "get_kernel_timings")
        out = dict(zip(TIMER_NAMES, list(v)))
        out.update(zip(KERNEL_TIMER_NAMES, list(k)))
        return out

    def export_tree(self):
        st = self.stats()
        nn = st["nodes"]
        perm = np.zeros(self.n, dtype=np.int64)
        nodes = np.zeros((nn, 10), dtype=np.int32)
        centers = np.zeros((nn, 3))
        lo = np.zeros(3)
        width = np.zeros(1)
        check(lib().aran_fmm3d_export_tree(
            self._ctx, perm.ctypes.data, nodes.ctypes.data,
            centers.ctypes.data, lo.ctypes.data, width.ctypes.data),
            "export_tree")
        return perm, nodes, centers, lo, float(width[0])

    def export_lists(self):
        st = self.stats()
        nn, nl = st["nodes"], st["leaves"]
        nf, nu = st["far_entries"], st["near_entries"]
        far_ofs = np.zeros(nn + 1, dtype=np.int32)
        far_src = np.zeros(max(nf, 1), dtype=np.int32)
        far_code = np.zeros(max(nf, 1), dtype=np.int16)
        leaves = np.zeros(max(nl, 1), dtype=np.int32)
        near_ofs = np.zeros(nl + 1, dtype=np.int32)
        nb = np.zeros(max(nu, 1), dtype=np.int32)
        ne = np.zeros(max(nu, 1), dtype=np.int32)
        check(lib().aran_fmm3d_export_lists(
            self._ctx, far_ofs.ctypes.data, far_src.ctypes.data,
            far_code.ctypes.data, leaves.ctypes.data, near_ofs.ctypes.data,
            nb.ctypes.data, ne.ctypes.data), "export_lists")
        return dict(far_ofs=far_ofs, far_src=far_src[:nf],
                    far_code=far_code[:nf], leaves=leaves[:nl],
                    near_ofs=near_ofs, near_begin=nb[:nu], near_end=ne[:nu])

    def export_semifar(self):
        """Semifar lists of the last solve: per leaf (leaf order of
        export_lists) the nodes it evaluates by M2P, per node the leaves it
        takes by P2L."""
        st = self.stats()
        nn, nl, ns = st["nodes"], st["leaves"], st["semifar_pairs"]
        w_ofs = np.zeros(nl + 1, dtype=np.int32)
        w_node = np.zeros(max(ns, 1), dtype=np.int32)
        x_ofs = np.zeros(nn + 1, dtype=np.int32)
        x_leaf = np.zeros(max(ns, 1), dtype=np.int32)
        check(lib().aran_fmm3d_export_semifar(
            self._ctx, w_ofs.ctypes.data, w_node.ctypes.data,
            x_ofs.ctypes.data, x_leaf.ctypes.data), "export_semifar")
        return dict(w_ofs=w_ofs, w_node=w_node[:ns], x_ofs=x_ofs,
                    x_leaf=x_leaf[:ns])

    def direct(self, x, y, z, q, targets):
        import torch
        t = targets.to(device=x.device, dtype=torch.int64).contiguous()
        nt = t.numel()
        pot = torch.empty(nt, dtype=torch.float64, device=x.device)
        grad = torch.empty((nt, 3), dtype=torch.float64, device=x.device)
        check(lib().aran_fmm3d_direct(self._ctx, x.numel(), _ptr(x), _ptr(y),
                                      _ptr(z), _ptr(q), nt, _ptr(t),
                                      _ptr(pot), _ptr(grad)), "direct")
        return pot, grad
