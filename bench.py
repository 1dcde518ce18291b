"""Benchmark: 3D Newton FMM particles/s on one MI355X.

Workload (one "step" = one full solve: tree build, interaction lists,
upward pass, M2L, downward pass, P2P + L2P, results in input order):
Plummer sphere (scale radius 0.1, truncated at radius 1, seeded), N = 2^20
equal masses 1/N, expansion order p = 10, leaf capacity 64, potential and
gradient.  Inputs are generated on the host, copied to HBM once, and stay
resident; the timed region covers K solves.

Also reported: the M2L kernel's roofline point (algorithmic flops per launch
/ kernel duration from HIP events on the engine's stream), the whole solve's
roofline time against the measured one with every phase's achieved rate,
the error of the timed solve against an FP64 direct sum on 4096 sampled
targets, the plain FP64 form on the same workload, Config 1 on the GPU, the
512-system batch (with the plain form's error on it), serving sessions at
several offered loads, and the FP64 CPU oracle (one core, median of 3 after
a warm-up) as the CPU baseline.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MI355X_FP64_TFLOPS = 78.6
MI355X_FP32_TFLOPS = 157.3
MI355X_HBM_GBS = 8000.0
MI355X_BF16_TFLOPS = 2500.0  # dense MFMA


def rot_macs(L):
    return (L + 1) * (L + 2) * (4 * L + 3) // 6


def m2l_flops(p):
    """Algorithmic flops of one rotate / z-translate / rotate-back M2L with
    multipole degrees < p and local degrees <= p (complex MAC = 8 flops,
    complex x real = 4)."""
    v = sum(p - m for l in range(p + 1) for m in range(min(l, p - 1) + 1))
    return 8 * (rot_macs(p - 1) + rot_macs(p)) + 4 * v


# sources of the far-field kernels whose HBM traffic a stored profile gives
TRAFFIC_SOURCES = {
    "bf16": ["aran_amd/csrc/m2l_mfma.hip", "aran_amd/csrc/m2l_prep.hip",
             "aran_amd/csrc/m2l_mfma.h"],
    "f64": ["aran_amd/csrc/m2l_f64.hip", "aran_amd/csrc/m2l_f64.h"],
    "tile": ["aran_amd/csrc/m2l_tile.hip", "aran_amd/csrc/m2l_tile.h",
             "aran_amd/csrc/m2l_prep.hip"],
}


def source_digest(kind):
    """SHA-256 (first 16 hex digits) of the kernel sources of `kind`."""
    import hashlib
    h = hashlib.sha256()
    for rel in TRAFFIC_SOURCES[kind]:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(workload):
    """HBM bytes per launch of the dominant kernel for this workload.  The
    bytes come from two rocprofv3 counter runs (FETCH_SIZE, WRITE_SIZE) of
    this bench command, summarised under profiles/ with a TRAFFIC line that
    records the digest of the kernel sources it measured; they are used
    only while the sources still have that digest (a counter run cannot
    happen inside the timed run itself).  Returns (record, path) or
    (None, reason)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.md"))):
        with open(path) as fh:
            for line in fh:
                if line.startswith("TRAFFIC "):
                    rec = json.loads(line[len("TRAFFIC "):])
                    if rec.get("workload") != workload:
                        continue
                    kind = rec.get("sources", "")
                    if kind not in TRAFFIC_SOURCES or \
                            rec.get("source_digest") != source_digest(kind):
                        return None, ("stored profile %s measured other "
                                      "kernel sources" %
                                      os.path.relpath(path, ROOT))
                    return rec, os.path.relpath(path, ROOT)
    return None, "no stored counter profile for this workload"


def host_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_sample, seed):
    """The FP64 serial oracle (this project's restatement of the reference
    AranSolver3d path; the reference itself needs LibVsg, absent here) on
    one core of the host: per case one warm-up run, then the median of 3,
    timing the tree (oracle tree build) and the whole solve.  Cases: a
    bounded sample of the headline workload (Plummer, p = 10, leaf 64,
    n_sample particles) and Config 1 (N = 10,000 uniform cube, p = 8,
    leaf 20, potential + gradient)."""
    import oracle as O
    from aran_amd import workloads as W
    cpus = sorted(os.sched_getaffinity(0))
    os.sched_setaffinity(0, {cpus[0]})
    try:
        def timed(x, y, z, m, lb, ub, p, q):
            runs, trees = [], []
            for k in range(4):
                t0 = time.perf_counter()
                O.tree(x, y, z, lb, ub, q)
                t1 = time.perf_counter()
                O.fmm_solve(x, y, z, m, lb, ub, p, q)
                t2 = time.perf_counter()
                if k:  # the first run is the warm-up
                    trees.append(t1 - t0)
                    runs.append(t2 - t1)
            n = len(x)
            return {"value": n / float(np.median(runs)),
                    "runs_s": [round(r, 3) for r in runs],
                    "median_total_s": round(float(np.median(runs)), 3),
                    "median_tree_s": round(float(np.median(trees)), 3)}
        x, y, z, m = W.plummer(n_sample, seed)
        head = timed(x, y, z, m, -np.ones(3), np.ones(3), 10, 64)
        x, y, z, m = W.uniform_cube(10_000, 0xA2A0 + 1)
        c1 = timed(x, y, z, m, np.zeros(3), np.ones(3), 8, 20)
    finally:
        os.sched_setaffinity(0, set(cpus))
    return {"value": head["value"], "unit": "particles/s", "cores": 1,
            "kind": "port",
            "sample": "FP64 serial oracle FMM (restatement of the reference "
                      "AranSolver3d path; LibVsg absent) on a %d-particle "
                      "Plummer system (a = 0.1, r <= 1), p = 10, leaf 64, "
                      "potential + gradient; median of 3 after a warm-up, "
                      "pinned to one core" % n_sample,
            "host": host_model(), "nproc": os.cpu_count(),
            "runs_s": head["runs_s"], "median_tree_s": head["median_tree_s"],
            "config1": dict(c1, workload="N=10000 uniform cube, p=8, leaf "
                                         "20, potential+gradient")}


def run_form(f, dev, pot, grad, lb, ub, p, q, precision, steps, warmup):
    """Warm-up solves (the first one also runs the far-field split autotuner
    when it is on), then `steps` timed solves."""
    import torch
    n = dev[0].numel()
    f.set_precision(precision)
    f.set_timings(False)
    for _ in range(warmup):
        f.solve(*dev, lb, ub, p, q, pot, grad)
    f.set_timings(True)
    torch.cuda.synchronize()
    t_m2l, phases = [], []
    t0 = time.perf_counter()
    for _ in range(steps):
        f.solve(*dev, lb, ub, p, q, pot, grad)
        tm = f.timings()
        t_m2l.append(tm["m2l"])
        phases.append(tm)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    mean_phase = {k: round(float(np.mean([ph[k] for ph in phases])), 4)
                  for k in phases[0]}
    return {"value": n * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "m2l_ms": float(np.mean(t_m2l)), "phases_ms": mean_phase,
            "stats": f.stats(), "plan": f.far_plan()}


def host_inclusive(f, host, lb, ub, p, q, precision, steps):
    """The same solve with inputs and outputs in pinned host memory: each
    step copies x, y, z, q to the GPU, solves, copies potential and
    gradient back and waits for them (the PCIe-inclusive rate of a caller
    that hands over host buffers)."""
    import torch
    n = host[0].numel()
    hin = [t.pin_memory() for t in host]
    hpot = torch.empty(n, dtype=torch.float64).pin_memory()
    hgrad = torch.empty((n, 3), dtype=torch.float64).pin_memory()
    din = [torch.empty_like(t, device="cuda") for t in hin]
    pot = torch.empty(n, dtype=torch.float64, device="cuda")
    grad = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    f.set_precision(precision)
    f.set_timings(False)

    def step():
        for d, h in zip(din, hin):
            d.copy_(h, non_blocking=True)
        f.solve(*din, lb, ub, p, q, pot, grad)
        hpot.copy_(pot, non_blocking=True)
        hgrad.copy_(grad, non_blocking=True)
        torch.cuda.current_stream().synchronize()

    step()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    elapsed = time.perf_counter() - t0
    return {"value": n * steps / elapsed,
            "ms_per_step": 1e3 * elapsed / steps,
            "bytes_per_step": 8 * n * (4 + 4),
            "note": "pinned host inputs and outputs, copies and the wait for "
                    "the results inside each step"}


def struct_session(x, y, z, q, lb, ub, p, leafcap, precision, steps):
    """The reference's own interface end to end: particles in caller structs
    (VsgVector3d position, charge, potential, field; 64 bytes each) inserted
    by pointer into an AranSolver3d with the Newton callbacks and the
    declared layout.  Each step gathers through the layout into pinned host
    memory (host threads), copies in, solves, copies out and adds the
    results into the structs."""
    import ctypes as C

    from aran_amd import solver as S
    L = S.lib()
    n = len(x)
    dt = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("density", "f8"),
                   ("pot", "f8"), ("fx", "f8"), ("fy", "f8"), ("fz", "f8")])
    pts = np.zeros(n, dtype=dt)
    pts["x"], pts["y"], pts["z"], pts["density"] = x, y, z, q
    vlb, vub = S.VsgVector3d(*lb), S.VsgVector3d(*ub)
    tree = L.vsg_prtree3d_new_full(C.byref(vlb), C.byref(vub),
                                   S.fn("vsg_vector3d_vector3d_locfunc"),
                                   S.fn("vsg_vector3d_dist"), None, leafcap)
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
    code = {"plain": 0, "mixed": 1, "bf16x3": 2, "bf16x6": 3}[precision]
    if L.aran_solver3d_set_precision(solver, code) != 0:
        raise RuntimeError("aran_solver3d_set_precision failed")
    base = pts.ctypes.data
    for i in range(n):
        L.aran_solver3d_insert_point(solver, base + i * dt.itemsize)
    L.aran_solver3d_solve(solver)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        L.aran_solver3d_solve(solver)
    elapsed = time.perf_counter() - t0
    ok = L.aran_solver3d_status(solver) == 0
    L.aran_solver3d_free(solver)
    if not ok:
        raise RuntimeError("aran_solver3d_solve failed")
    return {"value": n * steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
            "form": precision,
            "note": "AranSolver3d with particles in 64-byte caller structs: "
                    "gather through the layout into pinned memory, copies, "
                    "solve, copies, results added into the structs"}


def error_vs_direct(f, dev, pot, grad, n, seed):
    import torch
    from aran_amd import workloads as W
    tg = W.sample_targets(n, 4096, seed)
    dpot, dgrad = f.direct(*dev, torch.from_numpy(tg))
    dpot, dgrad = dpot.cpu().numpy(), dgrad.cpu().numpy()
    pp, gg = pot.cpu().numpy()[tg], grad.cpu().numpy()[tg]
    return {"pot_l2": float(np.linalg.norm(pp - dpot) / np.linalg.norm(dpot)),
            "pot_max": float(np.max(np.abs(pp - dpot) / np.abs(dpot))),
            "grad_l2": float(np.linalg.norm(gg - dgrad) / np.linalg.norm(dgrad)),
            "targets": 4096}


def batch_session(nsys, order, leafcap, steps, warmup, seed,
                  precision="mixed"):
    """One batched solve of nsys independent systems (shared launches, one
    octree per system), timed over `steps` solves; error against the FP64
    direct sum of each target's own system on 4096 sampled targets, beside
    the error of one plain FP64 solve of the same batch on the same targets
    (the truncation floor of the order and leaf capacity)."""
    import torch

    from aran_amd import workloads as W
    from aran_amd.fmm import Fmm3d
    cat, offs, lbs, ubs = W.batch_systems(nsys, seed)
    dev = [torch.from_numpy(a).cuda() for a in cat]
    n = offs[-1]
    f = Fmm3d()
    f.set_precision(precision)
    pot = torch.empty(n, dtype=torch.float64, device="cuda")
    grad = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    for _ in range(warmup):
        f.solve_batch(offs, *dev, lbs, ubs, order, leafcap, pot, grad)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.solve_batch(offs, *dev, lbs, ubs, order, leafcap, pot, grad)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    tg = W.sample_targets(n, 4096, seed)
    sysid = np.searchsorted(np.array(offs), tg, side="right") - 1
    pp, gg = pot.cpu().numpy()[tg], grad.cpu().numpy()[tg]
    dp, dg = np.zeros(len(tg)), np.zeros((len(tg), 3))
    for s in np.unique(sysid):
        sel = np.nonzero(sysid == s)[0]
        a, b = offs[s], offs[s + 1]
        loc = torch.from_numpy(tg[sel] - a)
        p_, g_ = f.direct(*[t[a:b] for t in dev], loc)
        dp[sel], dg[sel] = p_.cpu().numpy(), g_.cpu().numpy()
    st = f.stats()

    def errs(pp, gg):
        return {"pot_l2": float(np.linalg.norm(pp - dp) / np.linalg.norm(dp)),
                "pot_max": float(np.max(np.abs(pp - dp) / np.abs(dp))),
                "grad_l2": float(np.linalg.norm(gg - dg) /
                                 np.linalg.norm(dg))}
    out = {"systems": nsys, "particles": int(n), "order": order,
           "leafcap": leafcap, "form": precision,
           "ms_per_solve": 1e3 * dt, "particles_s": n / dt,
           **errs(pp, gg), "nodes": st["nodes"], "m2l": st["m2l"]}
    if precision != "plain":
        f.set_precision("plain")
        f.solve_batch(offs, *dev, lbs, ubs, order, leafcap, pot, grad)
        out["plain_fp64"] = errs(pot.cpu().numpy()[tg], grad.cpu().numpy()[tg])
    return out


def precision_sweep(n, orders, steps, seed):
    """Uniform cube, charges in [-1, 1), leaf 64: per order and form,
    particles/s and L2 / max errors of potential and gradient against the
    FP64 direct sum on 4096 sampled targets."""
    import torch

    from aran_amd import workloads as W
    from aran_amd.fmm import Fmm3d
    x, y, z, qv = W.signed_cube(n, seed)
    dev = [torch.from_numpy(a).cuda() for a in (x, y, z, qv)]
    f = Fmm3d()
    pot = torch.empty(n, dtype=torch.float64, device="cuda")
    grad = torch.empty((n, 3), dtype=torch.float64, device="cuda")
    tg = W.sample_targets(n, 4096, seed)
    dp, dg = f.direct(*dev, torch.from_numpy(tg))
    dp, dg = dp.cpu().numpy(), dg.cpu().numpy()
    out = []
    dtypes = {"plain": "f64", "mixed": "f32 M2L/P2P",
              "bf16x3": "bf16x3 M2L (MFMA), f32 P2P",
              "bf16x6": "bf16x6 M2L (MFMA), f32 P2P"}
    for p in orders:
        for form in ("plain", "mixed", "bf16x3", "bf16x6"):
            r = run_form(f, dev, pot, grad, (0, 0, 0), (1, 1, 1), p, 64, form,
                         steps, 1)
            pp, gg = pot.cpu().numpy()[tg], grad.cpu().numpy()[tg]
            out.append({"order": p, "form": form,
                        "dtype": dtypes[form],
                        "particles_s": r["value"],
                        "ms_per_solve": r["ms_per_step"],
                        "m2l_ms": r["m2l_ms"],
                        "pot_l2": float(np.linalg.norm(pp - dp) /
                                        np.linalg.norm(dp)),
                        "pot_max": float(np.max(np.abs(pp - dp) /
                                                np.abs(dp))),
                        "grad_l2": float(np.linalg.norm(gg - dg) /
                                         np.linalg.norm(dg))})
    return out


def serve_session(n_requests, rates, seed, precision="mixed"):
    """Poisson arrivals at each offered load in `rates` (requests/s); per
    request N log-uniform in [1e3, 2e6], p in {6, 8, 12}, leaf 64, half
    potential-only, half potential + gradient, uniform cube or Plummer
    (first N points of seeded 2e6-point pools, so each request is an i.i.d.
    sample), masses in (0, 1].  Admission under a 200 GB budget, at most 32
    requests per batch; the server is prepared for 2e6 particles and orders
    6, 8, 12 before the first arrival.  Per load: latency percentiles,
    throughput, batch counts, and every request's potential error against
    the FP64 direct sum on 64 sampled targets (L2 over the targets; the
    worst and the median request)."""
    import threading

    import torch

    from aran_amd import workloads as W
    from aran_amd.fmm import Fmm3d
    from aran_amd.server import Server
    nmax = 2_000_000
    reqs, gaps0 = W.serving_requests(n_requests, seed, nmax)
    # FP64 direct sums on 64 sampled targets of every request (the same
    # requests at every load)
    chk = Fmm3d()
    refs = []
    for k, r in enumerate(reqs):
        n = len(r[0])
        tg = W.sample_targets(n, min(64, n), seed + k)
        dev = [torch.from_numpy(np.ascontiguousarray(a)).cuda()
               for a in r[:4]]
        dp, _ = chk.direct(*dev, torch.from_numpy(tg))
        refs.append((tg, dp.cpu().numpy()))
        del dev
    out = []
    for rate in rates:
        gaps = gaps0 / rate
        srv = Server(budget_bytes=200e9, nworkers=2, precision=precision,
                     max_batch=32)
        # staging, workspaces and tables for the largest request and every
        # order, before the arrivals start
        srv.prepare(nmax, (6, 8, 12), 64)
        ids = [None] * n_requests
        done = {}

        def waiter():
            for k in range(n_requests):
                while ids[k] is None:
                    time.sleep(1e-4)
                done[k] = srv.wait(ids[k])
        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        t_next = t0
        for k, r in enumerate(reqs):
            t_next += gaps[k]
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids[k] = srv.submit(*r[:4], r[4], r[5], r[6], r[7],
                                want_grad=r[8])
        th.join()
        wall = time.perf_counter() - t0
        st = srv.stats()
        srv.close()
        lat = np.array([done[k][2]["latency"] for k in range(n_requests)])
        svc = np.array([done[k][2]["service"] for k in range(n_requests)])
        errs = []
        for k, (tg, dp) in enumerate(refs):
            pp = np.asarray(done[k][0])[tg]
            errs.append(float(np.linalg.norm(pp - dp) / np.linalg.norm(dp)))
        parts = sum(len(r[0]) for r in reqs)
        out.append({"offered_load_req_s": rate, "requests": n_requests,
                    "p50_ms": float(1e3 * np.percentile(lat, 50)),
                    "p99_ms": float(1e3 * np.percentile(lat, 99)),
                    "mean_ms": float(1e3 * lat.mean()),
                    "p99_service_ms": float(1e3 * np.percentile(svc, 99)),
                    "throughput_req_s": n_requests / wall,
                    "throughput_particles_s": parts / wall,
                    "batches": st["batches"], "max_batch": st["max_batch"],
                    "pot_l2_worst": max(errs),
                    "pot_l2_median": float(np.median(errs))})
    return {"form": precision, "workers": 2, "budget_bytes": 200e9,
            "max_batch_limit": 32, "loads": out}


def algorithmic_work(n, p, st):
    """Per-phase algorithmic work of one solve (the reference algorithm's
    operation counts; complex MAC = 8 flops, complex x real = 4): flops of
    P2P (20 per ordered interaction, potential + gradient), P2M, M2M, M2L,
    L2L, L2P, and the tree build's HBM bytes."""
    tri = lambda k: k * (k + 1) // 2
    R = rot_macs
    m2m = 8 * 2 * R(p - 1) + 4 * sum(l - m + 1 for l in range(p)
                                     for m in range(l + 1))
    l2l = 8 * 2 * R(p) + 4 * sum(p - l + 1 for l in range(p + 1)
                                 for m in range(l + 1))
    depth = st["depth"]
    return {"tree_bytes": n * (32 + 12 + 24 * -(-3 * depth // 8) + 64),
            "p2p": 20 * 2 * st["p2p"],
            "p2m": n * (6 * tri(p + 1) + 6 * tri(p)),
            "m2m": st["m2m"] * m2m,
            "m2l": st["m2l"] * m2l_flops(p),
            "l2l": st["l2l"] * l2l,
            "l2p": n * (6 + 4 + 20) * tri(p + 1)}


def whole_solve_roofline(form, n, p, st, ph, t_measured_ms):
    """t_roof = B_tree / HBM + F_p2p / P_p2p + (F_p2m + F_l2p) / P_fp64
    + F_m2l / P_m2l + (F_m2m + F_l2l) / P_fp64, against the measured time of
    a solve; and the achieved rate of each phase from the engine's HIP-event
    timers on the stream each runs on (P2P on the second stream overlaps
    the far-field chain, so the phases do not add up to the total)."""
    w = algorithmic_work(n, p, st)
    f64, f32 = MI355X_FP64_TFLOPS * 1e12, MI355X_FP32_TFLOPS * 1e12
    p_p2p = f64 if form == "plain" else f32
    p_m2l = {"plain": f64, "mixed": f32,
             "bf16x3": MI355X_BF16_TFLOPS * 1e12 / 3,
             "bf16x6": MI355X_BF16_TFLOPS * 1e12 / 6}[form]
    hbm = MI355X_HBM_GBS * 1e9
    t_roof = (w["tree_bytes"] / hbm + w["p2p"] / p_p2p +
              (w["p2m"] + w["l2p"]) / f64 + w["m2l"] / p_m2l +
              (w["m2m"] + w["l2l"]) / f64)

    def rate(work, ms, peak):
        if ms <= 0:
            return None
        v = work / (ms * 1e-3)
        return {"ms": round(ms, 4), "rate": round(v / 1e12, 4),
                "unit": "TFLOP/s", "frac": round(v / peak, 5)}
    tree = {"ms": round(ph["tree"], 4),
            "rate": round(w["tree_bytes"] / (ph["tree"] * 1e-3) / 1e9, 2),
            "unit": "GB/s", "frac": round(w["tree_bytes"] /
                                          (ph["tree"] * 1e-3) / hbm, 5),
            "bytes": w["tree_bytes"]}
    return {"t_roof_ms": round(1e3 * t_roof, 4),
            "t_measured_ms": round(t_measured_ms, 4),
            "frac": round(1e3 * t_roof / t_measured_ms, 5),
            "peaks_tflops": {"p2p": p_p2p / 1e12, "p2m_l2p_m2m_l2l": 78.6,
                             "m2l": round(p_m2l / 1e12, 2),
                             "hbm_gbs": MI355X_HBM_GBS},
            "phases": {
                "tree": tree,
                "lists": {"ms": round(ph["lists"], 4)},
                "p2p": rate(w["p2p"], ph["near"], p_p2p),
                "p2m": rate(w["p2m"], ph["p2m"], f64),
                "m2m": rate(w["m2m"], ph["m2m"], f64),
                "m2l": rate(w["m2l"], ph["m2l"], p_m2l),
                "l2l": rate(w["l2l"], ph["l2l"], f64),
                "l2p": rate(w["l2p"], ph["l2p"], f64)},
            "flops": {k: v for k, v in w.items() if k != "tree_bytes"}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--order", type=int, default=10)
    ap.add_argument("--leafcap", type=int, default=64)
    ap.add_argument("--precision",
                    choices=("mixed", "plain", "bf16x3", "bf16x6"),
                    default="bf16x3",
                    help="form of the timed solves: bf16x3 (M2L on MFMA, "
                         "BF16 operands in 3 products, FP32 P2P; the "
                         "default), bf16x6 (6 products), mixed (FP32 VALU "
                         "M2L), plain (FP64)")
    ap.add_argument("--cpu-sample", type=int, default=32768,
                    help="particles of the CPU baseline's headline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip the plain FP64 comparison run")
    ap.add_argument("--no-structs", action="store_true",
                    help="skip the end-to-end run through the AranSolver3d "
                         "pointer interface (particles in caller structs)")
    ap.add_argument("--autotune", action="store_true",
                    help="far-field split from the on-device autotuner (its "
                         "candidate launches run in the warm-up) instead of "
                         "the cost model")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one elementwise kernel that reads and "
                         "writes 512 MiB (the reference point for HBM byte "
                         "counters)")
    ap.add_argument("--no-config1", action="store_true",
                    help="skip the Config-1 GPU line (profiler runs of the "
                         "headline workload alone)")
    ap.add_argument("--batch-systems", type=int, default=512,
                    help="independent systems of the batched solve (0: skip)")
    ap.add_argument("--sweep", action="store_true",
                    help="only the precision sweep: N=4,194,304 uniform "
                         "cube, signed charges, leaf 64, p in {6, 10, 14}, "
                         "every form (one JSON line)")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--serve-rates", default="100,400,1000,2000,4000",
                    help="offered loads of the serving sessions, requests/s "
                         "(comma-separated; the last ones saturate)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; --gpus %d is not "
              "supported (one-GPU measurements only)" % args.gpus)
        sys.exit(2)

    import torch

    from aran_amd import workloads as W
    from aran_amd.fmm import Fmm3d

    torch.cuda.set_device(0)
    if args.sweep:
        pts = precision_sweep(4_194_304, (6, 10, 14), max(2, args.steps // 3),
                              0xA2A0 + 3)
        print(json.dumps({"sweep": "uniform cube N=4194304, charges in "
                                   "[-1,1), leaf 64, pot+grad",
                          "points": pts}))
        return
    seed = 0xA2A0 + 2
    n, p, q = args.n, args.order, args.leafcap
    x, y, z, m = W.plummer(n, seed)
    dev = [torch.from_numpy(a).cuda() for a in (x, y, z, m)]
    lb, ub = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    f = Fmm3d()
    f.set_autotune(args.autotune)
    pot = torch.empty(n, dtype=torch.float64, device="cuda")
    grad = torch.empty((n, 3), dtype=torch.float64, device="cuda")

    head = run_form(f, dev, pot, grad, lb, ub, p, q, args.precision,
                    args.steps, args.warmup)
    err = None if args.no_check else error_vs_direct(f, dev, pot, grad, n,
                                                     seed)
    plain = None
    if args.precision != "plain" and not args.no_plain:
        r = run_form(f, dev, pot, grad, lb, ub, p, q, "plain",
                     max(2, args.steps // 2), 1)
        m2l_tf = r["stats"]["m2l"] * m2l_flops(p) / (r["m2l_ms"] * 1e-3) / 1e12
        plain = {"value": r["value"], "ms_per_step": r["ms_per_step"],
                 "m2l_ms": round(r["m2l_ms"], 4),
                 "m2l_kernel": "k_m2l_f64 (FP64 MFMA)"
                 if r["plan"]["tiles"] > 0 else "k_m2l (FP64)",
                 "m2l_tflops": round(m2l_tf, 3),
                 "m2l_frac_fp64_peak": round(m2l_tf / MI355X_FP64_TFLOPS, 4),
                 "phases_ms": r["phases_ms"],
                 "whole_solve": whole_solve_roofline(
                     "plain", n, p, r["stats"], r["phases_ms"],
                     r["ms_per_step"])}
        if not args.no_check:
            plain["error"] = error_vs_direct(f, dev, pot, grad, n, seed)
        plain["mixed_over_plain"] = head["value"] / r["value"]

    # Config 1 on the GPU (plain FP64, N = 10,000 uniform cube, p = 8, leaf
    # 20), beside its CPU line
    config1 = None
    if not args.no_config1:
        x1, y1, z1, m1 = W.uniform_cube(10_000, 0xA2A0 + 1)
        d1 = [torch.from_numpy(a).cuda() for a in (x1, y1, z1, m1)]
        p1 = torch.empty(10_000, dtype=torch.float64, device="cuda")
        g1 = torch.empty((10_000, 3), dtype=torch.float64, device="cuda")
        r1 = run_form(f, d1, p1, g1, (0, 0, 0), (1, 1, 1), 8, 20, "plain",
                      max(5, args.steps), 2)
        config1 = {"workload": "N=10000 uniform cube, p=8, leaf 20, "
                               "potential+gradient, plain FP64",
                   "value": r1["value"], "ms_per_step": r1["ms_per_step"],
                   "error": None if args.no_check else
                   error_vs_direct(f, d1, p1, g1, 10_000, 0xA2A0 + 1)}

    hostinc = host_inclusive(f, [torch.from_numpy(a) for a in (x, y, z, m)],
                             lb, ub, p, q, args.precision,
                             max(2, args.steps // 2))
    structs = None if args.no_structs else struct_session(
        x, y, z, m, lb, ub, p, q, args.precision, max(2, args.steps // 2))

    st = head["stats"]
    form = args.precision
    plan = head["plan"]
    dense = plan["tiles"] > 0
    # peak of the arithmetic the M2L kernel issues: FP64 MFMA (plain, dense
    # tiles) or FP64 VALU (per-pair kernel), FP32 VALU (mixed), dense BF16
    # MFMA (bf16x3 / bf16x6; the split products are not counted as
    # algorithmic work, so frac is against the BF16 peak)
    peak = {"plain": MI355X_FP64_TFLOPS, "mixed": MI355X_FP32_TFLOPS,
            "bf16x3": MI355X_BF16_TFLOPS, "bf16x6": MI355X_BF16_TFLOPS}[form]
    kname = {"plain": "k_m2l_f64 (FP64 MFMA)" if dense else "k_m2l (FP64)",
             "mixed": "k_m2l_tile (FP32)",
             "bf16x3": "k_m2l_mfma1w (BF16 MFMA, 3 products, one wave "
                       "per SIMD)" if not os.environ.get(
                           "ARAN_MFMA_KERNEL", "").startswith("8") else
                       "k_m2l_mfma (BF16 MFMA, 3 products)",
             "bf16x6": "k_m2l_mfma (BF16 MFMA, 6 products)"}[form]
    flops = st["m2l"] * m2l_flops(p)
    achieved = flops / (head["m2l_ms"] * 1e-3) / 1e12
    workload = "plummer_2p20_p10_q64_%s" % args.precision
    tr, tr_src = measured_traffic(workload) \
        if (n, p, q) == (1 << 20, 10, 64) else (None, "other workload")
    roof = {"bound": "mfma" if (dense or form.startswith("bf16")) else
            "valu",
            "kernel": kname,
            "achieved": round(achieved, 4), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 5),
            "traffic": tr["bytes_per_launch"] if tr else None,
            "traffic_unit": "HBM bytes per launch (rocprofv3 FETCH_SIZE x "
                            "%.1f + WRITE_SIZE x %.1f)" % (
                                tr["fetch_factor"], tr["write_factor"])
            if tr else None,
            "traffic_source": ("%s: counter runs of this command on the "
                               "same kernel sources (digest %s), not of this "
                               "run" % (tr_src, tr["source_digest"]))
            if tr else tr_src,
            "algorithmic": "%d M2L x %d flops" % (st["m2l"], m2l_flops(p)),
            "kernel_ms": round(head["m2l_ms"], 4),
            "far_plan": plan}
    whole = whole_solve_roofline(form, n, p, st, head["phases_ms"],
                                 head["ms_per_step"])

    if args.calibrate_copy:
        # one elementwise kernel that reads 512 MiB and writes 512 MiB
        a = torch.ones(1 << 26, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        torch.add(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
    batch = None
    if args.batch_systems > 0:
        batch = batch_session(args.batch_systems, 8, 32, max(2, args.steps // 2),
                              1, 0xA2A0 + 4, args.precision)
    serve = None
    if args.serve_requests > 0:
        serve = serve_session(args.serve_requests,
                              [float(v) for v in args.serve_rates.split(",")],
                              seed + 3, args.precision)
    cpu = None if args.no_cpu else cpu_baseline(args.cpu_sample, seed)
    line = {
        "metric": "3D Newton FMM particles/s at N=2^20 (Plummer, p=10, "
                  "leaf 64), rel err vs FP64 direct sum",
        "value": head["value"], "unit": "particles/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"plain": "f64", "mixed": "f32", "bf16x3": "bf16x3",
                  "bf16x6": "bf16x6"}[form],
        "data": "synthetic (seeded Plummer sphere, a=0.1, r<=1)",
        "config": {"workload": workload,
                   "n": n, "order": p, "leafcap": q,
                   "form": {
                       "plain": "plain FP64",
                       "mixed": "mixed: FP32 M2L and P2P on per-level "
                                "scaled expansions, FP64 elsewhere",
                       "bf16x3": "M2L on MFMA, BF16 operands in 3 products "
                                 "with FP32 accumulation; FP32 P2P; FP64 "
                                 "elsewhere",
                       "bf16x6": "M2L on MFMA, BF16 operands in 6 products "
                                 "with FP32 accumulation; FP32 P2P; FP64 "
                                 "elsewhere"}[form],
                   "outputs": "potential+gradient"},
        "roofline": roof,
        "whole_solve": whole,
        "cpu_baseline": cpu,
        "error": err,
        "phases_ms": head["phases_ms"],
        "plain_fp64": plain,
        "config1_gpu": config1,
        "host_buffers": hostinc,
        "caller_structs": structs,
        "batched": batch,
        "serving": serve,
        "tree": {k: st[k] for k in ("depth", "nodes", "leaves", "m2l", "p2p")},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
