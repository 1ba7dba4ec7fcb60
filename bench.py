#!/usr/bin/env python3
"""Benchmark of the CT projector pair and the beam-hardening gradient on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one forward projection plus one adjoint (backprojection) of the 1024x1024
fan-beam slice (360 views over 360 degrees, 1024 bins, source distance 4096 px), inputs
resident in HBM.  The printed JSON line's `value` is voxel-ray updates per second of that
pair in units of 1e9 (GUPS): 2 * nnz(A) * K / t.  Extra fields:
  * bhc: polychromatic Poisson gradient evaluations per second (512x512 fan slice, 360
    views x 729 bins, d = 2000, E = 20 spline bases; one evaluation = forward
    projection, FP64 NLL + per-ray gradient weights, adjoint);
  * batched: 64 ragged slices (mixed parallel/fan) in one launch per operator vs one
    launch per slice, fwd+adj GUPS of both;
  * roofline: the dominant kernel's algorithmic bytes (4 B per voxel-ray update) per
    launch over its average launch time (HIP events on the launch stream), against
    8 TB/s; traffic from the committed rocprofv3 counter summary when present;
  * cpu_baseline: the reference's own pthreads cpuPrj (prj/cpuPrj.c, built into
    oracle/_ref) on the same slice, timed on this host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

from imgrecsrc_amd.workloads import FAN1024, FAN512, PAR256, STACK512, ragged_geoms  # noqa: E402
WORKLOADS = {"fan1024": FAN1024, "par256": PAR256, "fan512": FAN512, "par512": STACK512}


# Autotuner cache of every plan the bench creates (set from --tune-cache).
DEFAULT_TUNE_CACHE = os.path.join(ROOT, "imgrecsrc_amd", "tunes", "mi355x.tune")
TUNE_CACHE = None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def timed(fn, reps, stream):
    """Average ms per call of fn() with HIP events on `stream` (the launch stream)."""
    import torch
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    e1.synchronize()
    return e0.elapsed_time(e1) / reps


def graph_timed(fn, reps, stream):
    """fn() captured once as a HIP graph (torch.cuda.CUDAGraph) and replayed: average ms
    per replay on `stream`, without host launch overhead; also the captured call's
    outputs (valid after each replay)."""
    import torch
    fn()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    with torch.cuda.graph(g, stream=cap):
        out = fn()
    torch.cuda.synchronize()
    g.replay()
    torch.cuda.synchronize()
    t = timed(g.replay, reps, stream)
    torch.cuda.synchronize()
    return t, out, g


def traffic_from_profiles(kernel_tag):
    """HBM bytes per launch from profiles/<tag>_pmc.json (FETCH_SIZE and WRITE_SIZE of
    two rocprofv3 --pmc runs, KiB per dispatch, each divided by the counter's reading
    per KiB of a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", f"{kernel_tag}_pmc.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        p = json.load(f)
    try:
        return float(p["traffic_bytes"])
    except (KeyError, ZeroDivisionError):
        return None


def _csv_rows(path):
    import csv
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def _col(row, *names):
    low = {k.lower(): v for k, v in row.items()}
    for n in names:
        if n.lower() in low:
            return low[n.lower()]
    raise KeyError(names)


def pmc_summary(fetch_dir, write_dir, kernel_substr, tag, command):
    """Reduce two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE; -f csv) to
    profiles/<tag>_pmc.json: per-call KiB of the kernels of one operator call, named
    '+'-separated (each kernel's trailing dispatches at its last grid size: the timed
    workload), and the same counter for the largest plain device copy (the --calib-copy
    256 MiB copy) to scale by."""
    import glob
    import statistics
    parts = [k for k in kernel_substr.split("+") if k]
    out = {"command": command, "kernel": kernel_substr}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if not files:
            raise SystemExit(f"no counter_collection.csv under {d}")
        rows = [r for f in files for r in _csv_rows(f)
                if _col(r, "Counter_Name") == counter]
        total, per_kernel = 0.0, {}
        for part in parts:
            # truncated (-T) names are the bare function name; full ones carry <...>(...)
            kern = [r for r in rows
                    if _col(r, "Kernel_Name").split("<")[0].split("(")[0].split("::")[-1] == part]
            if not kern:
                continue
            # the trailing dispatches at the last grid size: the timed shape, after the
            # autotuner's trial shapes
            kern.sort(key=lambda r: int(float(_col(r, "Dispatch_Id"))))
            g = int(float(_col(kern[-1], "Grid_Size")))
            vals = []
            for r in reversed(kern):
                if int(float(_col(r, "Grid_Size"))) != g:
                    break
                vals.append(float(_col(r, "Counter_Value")))
            per_kernel[part] = {"KiB": statistics.mean(vals), "dispatches": len(vals), "grid": g}
            total += statistics.mean(vals)
        if not per_kernel:
            raise SystemExit(f"{counter}: no dispatch of {kernel_substr}")
        copies = [r for r in rows if "copyBuffer" in _col(r, "Kernel_Name")]
        big = max(copies, key=lambda r: int(float(_col(r, "Grid_Size"))))
        bg = int(float(_col(big, "Grid_Size")))
        cv = [float(_col(r, "Counter_Value")) for r in copies
              if int(float(_col(r, "Grid_Size"))) == bg]
        copy_kib = 256.0 * 1024.0  # 64 Mi floats read and written
        out[f"{counter}_KiB"] = total
        out[f"{counter}_per_kernel"] = per_kernel
        out[f"copy_{'fetch' if counter == 'FETCH_SIZE' else 'write'}_factor"] = \
            statistics.mean(cv) / copy_kib
    out["traffic_bytes"] = (out["FETCH_SIZE_KiB"] / out["copy_fetch_factor"] +
                            out["WRITE_SIZE_KiB"] / out["copy_write_factor"]) * 1024.0
    path = os.path.join(ROOT, "profiles", f"{tag}_pmc.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


CLOCK_GHZ = 2.4             # MI355X peak engine clock (GRBM_GUI_ACTIVE / kernel time agrees)
CUS = 256
VALU_LANES_PER_CU_CLK = 128  # 4 SIMDs; two VALU instructions can issue per quad-cycle
L1_BYTES_PER_CU_CLK = 64     # vector memory (TA/TD) data path per CU


def valu_peak_lane_ops():
    """VALU lane-ops per second of the whole GPU: 256 CUs x 128 lanes x 2.4 GHz, the
    FP32 vector FMA rate of the data sheet (157.3 TFLOP/s / 2)."""
    return CUS * VALU_LANES_PER_CU_CLK * CLOCK_GHZ * 1e9


def l1_peak_bytes():
    """Vector-memory (TA/TD) data rate of the whole GPU: 64 B/clk/CU (a 64-lane 16-byte
    buffer load occupies the TA for 16 cycles)."""
    return CUS * L1_BYTES_PER_CU_CLK * CLOCK_GHZ * 1e9


def valu_summary(sq_dir, kernel_substr, tag, command, tune_json=None, suffix="valu"):
    """Reduce a rocprofv3 --pmc run of SQ counters (-f csv) to profiles/<tag>_valu.json:
    per kernel of one operator call ('+'-separated names, each at its trailing dispatches
    of the last grid size: the timed shape) the mean per dispatch of every counter
    collected, and the call's sums.  lane_ops = SQ_INSTS_VALU * 64 (wave64 issue slots)."""
    import glob
    import statistics
    files = glob.glob(os.path.join(sq_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {sq_dir}")
    rows = [r for f in files for r in _csv_rows(f)]
    parts = [k for k in kernel_substr.split("+") if k]
    out = {"command": command, "kernel": kernel_substr, "per_kernel": {}}
    if tune_json:
        out["tune"] = json.loads(tune_json)
    totals = {}
    for part in parts:
        kern = [r for r in rows
                if _col(r, "Kernel_Name").split("<")[0].split("(")[0].split("::")[-1] == part]
        if not kern:
            continue
        kern.sort(key=lambda r: int(float(_col(r, "Dispatch_Id"))))
        g = int(float(_col(kern[-1], "Grid_Size")))
        by = {}
        for r in reversed(kern):
            if int(float(_col(r, "Grid_Size"))) != g:
                break
            by.setdefault(_col(r, "Counter_Name"), []).append(float(_col(r, "Counter_Value")))
        means = {c: statistics.mean(v) for c, v in by.items()}
        means["dispatches"] = max(len(v) for v in by.values())
        means["grid"] = g
        out["per_kernel"][part] = means
        for c, v in means.items():
            if c.startswith(("SQ_", "GRBM_", "TA_", "TD_", "TCP_")):
                totals[c] = totals.get(c, 0.0) + v
    if not out["per_kernel"]:
        raise SystemExit(f"no dispatch of {kernel_substr}")
    out["call"] = totals
    if "SQ_INSTS_VALU" in totals:
        out["lane_ops_per_call"] = totals["SQ_INSTS_VALU"] * 64.0
    path = os.path.join(ROOT, "profiles", f"{tag}_{suffix}.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


VALU_MIN_OPS_PER_UPDATE = 3.75  # one ~22-op weight shared by 8 images, plus one FMA each


def _summary(tag, suffix):
    path = os.path.join(ROOT, "profiles", f"{tag}_{suffix}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def roofline_entry(tag, kernel, kernel_ms, launches, nnz, tune, other_ms):
    """The dominant call against the ceiling that binds it.  The symmetric kernels reuse
    every loaded value for 2-8 updates and move little HBM traffic; their counters
    (committed under profiles/, taken at the same tuned shapes) put the vector-memory
    pipeline (TA address and TD data units, 64 B/clk/CU) or the VALU highest.  Vector
    memory: busy cycles per call (TA_TA_BUSY or TD_TD_BUSY summed over the 256 units,
    profiles/<tag>_l1.json) / (256 x the call's live duration x 2.4 GHz), the larger of
    the two; achieved = that fraction x 39.3 TB/s, the data rate the busy time carries.
    VALU: SQ_INSTS_VALU x 64 lane-ops per call over the live duration against
    78.6 T lane-op/s.  `bound` is the busier of the two, `frac` its fraction; the other
    and HBM (measured FETCH_SIZE + WRITE_SIZE over the same duration) are reported beside
    it.  Without summaries at these shapes: HBM from the measured traffic, or nothing."""
    traffic = traffic_from_profiles(tag)
    sec = kernel_ms * 1e-3
    hbm = None
    if traffic is not None:
        gbs = traffic / sec / 1e9
        hbm = {"achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": round(gbs / HBM_PEAK_GBS, 4), "bytes_per_call": round(traffic)}
    base = {"kernel": kernel, "kernel_ms": round(kernel_ms, 4), "launches": launches,
            "other_kernel_ms": round(other_ms, 4), "traffic": traffic,
            "updates_per_s": round(nnz / sec, 1)}
    valu, l1 = _summary(tag, "valu"), _summary(tag, "l1")
    if (valu is not None and l1 is not None and valu.get("tune") == tune
            and l1.get("tune") == tune):
        lane_ops = float(valu["lane_ops_per_call"])
        vfrac = lane_ops / sec / valu_peak_lane_ops()
        cyc = CUS * sec * CLOCK_GHZ * 1e9
        ta = float(l1["call"]["TA_TA_BUSY"])
        td = float(l1["call"].get("TD_TD_BUSY", 0.0))
        mfrac = max(ta, td) / cyc
        vmem = {"achieved": round(mfrac * l1_peak_bytes() / 1e12, 2),
                "peak": round(l1_peak_bytes() / 1e12, 2), "unit": "TB/s",
                "frac": round(mfrac, 4), "ta_frac": round(ta / cyc, 4),
                "td_frac": round(td / cyc, 4), "ta_busy_cycles_per_call": ta,
                "td_busy_cycles_per_call": td}
        vl = {"achieved": round(lane_ops / sec / 1e12, 2),
              "peak": round(valu_peak_lane_ops() / 1e12, 2), "unit": "Tlane-op/s",
              "frac": round(vfrac, 4), "lane_ops_per_update": round(lane_ops / nnz, 3),
              "min_lane_ops_per_update": VALU_MIN_OPS_PER_UPDATE}
        if mfrac >= vfrac:
            out = {"bound": "l1", "achieved": vmem["achieved"], "peak": vmem["peak"],
                   "unit": "TB/s", "frac": vmem["frac"]}
        else:
            out = {"bound": "valu", "achieved": vl["achieved"], "peak": vl["peak"],
                   "unit": "Tlane-op/s", "frac": vl["frac"]}
        out.update(base)
        out.update({"l1": vmem, "valu": vl, "hbm": hbm, "clock_GHz": CLOCK_GHZ,
                    "profiles": [f"profiles/{tag}_l1.json", f"profiles/{tag}_valu.json"]})
        return out
    if hbm is not None:
        out = {"bound": "hbm", "achieved": hbm["achieved"], "peak": HBM_PEAK_GBS,
               "unit": "GB/s", "frac": hbm["frac"]}
    else:
        out = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
               "frac": None}
    out.update(base)
    return out


def cpu_model():
    """The host CPU's model name and logical CPU count (/proc/cpuinfo)."""
    name, count = "unknown", 0
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    count += 1
                    name = line.split(":", 1)[1].strip()
    except OSError:
        pass
    return f"{name} ({count} logical CPUs on the host)"


def cpu_baseline(geo, img, sino, mode, gpu_sino, gpu_back, reps=3):
    """Reference cpuPrj (64 pthreads, prj/cpuPrj.c:40) forward and adjoint of the same
    slice, timed separately (wall clock, median of `reps` after a warm-up).  With the GPU
    outputs given, also their relative L2 errors against the reference's outputs and the
    FP64 oracle's (at the kernels' angles; the adjoint of the GPU's own sinogram)."""
    import numpy as np
    import oracle
    n, W, npj, pf, ds, er, d = geo
    cores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    if oracle.ref_available():
        kind, threads = "reference", 64
        fwd = lambda: oracle.ref_run(n, W, npj, pf, ds, er, d, img, oracle.FWD_BIT)  # noqa: E731
        adj = lambda: oracle.ref_run(n, W, npj, pf, ds, er, d, sino, oracle.BWD_BIT)  # noqa: E731
    else:
        kind, threads = "port", oracle._nth()
        cf = oracle.conf(*geo)
        fwd = lambda: oracle.forward(cf, img)[0]  # noqa: E731
        adj = lambda: oracle.adjoint(cf, sino)[0]  # noqa: E731

    def med(fn):
        out = fn()  # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), out

    t_f, ref_f = med(fwd)
    t_a, ref_a = med(adj)
    res = {"fwd_s": round(t_f, 4), "adj_s": round(t_a, 4), "kind": kind, "threads": threads,
           "cores": cores, "cpu_model": cpu_model(), "vs_reference": None, "vs_fp64": None}
    if gpu_sino is not None:
        def rel(a, b):
            a = np.asarray(a, np.float64).ravel()
            b = np.asarray(b, np.float64).ravel()
            return float(f"{np.linalg.norm(a - b) / np.linalg.norm(b):.3e}")
        if kind == "reference":
            res["vs_reference"] = {"fwd": rel(gpu_sino, ref_f), "adj": rel(gpu_back, ref_a)}
        cf = oracle.conf(*geo, angle_mode={0: 1, 1: 2, 2: 3, 3: 4}[mode])
        res["vs_fp64"] = {"fwd": rel(gpu_sino, oracle.forward(cf, img)[0]),
                          "adj": rel(gpu_back, oracle.adjoint(cf, sino)[0])}
    return res


STACK_CHECK = (0, 137, 300, 511)  # slices of the stack whose iterates are checked in FP64


def stack_problem(nslices, dev, tune_reps=2):
    """The slice stack's polychromatic Poisson problem: 512^2 parallel slices (720 views
    over 180 degrees, 729 bins), E = 20 spline bases, counts at I0 = 1e5."""
    import numpy as np
    import torch

    from imgrecsrc_amd.bhc import Spline, kappa_logspace, normalize_ie, spline_knots
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.workloads import STACK512, stack_truth

    P = Projector([STACK512] * nslices)
    P.autotune(reps=tune_reps, cache_path=TUNE_CACHE)
    truth = torch.from_numpy(stack_truth(nslices)).to(dev).reshape(-1)
    knots = spline_knots(kappa_logspace(20, 0.03, 30.0))
    Ie = normalize_ie(np.exp(-0.5 * ((np.arange(20) - 7.5) / 3.0) ** 2) + 0.05, knots)
    raw = P.forward(truth)
    Ts = float(3.0 / raw.max().item())
    BL = Spline("b1", knots).polyIout(raw * Ts, Ie, want=1)[0]
    g = torch.Generator(device=dev).manual_seed(6)
    counts = torch.poisson(1e5 * BL, generator=g)
    Imea = (counts / counts.max()).float()
    del raw, BL, counts, truth
    return P, Imea, knots, Ie, Ts


def stack_batched_vs_plain(P, Imea, knots, Ie, Ts, dev, stream):
    """The stack's forward + adjoint pair and its beam-hardening gradient (poissLAlpha),
    batched (one call over all slices: one launch per kernel) against the plain path (a
    one-slice plan called once per slice, on the slice's part of the same buffers)."""
    import torch

    from imgrecsrc_amd.bhc import poissLAlpha
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.workloads import STACK512, stack_truth

    ns = P.nslices
    one = Projector([STACK512])
    one.autotune(reps=2, cache_path=TUNE_CACHE)
    ni, nr = one.img_elems, one.sino_elems
    x = torch.from_numpy(stack_truth(ns)).to(dev).reshape(-1)
    y = torch.empty(P.sino_elems, dtype=torch.float32, device=dev)
    b = torch.empty(P.img_elems, dtype=torch.float32, device=dev)
    g = torch.empty(P.img_elems, dtype=torch.float32, device=dev)

    def batched():
        P.forward(x, out=y)
        P.adjoint(y, out=b)

    def plain():
        for k in range(ns):
            one.forward(x[k * ni:(k + 1) * ni], out=y[k * nr:(k + 1) * nr])
            one.adjoint(y[k * nr:(k + 1) * nr], out=b[k * ni:(k + 1) * ni])

    def grad_batched():
        poissLAlpha(Imea, Ie, x, P, Ts, knots)

    def grad_plain():
        for k in range(ns):
            poissLAlpha(Imea[k * nr:(k + 1) * nr], Ie, x[k * ni:(k + 1) * ni], one, Ts, knots)

    for fn in (batched, plain, grad_batched, grad_plain):
        fn()
    t_b, t_p = timed(batched, 3, stream), timed(plain, 3, stream)
    t_gb, t_gp = timed(grad_batched, 3, stream), timed(grad_plain, 3, stream)
    nnz = one.nnz() * ns
    del g
    return {"workload": f"{ns} x 512^2 parallel, 720 views x 729 bins",
            "fwd_adj_batched_ms": round(t_b, 3), "fwd_adj_plain_ms": round(t_p, 3),
            "fwd_adj_batched_GUPS": round(2 * nnz / t_b / 1e6, 1),
            "fwd_adj_plain_GUPS": round(2 * nnz / t_p / 1e6, 1),
            "fwd_adj_speedup": round(t_p / t_b, 3),
            "grad_batched_ms": round(t_gb, 3), "grad_plain_ms": round(t_gp, 3),
            "grad_batched_GUPS": round(2 * nnz / t_gb / 1e6, 1),
            "grad_speedup": round(t_gp / t_gb, 3),
            "tune_batched": P.get_tune(), "tune_plain": one.get_tune()}


def pnpg_leg(nslices, iters, dev, stream, compare=True):
    """PNPG alpha step (npg/PNPG.m:50-167, non-negativity prox) on the slice stack: one
    line-search trial (gradient evaluation: forward + FP64 NLL + adjoint; cost
    evaluation: forward + FP64 NLL; decision) captured once in a HIP graph and replayed
    until `iters` iterations are accepted; the alpha of STACK_CHECK's slices is kept after
    every accepted iteration, and the device's f and grad_alpha of those slices at each
    kept iterate are evaluated afterwards (for the FP64 comparison of the CPU leg)."""
    import numpy as np
    import torch

    from imgrecsrc_amd.bhc import poissLAlpha
    from imgrecsrc_amd.pnpg import PNPG
    from imgrecsrc_amd.projector import Projector
    from imgrecsrc_amd.workloads import STACK512

    log(f"pnpg leg: {nslices} slices, {iters} iterations")
    P, Imea, knots, Ie, Ts = stack_problem(nslices, dev)
    stack_cmp = None
    if compare:
        stack_cmp = stack_batched_vs_plain(P, Imea, knots, Ie, Ts, dev, stream)
        log(f"stack batched vs plain: {stack_cmp}")
    check = [k for k in STACK_CHECK if k < nslices]
    img = STACK512[0] ** 2
    alpha = torch.full((P.img_elems,), 0.02, dtype=torch.float32, device=dev)
    solver = PNPG(P, Imea, knots, Ie, Ts, alpha)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    with torch.cuda.graph(graph, stream=cap):
        solver.step(alpha)
    torch.cuda.synchronize()
    st0 = solver.stats()
    c0 = st0["cost"]
    snaps = [alpha.view(-1, img)[check].clone()]
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    ms, trials, last, costs = 0.0, 0, 0, [c0]
    t_wall = time.perf_counter()
    while last < iters and trials < 4 * iters:
        e0.record(stream)
        graph.replay()
        e1.record(stream)
        e1.synchronize()
        ms += e0.elapsed_time(e1)
        trials += 1
        st = solver.stats()
        if st["iterations"] > last:
            last = st["iterations"]
            costs.append(st["cost"])
            snaps.append(alpha.view(-1, img)[check].clone())
    wall = time.perf_counter() - t_wall
    st = solver.stats()
    del solver, graph
    nnz = Projector([STACK512]).nnz() * nslices
    # the device's f and grad_alpha of the checked slices at every kept iterate
    P4 = Projector([STACK512] * len(check))
    Im4 = Imea.view(-1, STACK512[1] * STACK512[2])[check].reshape(-1).contiguous()
    dev_f, dev_g, alphas = [], [], []
    for a in snaps:
        f, g = poissLAlpha(Im4, Ie, a.reshape(-1), P4, Ts, knots)
        dev_f.append(f.item())
        dev_g.append(g.cpu().numpy())
        alphas.append(a.cpu().numpy())
    out = {"workload": f"{nslices} x 512^2 parallel, 720 views x 729 bins, E=20, I0=1e5",
           "iterations": st["iterations"], "trials": trials,
           "ms_per_trial": round(ms / trials, 3),
           "ms_per_iteration": round(ms / max(1, st["iterations"]), 3),
           "wall_s": round(wall, 3),
           "GUPS": round(3 * nnz * trials / (ms * 1e-3) / 1e9, 1),
           "cost_first": c0, "cost_last": st["cost"],
           "cost_monotone": bool(all(b <= a for a, b in zip(costs, costs[1:]))),
           "rejects": st["rejects"], "restarts": st["restarts"], "t": st["t"],
           "checked_slices": check, "batched_vs_plain": stack_cmp}
    check_data = {"alphas": alphas, "f": dev_f, "g": dev_g,
                  "Imea": Im4.cpu().numpy().astype(np.float64), "Ie": Ie, "knots": knots,
                  "Ts": Ts, "mode": P4.modes()[0]}
    return out, check_data


def stack_fp64_check(cd):
    """FP64 f and grad_alpha (bhc/poissLAlpha.m restated, over the oracle's stored A) of
    the checked slices at every kept iterate; relative errors of the device's values."""
    import numpy as np

    import oracle
    from oracle import bhc_oracle as B
    from imgrecsrc_amd.workloads import STACK512

    ANGLE_MODE = {0: 1, 1: 2, 2: 3, 3: 4}
    t0 = time.perf_counter()
    M = oracle.Matrix(oracle.conf(*STACK512, angle_mode=ANGLE_MODE[cd["mode"]]))
    t_build = time.perf_counter() - t0
    Ts = cd["Ts"]
    Phi = lambda a: Ts * M.forward(a).ravel()  # noqa: E731
    Phit = lambda y: Ts * M.adjoint(y).ravel()  # noqa: E731
    pol = lambda k, s, I: B.b1_iout_stable(k, s, I)  # noqa: E731
    from concurrent.futures import ThreadPoolExecutor

    def check(k):
        a, f, g = cd["alphas"][k], cd["f"][k], cd["g"][k]
        f64, g64, _ = B.poiss_l_alpha(cd["Imea"], cd["Ie"], a.astype(np.float64).ravel(), Phi,
                                      Phit, pol, cd["knots"])
        return abs(f - f64) / abs(f64), float(np.linalg.norm(g.ravel() - g64) / np.linalg.norm(g64))

    # numpy's array operations and the oracle's sparse products release the GIL: the
    # iterates are checked by a pool of threads
    n_it = len(cd["alphas"])
    ef, eg = [0.0] * n_it, [0.0] * n_it
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=8) as ex:
        for k, (e1, e2) in enumerate(ex.map(check, range(n_it))):
            ef[k], eg[k] = e1, e2
            log(f"FP64 check iterate {k}: f {e1:.2e}, grad {e2:.2e}")
    t_eval = time.perf_counter() - t0
    return {"f_rel_err": [float(f"{e:.3e}") for e in ef],
            "grad_rel_l2": [float(f"{e:.3e}") for e in eg],
            "max_f_rel_err": max(ef), "max_grad_rel_l2": max(eg),
            "fp64_s_per_eval": round(t_eval / len(ef), 3), "matrix_build_s": round(t_build, 2),
            "matrix_nnz": M.nnz}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--settle-s", type=float, default=0.25,
                    help="seconds of untimed fwd+adj pairs before the warm-up steps: the "
                         "device clock settles over the first ~0.1 s of load (0: none)")
    ap.add_argument("--workload", default="fan1024", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="only the headline slice")
    ap.add_argument("--legs", default="bhc,npg,batched,pcie,stack,par256",
                    help="comma-separated extra legs to run (bhc, npg, batched, pcie, stack, "
                         "par256); npg and batched also run the bhc set-up")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (PMC scaling reference)")
    ap.add_argument("--pnpg-slices", type=int, default=512,
                    help="slices of the 512^2 parallel stack in the PNPG leg (0: skip)")
    ap.add_argument("--pnpg-iters", type=int, default=50,
                    help="PNPG iterations of the stack leg (graph-replayed trials)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_DIR", "WRITE_DIR", "KERNEL", "TAG"),
                    help="reduce two rocprofv3 --pmc csv runs to profiles/TAG_pmc.json")
    ap.add_argument("--pmc-command", default="", help="command recorded in the summary")
    ap.add_argument("--valu-summary", nargs=3, metavar=("SQ_DIR", "KERNEL", "TAG"),
                    help="reduce a rocprofv3 --pmc csv run of SQ counters to "
                         "profiles/TAG_valu.json")
    ap.add_argument("--summary-suffix", default="valu",
                    help="file suffix of a --valu-summary (profiles/TAG_SUFFIX.json)")
    ap.add_argument("--valu-tune", default=None,
                    help="JSON of the tuned shapes the counter run used (recorded in the "
                         "summary; the bench uses a summary only at the same shapes)")
    ap.add_argument("--tune-cache", default=DEFAULT_TUNE_CACHE,
                    help="autotuner cache file (every leg's plans): shapes recorded there "
                         "for this device and geometry are reused, others are tuned and "
                         "appended; the default is the committed MI355X cache, whose shapes "
                         "the committed counter summaries were taken at; 'none' tunes "
                         "afresh")
    args = ap.parse_args()
    global TUNE_CACHE
    TUNE_CACHE = None if args.tune_cache in (None, "", "none") else args.tune_cache
    legs = set(x.strip() for x in args.legs.split(",") if x.strip())
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, args.pmc_command)
        return
    if args.valu_summary:
        valu_summary(*args.valu_summary, args.pmc_command, args.valu_tune, args.summary_suffix)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(slices shard across GPUs with no exchange; see the design notes). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from imgrecsrc_amd.bhc import kappa_logspace, normalize_ie, poissLAlpha, poissLI, spline_knots
    from imgrecsrc_amd.phantom import ellipse_phantom, gaussian, rasterise, ellipse_params
    from imgrecsrc_amd.projector import Projector

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()
    geo = WORKLOADS[args.workload]
    n, W, npj = geo[0], geo[1], geo[2]
    log(f"workload {args.workload} {geo}")
    P = Projector([geo])
    tune = P.autotune(reps=5, cache_path=TUNE_CACHE)
    log(f"autotuned shapes {tune}")
    img_np = ellipse_phantom(n, 2)
    img = torch.from_numpy(img_np).to(dev).reshape(-1)
    sino = torch.empty(P.sino_elems, dtype=torch.float32, device=dev)
    back = torch.empty(P.img_elems, dtype=torch.float32, device=dev)
    nnz = P.nnz()
    log(f"nnz(A) = {nnz}")

    if args.calib_copy:
        a = torch.empty(64 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("calibration copy: 3 x 256 MiB")

    def step():
        P.forward(img, out=sino)
        P.adjoint(sino, out=back)

    # untimed load until the device clock has settled (a cold device runs the first
    # ~0.1 s of work slower), then the W warm-up steps and the K timed steps
    settle_steps = 0
    t_settle = time.perf_counter()
    while time.perf_counter() - t_settle < args.settle_s:
        for _ in range(50):
            step()
        settle_steps += 50
        torch.cuda.synchronize()
    if settle_steps:
        log(f"settle: {settle_steps} untimed pairs in {time.perf_counter() - t_settle:.2f} s")
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    value = 2.0 * nnz * args.steps / wall / 1e9
    log(f"fwd+adj: {ms_step:.3f} ms/step, {value:.1f} GUPS")

    # per-kernel averages with HIP events on the launch stream
    reps = max(5, min(args.steps, 50))
    t_fwd = timed(lambda: P.forward(img, out=sino), reps, stream)
    t_adj = timed(lambda: P.adjoint(sino, out=back), reps, stream)
    sfx = "_sym" if P.modes()[0] != 0 else ""
    # the forward call also runs the 32x32 transpose of the image (a few microseconds)
    dom_name, dom_ms = (f"k_adjoint{sfx}", t_adj) if t_adj >= t_fwd else (f"k_forward{sfx}", t_fwd)
    launches = (["k_adjoint_sym"] + (["k_adjoint_sym_combine"] if tune["adj_split"] > 1 else [])
                if dom_name.startswith("k_adjoint") else
                ["k_transpose", "k_forward_sym"] + (["k_forward_sym_combine"]
                                                    if tune["fwd_split"] > 1 else []))
    tag = f"{'adjoint' if dom_name.startswith('k_adjoint') else 'forward'}_{args.workload}"
    roofline = roofline_entry(tag, dom_name, dom_ms, launches, nnz, tune,
                              t_fwd if dom_name.startswith("k_adjoint") else t_adj)

    # the pair replayed as one HIP graph next to the same pair launched eagerly: what
    # the call's launches (transpose, forward, its combine, adjoint, its combine) cost
    step()
    torch.cuda.synchronize()
    back_eager = back.clone()
    t_pair = timed(step, reps, stream)
    t_pair_graph, _, pair_graph = graph_timed(step, reps, stream)
    pair_replay = {"eager_ms": round(t_pair, 4), "graph_ms": round(t_pair_graph, 4),
                   "launches_per_pair": (2 + (tune["fwd_split"] > 1)) + (1 + (tune["adj_split"] > 1)),
                   "replay_equals_eager": bool(torch.equal(back, back_eager))}
    del pair_graph
    log(f"pair eager {t_pair:.4f} ms, graph replay {t_pair_graph:.4f} ms")

    # <Ax, y> = <x, A^T y> on seeded Gaussian x, y
    x = torch.from_numpy(gaussian((n, n), 21)).to(dev).reshape(-1)
    y = torch.from_numpy(gaussian((npj, W), 22)).to(dev).reshape(-1)
    lhs = torch.dot(P.forward(x).double(), y.double()).item()
    rhs = torch.dot(x.double(), P.adjoint(y).double()).item()
    adj_rel = abs(lhs - rhs) / abs(lhs)

    out = {
        "metric": "fwd+adjoint voxel-ray updates/s (GUPS), 1024^2 fan slice",
        "value": round(value, 2), "unit": "GUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "settle": {"seconds": args.settle_s, "pairs": settle_steps},
        "ms_per_step": round(ms_step, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded ellipse phantom, FP32)",
        "config": {"workload": f"{args.workload}: n={n} prjWidth={W} np={npj} "
                               f"prjFull={geo[3]} d={geo[6]}", "nnz": nnz,
                   "parallelism": "single GPU", "symmetry_group": P.modes()[0],
                   "tune": tune},
        "adjoint_test_rel": adj_rel,
        "pair_graph_replay": pair_replay,
        "roofline": roofline,
    }

    if not args.no_extras and legs & {"bhc", "npg", "batched"}:
        # ---- beam-hardening gradient evaluations (512^2 fan, E = 20)
        log("bhc leg")
        bg = FAN512
        PB = Projector([bg])
        PB.autotune(reps=3, cache_path=TUNE_CACHE)
        knots = spline_knots(kappa_logspace(20, 0.03, 30.0))
        xs = np.arange(20)
        Ie = normalize_ie(np.exp(-0.5 * ((xs - 7.5) / 3.0) ** 2) + 0.05, knots)
        alpha_np = rasterise(ellipse_params(3, bg[0]), bg[0]).astype(np.float32)
        alpha = torch.from_numpy(alpha_np).to(dev).reshape(-1)
        raw = PB.forward(alpha)
        Ts = float(3.0 / raw.max().item())
        from imgrecsrc_amd.bhc import Spline
        BL = Spline("b1", knots).polyIout(raw * Ts, Ie, want=1)[0]
        g = torch.Generator(device="cpu").manual_seed(5)
        counts = torch.poisson(1e5 * BL.double().cpu(), generator=g)
        Imea = (counts / counts.max()).float().to(dev)
        a_eval = alpha * 0.8 + 0.05
        for _ in range(3):
            poissLAlpha(Imea, Ie, a_eval, PB, Ts, knots)
        bnnz = PB.nnz()
        # the same plan's projector pair: the gradient's floor (its ray passes and
        # reductions come on top of one forward and one adjoint)
        bs = torch.empty(PB.sino_elems, dtype=torch.float32, device=dev)
        bb = torch.empty(PB.img_elems, dtype=torch.float32, device=dev)
        calls = {"ga": lambda: poissLAlpha(Imea, Ie, a_eval, PB, Ts, knots),
                 "gi": lambda: poissLI(Imea, Ie, PB, Ts, knots, alpha=a_eval),
                 "pair": lambda: (PB.forward(a_eval, out=bs), PB.adjoint(bs, out=bb)),
                 "fw": lambda: PB.forward(a_eval, out=bs)}
        # the same calls replayed from HIP graphs: device time without the host's launch
        # overhead (every entry point here allocates nothing, so each call captures)
        graphs = {k: graph_timed(fn, 1, stream) for k, fn in calls.items()}
        # the device clock settles over the first ~100 ms of load: warm up, then time the
        # calls in interleaved rounds and keep each one's median
        for _ in range(200):
            calls["pair"]()
        ts = {k: [] for k in calls}
        tg = {k: [] for k in calls}
        for _ in range(5):
            for k, fn in calls.items():
                ts[k].append(timed(fn, 20, stream))
                tg[k].append(timed(graphs[k][2].replay, 50, stream))
        med = {k: float(np.median(v)) for k, v in ts.items()}
        gmed = {k: float(np.median(v)) for k, v in tg.items()}
        t_ga, t_gi, t_pair, t_fw = med["ga"], med["gi"], med["pair"], med["fw"]
        g_ga, g_gi, g_pair, g_fw = gmed["ga"], gmed["gi"], gmed["pair"], gmed["fw"]
        ga_out, gga = graphs["ga"][1], graphs["ga"][2]
        f_eager = poissLAlpha(Imea, Ie, a_eval, PB, Ts, knots)[0].item()
        gga.replay()
        torch.cuda.synchronize()
        f_graph = ga_out[0].item()
        del gga, graphs
        out["bhc"] = {"workload": "fan 512^2, 360 views x 729 bins, d=2000, E=20",
                      "grad_alpha_evals_per_s": round(1e3 / t_ga, 1),
                      "grad_alpha_ms": round(t_ga, 4),
                      "grad_ie_evals_per_s": round(1e3 / t_gi, 1), "grad_ie_ms": round(t_gi, 4),
                      "fwd_adj_ms": round(t_pair, 4), "fwd_ms": round(t_fw, 4),
                      "grad_alpha_over_fwd_adj": round(t_ga / t_pair, 3),
                      "grad_ie_over_fwd": round(t_gi / t_fw, 3),
                      "grad_GUPS": round(2 * bnnz / t_ga / 1e6, 1), "nnz": bnnz,
                      "graph_replay": {"grad_alpha_ms": round(g_ga, 4),
                                       "fwd_adj_ms": round(g_pair, 4),
                                       "grad_alpha_over_fwd_adj": round(g_ga / g_pair, 3),
                                       "grad_ie_ms": round(g_gi, 4), "fwd_ms": round(g_fw, 4),
                                       "grad_ie_over_fwd": round(g_gi / g_fw, 3),
                                       "grad_alpha_evals_per_s": round(1e3 / g_ga, 1),
                                       "f_replay_equals_eager": f_graph == f_eager},
                      "tune": PB.get_tune()}

        # ---- I-epsilon step (20 iterations) and the measurement simulator; both block
        from imgrecsrc_amd.bhc import gen_beam_harden, ie_solve
        Ie0 = normalize_ie(np.ones(20), knots)
        ie_solve(Imea, Ie0, PB, Ts, knots, alpha=a_eval, maxit=2)
        t0 = time.perf_counter()
        _, info = ie_solve(Imea, Ie0, PB, Ts, knots, alpha=a_eval, maxit=20, tol=0.0)
        t_ie = (time.perf_counter() - t0) * 1e3
        eps_s = np.arange(8.0, 141.0)  # synthetic 140 kV tungsten-like spectrum, 1 keV bins
        iota_s = np.maximum(140.0 - eps_s, 0.0) / eps_s * np.exp(-(12.0 / eps_s) ** 3)
        kap_s = (0.12 * (eps_s / 60.0) ** -2.7 + 0.03)[None, :]
        gen_beam_harden(PB, alpha, [7.874], kap_s, eps_s, iota_s)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            sim = gen_beam_harden(PB, alpha, [7.874], kap_s, eps_s, iota_s)
        torch.cuda.synchronize()
        t_sim = (time.perf_counter() - t0) * 1e3 / 5
        out["bhc"].update({
            "ie_solve_ms": round(t_ie, 3), "ie_solve_iterations": info["iterations"],
            "ie_solve_passes": info["passes"],
            "gen_beam_harden_ms": round(t_sim, 3), "gen_beam_harden_Ts": sim["Ts"],
            "gen_beam_harden_trials": int(round(np.log(sim["Ts"]) / np.log(0.8))) + 1,
            "gen_beam_harden_energies": int(eps_s.size)})

        if "npg" in legs:
            # ---- the reference's default alpha step: NPG with the wavelet-ADMM prior (Haar,
            #      6 levels) on a circle support, same 512^2 fan problem; trials synchronise
            log("npg + wavelet leg")
            from imgrecsrc_amd.pnpg import PNPG
            from imgrecsrc_amd.wavelet import Wavelet, daubcqf, wvlt_mask
            nb = bg[0]
            cc = np.arange(nb) - (nb - 1) / 2
            XX, YY = np.meshgrid(cc, cc)
            disc = XX ** 2 + YY ** 2 <= (nb / 2 - 1) ** 2
            hw = daubcqf(2)[0]
            wv = Wavelet(nb, nb, hw, 6)
            wv.set_masks(disc, wvlt_mask(disc, hw, 6))
            a_w = (alpha * torch.from_numpy(disc.ravel().astype(np.float32)).to(dev)).contiguous()
            npg = PNPG(PB, Imea, knots, Ie, Ts, a_w)
            npg.set_method("NPG")
            npg.set_mask(disc)
            npg.set_wavelet(wv, 1e-3 * npg.t0)
            ntr, inner = 10, 0
            npg.step(a_w)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(ntr):
                npg.step(a_w)
                inner += npg.stats()["innerSearch"]
            torch.cuda.synchronize()
            t_npg = (time.perf_counter() - t0) * 1e3 / ntr
            st = npg.stats()
            out["npg_wavelet"] = {
                "workload": "NPG + wavelet-ADMM prior (daubcqf(2), 6 levels, circle support), "
                            "fan 512^2, E=20",
                "ms_per_trial": round(t_npg, 3), "admm_iterations_per_trial": inner / ntr,
                "iterations": st["iterations"], "trials": st["trials"], "cost": st["cost"]}
            del npg, wv

            # ---- the same step with the isotropic TV prior ('tviso', utils/TV.m denoise)
            log("npg + tv leg")
            from imgrecsrc_amd.tv import TV
            tvp = TV(nb, nb, tv="iso")
            tvp.set_mask(disc)
            npg = PNPG(PB, Imea, knots, Ie, Ts, a_w)
            npg.set_method("NPG")
            npg.set_mask(disc)
            npg.set_tv(tvp, 1e-3 * npg.t0)
            inner = 0
            npg.step(a_w)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(ntr):
                npg.step(a_w)
                inner += npg.stats()["innerSearch"]
            torch.cuda.synchronize()
            t_tv = (time.perf_counter() - t0) * 1e3 / ntr
            st = npg.stats()
            out["npg_tv"] = {
                "workload": "NPG + isotropic TV prior (FGP), circle support, fan 512^2, E=20",
                "ms_per_trial": round(t_tv, 3), "fgp_iterations_per_trial": inner / ntr,
                "iterations": st["iterations"], "trials": st["trials"], "cost": st["cost"]}
            del npg, tvp

        if "batched" in legs:
            # ---- 64 ragged slices: one launch per operator vs one launch per slice
            log("batched leg")
            geoms = ragged_geoms()
            PBat = Projector(geoms)
            PBat.autotune(reps=3, cache_path=TUNE_CACHE)
            imgs = [torch.from_numpy(ellipse_phantom(gq[0], 100 + i)).to(dev).reshape(-1)
                    for i, gq in enumerate(geoms)]
            bimg = PBat.pack_images(imgs)
            bsino = torch.empty(PBat.sino_elems, dtype=torch.float32, device=dev)
            bback = torch.empty(PBat.img_elems, dtype=torch.float32, device=dev)
            tot_nnz = PBat.nnz()
            singles = [Projector([gq]) for gq in geoms]
            for s1 in singles:
                s1.autotune(reps=3, cache_path=TUNE_CACHE)
            # the plain path on the general (per-view) kernels, one slice per launch
            generals = [Projector([gq], symmetric=False) for gq in geoms]
            for s1 in generals:
                s1.autotune(reps=3, cache_path=TUNE_CACHE)
            s_out = [torch.empty(s.sino_elems, dtype=torch.float32, device=dev) for s in singles]
            i_out = [torch.empty(s.img_elems, dtype=torch.float32, device=dev) for s in singles]

            def batched():
                PBat.forward(bimg, out=bsino)
                PBat.adjoint(bsino, out=bback)

            def plain():
                for k, s in enumerate(singles):
                    s.forward(imgs[k], out=s_out[k])
                    s.adjoint(s_out[k], out=i_out[k])

            def plain_general():
                for k, s in enumerate(generals):
                    s.forward(imgs[k], out=s_out[k])
                    s.adjoint(s_out[k], out=i_out[k])

            batched()
            plain()
            plain_general()
            t_b = timed(batched, 3, stream)
            t_p = timed(plain, 3, stream)
            t_g = timed(plain_general, 3, stream)
            # both paths captured as HIP graphs (the batched one is a handful of launches,
            # the plain one 128 pairs of launch sequences) and replayed
            tg_b, _, gb_ = graph_timed(batched, 5, stream)
            tg_p, _, gp_ = graph_timed(plain, 5, stream)
            del gb_, gp_
            groups = PBat.modes()
            # per symmetry group: the group's slices in one plan against one plan per slice
            # (the groups differ in how often a weight is reused: D4 8, V4 4, Z2 2 times)
            per_group = {}
            for gid, gname in ((1, "D4"), (2, "V4"), (3, "Z2"), (0, "general")):
                idx = [k for k, m in enumerate(groups) if m == gid]
                if not idx:
                    continue
                Pg = Projector([geoms[k] for k in idx])
                Pg.autotune(reps=3, cache_path=TUNE_CACHE)
                gimg = Pg.pack_images([imgs[k] for k in idx])
                gs = torch.empty(Pg.sino_elems, dtype=torch.float32, device=dev)
                gb = torch.empty(Pg.img_elems, dtype=torch.float32, device=dev)

                def gbat():
                    Pg.forward(gimg, out=gs)
                    Pg.adjoint(gs, out=gb)

                def gplain():
                    for k in idx:
                        singles[k].forward(imgs[k], out=s_out[k])
                        singles[k].adjoint(s_out[k], out=i_out[k])

                gbat()
                gplain()
                tb_g, tp_g = timed(gbat, 3, stream), timed(gplain, 3, stream)
                gnnz = Pg.nnz()
                per_group[gname] = {"slices": len(idx), "batched_ms": round(tb_g, 3),
                                    "plain_ms": round(tp_g, 3),
                                    "batched_GUPS": round(2 * gnnz / tb_g / 1e6, 1),
                                    "plain_GUPS": round(2 * gnnz / tp_g / 1e6, 1),
                                    "speedup": round(tp_g / tb_g, 3)}
                del Pg, gimg, gs, gb
            out["batched"] = {"workload": "64 ragged slices, n in {128..1024}, views 90..720, "
                                          "mixed parallel/fan", "nnz": tot_nnz,
                              "batched_ms": round(t_b, 3), "plain_ms": round(t_p, 3),
                              "batched_GUPS": round(2 * tot_nnz / t_b / 1e6, 1),
                              "plain_GUPS": round(2 * tot_nnz / t_p / 1e6, 1),
                              "speedup": round(t_p / t_b, 3),
                              "plain_general_ms": round(t_g, 3),
                              "plain_general_GUPS": round(2 * tot_nnz / t_g / 1e6, 1),
                              "speedup_vs_general": round(t_g / t_b, 3),
                              "graph_replay": {"batched_ms": round(tg_b, 3),
                                               "plain_ms": round(tg_p, 3),
                                               "batched_GUPS": round(2 * tot_nnz / tg_b / 1e6, 1),
                                               "speedup": round(tg_p / tg_b, 3)},
                              "slices_per_group": {str(k): groups.count(k) for k in range(4)},
                              "per_group": per_group}

    if not args.no_extras and "pcie" in legs:
        # ---- the reference-compatible boundary: host arrays through prj_run, PCIe included
        import ctypes
        from imgrecsrc_amd._lib import BWD_BIT, FWD_BIT, check, lib
        L = lib()
        check(L.prj_setup(n, W, npj, geo[3], geo[4], geo[5], geo[6]))
        h_img = np.ascontiguousarray(img_np, dtype=np.float32)
        h_sino = np.empty(npj * W, np.float32)
        h_back = np.empty(n * n, np.float32)

        def host_pair():
            check(L.prj_run(h_img.ctypes.data, h_sino.ctypes.data, FWD_BIT))
            check(L.prj_run(h_back.ctypes.data, h_sino.ctypes.data, BWD_BIT))

        host_pair()
        reps_h = 10
        t0 = time.perf_counter()
        for _ in range(reps_h):
            host_pair()
        t_h = (time.perf_counter() - t0) / reps_h
        out["pcie_inclusive"] = {"GUPS": round(2.0 * nnz / t_h / 1e9, 1),
                                 "ms_per_pair": round(1e3 * t_h, 3),
                                 "path": "prj_setup + prj_run(FWD), prj_run(BWD) on pageable "
                                         "host arrays (the MEX boundary), default shapes"}

    check_data = None
    if not args.no_extras and args.pnpg_slices > 0 and "stack" in legs:
        out["pnpg"], check_data = pnpg_leg(args.pnpg_slices, args.pnpg_iters, dev, stream)

    if not args.no_extras and "par256" in legs:
        # ---- the 256^2 parallel slice (180 views over 360, 367 bins): plain GPU times
        PC1 = Projector([PAR256])
        PC1.autotune(reps=3, cache_path=TUNE_CACHE)
        c1_img = torch.from_numpy(ellipse_phantom(256, 1)).to(dev).reshape(-1)
        c1_s = PC1.forward(c1_img)
        c1_b = PC1.adjoint(c1_s)
        out["par256"] = {"workload": "par 256^2, 180 views (prjFull 360), 367 bins",
                         "nnz": PC1.nnz(),
                         "gpu_fwd_ms": round(timed(lambda: PC1.forward(c1_img, out=c1_s), 20,
                                                   stream), 4),
                         "gpu_adj_ms": round(timed(lambda: PC1.adjoint(c1_s, out=c1_b), 20,
                                                   stream), 4)}

    if not args.no_cpu_baseline:
        log("cpu baseline (reference cpuPrj)")
        sino_np = sino.cpu().numpy()
        back_np = back.cpu().numpy()
        cb = cpu_baseline(geo, img_np, sino_np, P.modes()[0], sino_np, back_np)
        out["cpu_baseline"] = {"value": round(2.0 * nnz / (cb["fwd_s"] + cb["adj_s"]) / 1e9, 3),
                               "unit": "GUPS", "cores": cb["cores"], "threads": cb["threads"],
                               "kind": cb["kind"], "cpu_model": cb["cpu_model"],
                               "fwd_s": cb["fwd_s"], "adj_s": cb["adj_s"],
                               "sample": f"one fwd + one adj of the {args.workload} slice, "
                                         f"median of 3 after 1 warm-up each",
                               "gpu_vs_reference": cb["vs_reference"],
                               "gpu_vs_fp64": cb["vs_fp64"]}
        if "par256" in out:
            c1 = cpu_baseline(PAR256, ellipse_phantom(256, 1), c1_s.cpu().numpy(), None,
                              None, None)
            out["par256"].update({"cpu_fwd_ms": round(1e3 * c1["fwd_s"], 3),
                                  "cpu_adj_ms": round(1e3 * c1["adj_s"], 3),
                                  "cpu_kind": c1["kind"], "cpu_threads": c1["threads"]})
        if check_data is not None:
            log("FP64 check of the stack's iterates")
            out["pnpg"]["fp64_check"] = stack_fp64_check(check_data)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
