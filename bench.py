#!/usr/bin/env python3
"""Benchmark of the local analysis: analysis gridpoints/sec on one MI355X.

Workload (default): the 0.25-degree 1440x720 surface analysis with k = 50,
2M observations on 30 synthetic swath arcs and a 300 km cut-off radius
(seeded synthetic data, inputs resident in HBM before the timed region).
A step is one full analysis (letkf_gpu_run) over all 1,036,800 gridpoints,
on the batched path with split-BF16 products of all four terms and
compensated accumulation (bf16x4); the three-term split (bf16x3) is timed
beside it on the same inputs.

Printed: one JSON line with the metric, the roofline of the product kernel
(the largest single kernel; from the library's HIP events on its own
stream, with the HBM traffic recorded from rocprofv3 counters in
profiles/traffic.json), the CPU baseline (the FP64 oracle on a bounded
sample of gridpoints, all host threads up to 16) and the error of the GPU
analysis against that oracle on the same sample.

--config c5 runs the cycled analysis instead (ten cycles on the same grid,
fresh observations each cycle) and prints per-cycle latency percentiles;
--config c4 the 0.5-degree 75-level four-variable analysis at k = 100 with
the error of every precision against the oracle.
"""
import argparse
import dataclasses
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "umd-letkf_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0 / 3.0,
        "bf16x4": 2500.0 / 4.0}
HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_problem(name):
    from letkf_amd import synth
    if name == "c3":
        return synth.config3(), ("C3: 0.25deg 1440x720 surface, k=50, 2M swath obs, "
                                 "300 km cut-off (L=82.2 km), RTPS 0.6")
    if name == "c1":
        return synth.config1(), "C1: 4deg 90x45 SST, k=20, 5000 obs, GC 500 km, RTPS 0.6"
    if name == "c2":
        return synth.config2(), ("C2: 1deg 360x180x50 T/S, k=40, 4000 profiles x 50, "
                                 "700->200 km, vertical Gaussian 300 m")
    raise SystemExit(f"unknown config {name}")


PROFILES = os.path.join(ROOT, "profiles")
TRAFFIC_FILE = os.path.join(PROFILES, "traffic.json")
# rocprofv3 counter scale factors measured on a 1 GiB device copy (FETCH_SIZE
# read 524308 KiB for 1048576 KiB moved; WRITE_SIZE read 1048576 KiB)
FETCH_SCALE = 2.0
WRITE_SCALE = 1.0


def product_kernel(path, precision, tun=None, k=None):
    """The product kernel a run launches: the plain path's FP64 kernel, or the
    batched path's 16x16-tile kernel when the launch choice asks for it and
    it exists (BF16 and split-BF16 products, k <= 111), else the 32x32 one."""
    if path == "plain":
        return "k_product_f64"
    t16 = (tun or {}).get("tile", 1) and precision != "fp32" and (
        k is None or (k + 1 + 15) // 16 <= 7)
    return "k_product_t16" if t16 else "k_product_mx"


def tuning_key(tun):
    """The launch choices that shape the product kernel's memory traffic:
    problem order, bucket edges, XCD interleave, tile shape and (when not
    the default 2^20) the chunk size."""
    if not tun:
        return "default"
    t = {"xcd": 2, "tile": 1, "chunk_log2": 20, **tun}
    return ("order{order}_bb{bucket_bits}_xcd{xcd}".format(**t) + ("_t16" if t["tile"] else "")
            + (f"_c{t['chunk_log2']}" if t["chunk_log2"] != 20 else ""))


def sweep_tunings(spec):
    """Launch choices for --product-sweep: "tuner" is every product-side
    combination the autotuner can reach (order 0..3 with bucket edges 0..2
    for the bucketed orders, both tile shapes, XCD interleave 0..5), else a
    ';'-separated list of order,bucket_bits,xcd,tile."""
    base = dict(tridiag_reg=2, overlap=0, chunk_log2=20)
    out = []
    if spec == "tuner":
        for order in range(4):
            for bb in range(3 if order >= 2 else 1):
                for tile in (1, 0):
                    for xcd in range(6):
                        out.append(dict(order=order, bucket_bits=bb, xcd=xcd, tile=tile, **base))
        return out
    for item in spec.split(";"):
        o, b, x, t = (int(v) for v in item.split(","))
        out.append(dict(order=o, bucket_bits=b, xcd=x, tile=t, **base))
    return out


def product_sweep(args):
    """One analysis per launch choice of the sweep, in order (for per-dispatch
    counters of the product kernel under rocprofv3 --pmc)."""
    from letkf_amd.solver import LetkfGpuSolver
    prob, _ = make_problem(args.config)
    s = LetkfGpuSolver(prob, precision=args.precision, path=args.path)
    keys, problems = [], 0
    try:
        for t in sweep_tunings(args.product_sweep):
            s.set_tuning(**t)
            s.run()
            tm = s.timing()
            keys.append(tuning_key(t))
            problems = int(tm["n_problems"])
            log(f"[sweep] {len(keys) - 1}: {keys[-1]} product {tm['product_ms']:.2f} ms "
                f"launches {int(tm['n_launch'])}")
    finally:
        s.final()
    print(json.dumps({"sweep": keys, "problems_per_dispatch": problems}))


def roofline(t, k, path, precision, workload, tun=None, slabs_per_problem=1.0, ns=1,
             n_gp=0, n_obs=0):
    """Roofline of the product kernel (the largest single kernel): the
    algorithmic flops F_prod = sum over problems of 2k^2 m + 2k m per launch
    over the kernel's own HIP-event time on the library stream.  Also the
    whole analysis: T_roof = max(F_prod / P_prod + F_eig / P_fp64 +
    F_apply / P_apply, B_comp / BW) over the measured analysis time, with
    F_eig = 11k^3 + 4k^2 per problem (symmetric QR with vectors, the
    transform W, w), F_apply = 2k^2 per slab of each problem and the
    compulsory bytes B_comp = N_gp (8 ns k + 16 ns + 12) + N_obs (k+3) 4."""
    flop = (2.0 * k * k + 2.0 * k) * t["n_pairs"]
    ms = t["product_ms"]
    nl = max(int(t["n_launch"]), 1)
    peak = PEAK["fp64"] if path == "plain" else PEAK[precision]
    ach = flop / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    kern = product_kernel(path, precision, tun, k)
    gather = t["n_pairs"] * (4.0 * k + 8.0)          # hx column + (index, weight)
    roof = dict(kernel=kern, bound="mfma", achieved=round(ach, 3), peak=peak, unit="TFLOP/s",
                frac=round(ach / peak, 5), traffic=None,
                launches=nl, avg_launch_ms=round(ms / nl, 4),
                algorithmic=f"{flop / nl:.4e} flop and {gather / nl:.4e} gathered bytes per launch",
                gather_GBps=round(gather / (ms * 1e-3) / 1e9, 1) if ms > 0 else 0.0)
    # the whole analysis
    npb = float(t["n_problems"])
    f_eig = npb * (11.0 * k ** 3 + 4.0 * k * k)
    f_apply = npb * 2.0 * k * k * slabs_per_problem
    apply_peak = PEAK["fp64"] if (path == "plain" or k < 64) else peak
    b_comp = n_gp * (8.0 * ns * k + 16.0 * ns + 12.0) + n_obs * (k + 3) * 4.0
    t_cmp = flop / (peak * 1e12) + f_eig / (PEAK["fp64"] * 1e12) + f_apply / (apply_peak * 1e12)
    t_mem = b_comp / (HBM_GBS * 1e9)
    t_roof = max(t_cmp, t_mem)
    if t.get("total_ms", 0) > 0:
        roof["path_frac"] = round(t_roof / (t["total_ms"] * 1e-3), 5)
        roof["path"] = {"t_roof_ms": round(1e3 * t_roof, 3), "t_measured_ms": round(t["total_ms"], 3),
                        "F_prod": f"{flop:.4e}", "F_eig": f"{f_eig:.4e}",
                        "F_apply": f"{f_apply:.4e}", "B_comp": f"{b_comp:.4e}",
                        "bound": "compute" if t_cmp >= t_mem else "hbm",
                        "peaks_TFLOPs": {"prod": peak, "eig": PEAK["fp64"], "apply": apply_peak}}
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
        key = f"{workload}/{path}/{precision}/{tuning_key(tun)}"
        ent = tr.get(key)
        if ent and ent.get("kernel") == kern:
            roof["traffic"] = ent["bytes_per_dispatch"]
            roof["traffic_note"] = (f"HBM bytes per dispatch from rocprofv3 --pmc FETCH_SIZE x "
                                    f"{FETCH_SCALE} + WRITE_SIZE x {WRITE_SCALE} "
                                    f"({ent['source']}); launches of {ent['problems_per_dispatch']} "
                                    f"problems under the same launch choice ({tuning_key(tun)})")
        else:
            roof["traffic_note"] = f"no counter record for {key}"
    except (OSError, ValueError):
        pass
    return roof


def summarize_pmc(fetch_csv, write_csv, workload, path, precision, problems, tun=None,
                  sweep=None):
    """Per-dispatch HBM traffic of the product kernel from two rocprofv3
    --pmc runs, merged into profiles/traffic.json.  With a sweep (the
    --product-sweep list the runs made) the product dispatches, in dispatch
    order, are one per launch choice of the list; else they are averaged
    under one launch choice."""
    import csv
    tuns = sweep_tunings(sweep) if sweep else [tun]
    kerns = [product_kernel(path, precision, t) for t in tuns]

    def per_dispatch(fn, counter):
        rows = [r for r in csv.DictReader(open(fn))
                if r["Counter_Name"] == counter and any(kn in r["Kernel_Name"] for kn in kerns)]
        rows.sort(key=lambda r: int(r.get("Dispatch_Id", 0)))
        if not rows:
            raise SystemExit(f"no {counter} rows for {sorted(set(kerns))} in {fn}")
        return [(float(r["Counter_Value"]), r["Kernel_Name"]) for r in rows]
    fr = per_dispatch(fetch_csv, "FETCH_SIZE")
    wr = per_dispatch(write_csv, "WRITE_SIZE")
    src = (f"{os.path.basename(os.path.dirname(fetch_csv))}, "
           f"{os.path.basename(os.path.dirname(write_csv))}")
    if sweep:
        if len(fr) != len(tuns) or len(wr) != len(tuns):
            raise SystemExit(f"{len(fr)}/{len(wr)} product dispatches for {len(tuns)} launch choices")
        groups = [([f], [w]) for f, w in zip(fr, wr)]
    else:
        groups = [(fr, wr)]
    try:
        with open(TRAFFIC_FILE) as f:
            tr = json.load(f)
    except (OSError, ValueError):
        tr = {}
    for t, kern, (fs, ws) in zip(tuns, kerns, groups):
        if any(kern not in n for _, n in fs + ws):
            raise SystemExit(f"dispatch of another kernel where {kern} ({tuning_key(t)}) ran")
        fk = sum(v for v, _ in fs) / len(fs)
        wk = sum(v for v, _ in ws) / len(ws)
        ent = dict(kernel=kern, fetch_kib=fk, write_kib=wk, dispatches=[len(fs), len(ws)],
                   bytes_per_dispatch=round(1024.0 * (fk * FETCH_SCALE + wk * WRITE_SCALE)),
                   problems_per_dispatch=problems, source=src)
        tr[f"{workload}/{path}/{precision}/{tuning_key(t)}"] = ent
        print(tuning_key(t), json.dumps(ent))
    with open(TRAFFIC_FILE, "w") as f:
        json.dump(tr, f, indent=1, sort_keys=True)


def host_cpu():
    """Host cores as the OS reports them: nproc (online CPUs), the CPUs this
    process may run on, and the model name."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"nproc": os.cpu_count(), "affinity": aff, "model": model}


def cpu_threads():
    """Worker threads of the CPU legs: every CPU this process may use, at
    most 16 (one GPU's share of the host on the GPU machines)."""
    return max(1, min(16, host_cpu()["affinity"] or 1))


def cpu_baseline(prob, solver, seconds, nthreads):
    """The reference's CPU path: its gridpoint loop in FP32 with the
    reference's order of operations in letkf_core_solve (sgemm products,
    a symmetric eigensolve, the Pa hdxb_rinv product, the per-slab apply,
    src/letkf_core.F90:92-141, src/letkf_solver.F90:203-379), on a bounded
    random sample of unmasked gridpoints, one worker per host thread with
    the reference's round-robin gridpoint ownership.  A pilot run sizes the
    sample (and warms up), then the best of three runs over it is reported
    with the share of the reference's timer phases (obs_search,
    letkf_core, letkf_trans).  The search-structure set-up is timed apart
    and left out, like the GPU's set_obs.  The FP64 oracle then analyses
    the same sample once: the GPU's error (and the FP32 path's own) is
    measured against it."""
    import oracle as O
    g = np.random.default_rng(1234)
    pool = np.nonzero(prob.mask == 0)[0]
    pilot = np.sort(g.choice(pool, min(256, pool.size), replace=False))
    r = O.analysis(prob, points=pilot, nthreads=nthreads, f32=True)
    per = r["loop_s"] / pilot.size
    n = int(min(pool.size, max(256, seconds / 3.0 / max(per, 1e-7))))
    pts = np.sort(g.choice(pool, n, replace=False))
    runs = [O.analysis(prob, points=pts, nthreads=nthreads, f32=True) for _ in range(3)]
    ref32 = min(runs, key=lambda x: x["loop_s"])
    ph = np.array([ref32["obs_search_s"], ref32["letkf_core_s"], ref32["letkf_trans_s"]])
    ph = ph / max(ph.sum(), 1e-12)
    hc = host_cpu()
    orc = O.analysis(prob, points=pts, nthreads=nthreads)

    def rel(a, b):
        den = np.abs(b).max()
        return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
    base = {"value": round(n / ref32["loop_s"], 2), "unit": "gridpoints/s", "cores": nthreads,
            "kind": "port",
            "sample": f"{n} random unmasked gridpoints of the same analysis; the reference's "
                      f"loop in FP32 in its order of operations (letkf_core_solve: sgemm, "
                      f"FP32 Householder + QL eigensolve, Pa hdxb_rinv; per-slab apply), "
                      f"best of 3 gridpoint loops {ref32['loop_s']:.2f} s after a warm-up "
                      f"(+ {ref32['setup_s']:.1f} s search set-up, not counted)",
            "host": hc,
            "threads_note": f"{nthreads} worker threads: the CPUs this process may use "
                            f"({hc['affinity']} of nproc {hc['nproc']}), at most 16",
            "phase_share": {"obs_search": round(float(ph[0]), 3),
                            "letkf_core": round(float(ph[1]), 3),
                            "letkf_trans": round(float(ph[2]), 3)},
            "fp64_oracle_gridpoints_per_s": round(n / orc["loop_s"], 2),
            "fp32_reference_error_vs_fp64_oracle": {
                "mean_rel": rel(ref32["mean"], orc["mean"]),
                "sprd_rel": rel(ref32["sprd"], orc["sprd"]),
                "perturbations_rel": rel(ref32["state"] - ref32["mean"][None],
                                         orc["state"] - orc["mean"][None])}}
    st, mn, sp = solver.state_f64()
    d = solver.diag()
    sel_ok = bool(np.array_equal(d["lg_selhash"][pts, :].T, orc["lg_hash"]) and
                  np.array_equal(d["lg_obscount"][pts, :].T.astype(np.int64), orc["lg_m"]))
    inc_o = orc["mean"] - prob.mean[:, pts]
    pert_o = orc["state"] - orc["mean"][None]
    err = {"sample_gridpoints": n, "selection_bit_exact": sel_ok,
           "mean_rel": rel(mn[:, pts], orc["mean"]), "sprd_rel": rel(sp[:, pts], orc["sprd"]),
           "members_rel": rel(st[:, :, pts], orc["state"]),
           "perturbations_rel": rel(st[:, :, pts] - mn[None, :, pts], pert_o),
           "increment_rel": rel(mn[:, pts] - prob.mean[:, pts], inc_o),
           "mean_maxabs": float(np.abs(mn[:, pts] - orc["mean"]).max())}
    return base, err, (pts, orc)


def bench_cycles(args):
    """Cycled analysis on the 0.25-degree grid (k=50): ``args.cycles``
    cycles, each background the previous batched analysis (file order)
    moved one gridpoint east plus N(0, 0.1^2), each with fresh swath obs.
    The batched and the plain FP64 handle each take the same cycle inputs
    through set_state_members / set_obs_members, run, and return the
    analysis through get_state_members; the batched analysis drives the
    next cycle.  Per cycle: wall latency of the run and of the whole cycle
    (uploads, device preprocessing, run, egress), and the FP64 analysis of
    both paths against the oracle on a random sample of gridpoints.  Each
    handle is created, autotuned and warmed up on cycle 0, outside the
    per-cycle timings."""
    import oracle as O
    from letkf_amd import synth
    from letkf_amd.solver import LetkfGpuSolver

    nt = cpu_threads()
    t0 = time.time()
    prob = synth.config3()
    mem = np.asfortranarray((prob.state + prob.mean[None]).transpose(2, 1, 0), np.float32)
    log(f"[cycles] cycle 0 inputs ({time.time() - t0:.1f} s)")
    truth = None
    paths = (("batched", "batched", args.precision), ("plain", "plain", "fp64"))
    solvers = {}
    rows = []
    g = np.random.default_rng(4321)
    pool = np.nonzero(prob.mask == 0)[0]

    def rel(a, b):
        den = np.abs(b).max()
        return float(np.abs(a - b).max() / (den if den > 0 else 1.0))

    for c in range(args.cycles):
        if c > 0:
            t0 = time.time()
            prob = synth.config5_cycle(prob, mem, c)
            log(f"[cycles] cycle {c} inputs: {prob.nobs} obs ({time.time() - t0:.1f} s)")
        truth = prob.meta["truth"].ravel()
        qc = np.zeros(prob.nobs, np.int32)
        hxm = np.asfortranarray((prob.hx + prob.hxmean[None]).astype(np.float32))
        # the oracle's inputs as the device's ingest makes them (FP32, member order)
        ost, omn, osp = O.state_from_members(mem)
        _, ohx, ohm = O.obs_preprocess(hxm, qc)
        oprob = dataclasses.replace(prob, state=ost, mean=omn, sprd=osp, hx=ohx, hxmean=ohm)
        pts = np.sort(g.choice(pool, min(args.cycle_sample, pool.size), replace=False))
        orc = O.analysis(oprob, points=pts, nthreads=nt)
        bg = omn[0]
        rec = {"cycle": c, "n_obs": int(prob.nobs),
               "bg_rmse": float(np.sqrt(np.mean((bg[pool] - truth[pool]) ** 2)))}
        for name, path, prec in paths:
            s = solvers.get(name)
            if s is None:
                # handle creation and one untimed warm-up run on cycle 0
                s = LetkfGpuSolver(prob, precision=prec, path=path, members=mem,
                                   obs_members=(qc, hxm))
                if args.no_autotune:
                    s.run()
                else:
                    # tuned once on cycle 0, kept for every cycle
                    log(f"[cycles] {name} autotune: {s.autotune(repeats=1)}")
                if not args.no_graph:
                    s.run_graph()      # eager run + capture of the analysis
                solvers[name] = s
            ta = time.perf_counter()
            s.cycle(prob, members=mem, obs_members=(qc, hxm))
            tb = time.perf_counter()
            if args.no_graph:
                s.run()
            else:
                s.run_graph()          # captured on cycle 0, replayed after
            tc = time.perf_counter()
            replayed = s.timing()["search_ms"] == 0.0
            amem, amn, asp = s.state_members()
            td = time.perf_counter()
            st, mn, sp = s.state_f64()
            d = s.diag()
            sel_ok = bool(np.array_equal(d["lg_selhash"][pts, :].T, orc["lg_hash"]))
            rec[name] = {"input_ms": round(1e3 * (tb - ta), 2), "run_ms": round(1e3 * (tc - tb), 2),
                         "graph_replay": bool(replayed and not args.no_graph),
                         "egress_ms": round(1e3 * (td - tc), 2),
                         "cycle_ms": round(1e3 * (td - ta), 2),
                         "mean_rel": rel(mn[:, pts], orc["mean"]),
                         "sprd_rel": rel(sp[:, pts], orc["sprd"]),
                         "selection_bit_exact": sel_ok,
                         "ana_rmse": float(np.sqrt(np.mean((amn[pool, 0] - truth[pool]) ** 2)))}
            if name == "batched":
                nxt = amem
        log(f"[cycles] {json.dumps(rec)}")
        rows.append(rec)
        mem = nxt
    for s in solvers.values():
        s.final()

    def pct(name, key):
        v = np.array([r[name][key] for r in rows])
        return {"p50": round(float(np.percentile(v, 50)), 2),
                "p99": round(float(np.percentile(v, 99)), 2)}
    summ = {nm: {kk: pct(nm, kk) for kk in ("run_ms", "cycle_ms", "input_ms", "egress_ms")}
            for nm, _, _ in paths}
    worst = {nm: {"mean_rel": max(r[nm]["mean_rel"] for r in rows),
                  "sprd_rel": max(r[nm]["sprd_rel"] for r in rows),
                  "selection_bit_exact": all(r[nm]["selection_bit_exact"] for r in rows)}
             for nm, _, _ in paths}
    replays = {nm: sum(1 for r in rows if r[nm].get("graph_replay")) for nm, _, _ in paths}
    out = {"metric": "per-cycle analysis latency, cycled 0.25deg k=50 (p50/p99 over cycles)",
           "value": summ["batched"]["run_ms"]["p50"], "unit": "ms", "n_gpus": 1,
           "higher_is_better": False, "cycles": args.cycles,
           "dtype": args.precision, "data": "synthetic (seeded Philox)",
           "config": {"workload": "C5: 10-cycle 0.25deg 1440x720 surface, k=50, 1-2M fresh "
                                  "swath obs per cycle, background = analysis shifted 1 gp "
                                  "east + N(0, 0.1^2)", "n_gp": prob.nij, "k": prob.k,
                      "precision": args.precision, "parallelism": "single GPU"},
           "latency_ms": summ, "error_vs_fp64_oracle_worst": worst,
           "graph": {"captured_once_replayed_cycles": replays,
                     "note": "letkf_gpu_run_graph: the analysis captured on cycle 0 and "
                             "replayed on every cycle's new inputs" if not args.no_graph
                             else "eager runs (--no-graph)"},
           "oracle_sample_per_cycle": int(min(args.cycle_sample, pool.size)),
           "cycles_detail": rows}
    print(json.dumps(out), flush=True)


def bench_c4(args):
    """The 0.5-degree 720x360 x 75-level T/S/u/v analysis (ns = 300, k = 100,
    31.1 GB FP32 state, 1M obs): the batched path timed at args.precision,
    the plain FP64 path timed beside it, and the analysis mean and spread of
    every precision (FP64 plain, FP32, bf16x3, BF16) against the FP64
    oracle on a random sample of unmasked gridpoints.  One handle at a time
    (each holds the state, the FP64 analysis and the chunk buffers, about
    105 GB of HBM)."""
    import oracle as O
    from letkf_amd import synth
    from letkf_amd.solver import LetkfGpuSolver

    nt = cpu_threads()
    t0 = time.time()
    prob = synth.config4()
    log(f"[c4] inputs: nij={prob.nij} ns={prob.ns} k={prob.k} nobs={prob.nobs} "
        f"({time.time() - t0:.1f} s)")
    g = np.random.default_rng(4444)
    pool = np.nonzero(prob.mask == 0)[0]
    pts = np.sort(g.choice(pool, min(args.cycle_sample, pool.size), replace=False))
    if args.no_cpu:
        pts = pts[:0]
    t0 = time.time()
    orc = O.analysis(prob, points=pts, nthreads=nt)
    log(f"[c4] oracle on {pts.size} gridpoints: {time.time() - t0:.1f} s")

    def rel(a, b):
        if b.size == 0:
            return None
        den = np.abs(b).max()
        return float(np.abs(a - b).max() / (den if den > 0 else 1.0))

    def run(prec, path, steps, tune=False):
        tu = time.perf_counter()
        s = LetkfGpuSolver(prob, precision=prec, path=path)
        up = time.perf_counter() - tu
        try:
            tuning = None
            if tune and args.tun_fixed:
                s.set_tuning(**args.tun_fixed)
                tuning = s.tuning()
                log(f"[c4] tuning set: {tuning}")
            elif tune and not args.no_autotune:
                tuning = s.autotune(repeats=1)
                log(f"[c4] autotune: {tuning}")
            s.run()
            t1 = time.perf_counter()
            for _ in range(steps):
                s.run()
            dt = (time.perf_counter() - t1) / max(steps, 1)
            tm = s.timing()
            _, mn, sp = s.state_f64(members=False)
            d = s.diag()
        finally:
            s.final()
        err = {"mean_rel": rel(mn[:, pts], orc["mean"]), "sprd_rel": rel(sp[:, pts], orc["sprd"]),
               "mean_maxabs": float(np.abs(mn[:, pts] - orc["mean"]).max()) if pts.size else None,
               "selection_bit_exact": bool(np.array_equal(d["lg_selhash"][pts, :].T,
                                                          orc["lg_hash"]))}
        res = {"ms_per_step": round(1e3 * dt, 2), "gridpoints_per_s": round(prob.nij / dt, 1),
               "upload_s": round(up, 2),
               "phases_ms": {"search": round(tm["search_ms"], 2), "order": round(tm["order_ms"], 2),
                             "product": round(tm["product_ms"], 2),
                             "transform": round(tm["eigen_ms"], 2),
                             "finalize": round(tm["finalize_ms"], 2)},
               "work": {"solves": int(tm["n_problems"]), "kept_pairs": int(tm["n_pairs"])},
               "error_vs_fp64_oracle": err}
        if tuning:
            res["tuning"] = tuning
        log(f"[c4] {path} {prec}: {json.dumps(res)}")
        return res, tm

    main_res, tm = run(args.precision, "batched", args.steps, tune=True)
    out = {"metric": "analysis gridpoints/sec, 0.5deg 75-level T/S/u/v, k=100, batched "
                     "mixed-prec vs FP64 GPU; rel err vs FP64 oracle",
           "value": main_res["gridpoints_per_s"], "unit": "gridpoints/s", "n_gpus": 1,
           "steps": args.steps, "ms_per_step": main_res["ms_per_step"],
           "higher_is_better": True, "dtype": args.precision, "data": "synthetic (seeded Philox)",
           "config": {"workload": "C4: 0.5deg 720x360x75 T/S/u/v (ns=300), k=100, 1M obs "
                                  "(13,000 profiles x 75 + 25,000 surface), LOC_NOVRT 700->200 km "
                                  "cut-off, RTPS 0.6", "n_gp": prob.nij, "ns": prob.ns,
                      "k": prob.k, "n_obs": prob.nobs, "parallelism": "single GPU"},
           "roofline": roofline(tm, prob.k, "batched", args.precision, "c4",
                                main_res.get("tuning"), prob.ns,
                                prob.ns, prob.nij, prob.nobs),
           "batched": main_res, "oracle_sample": int(pts.size)}
    if args.no_fp64_compare:
        # the batched run alone (profiling)
        print(json.dumps(out), flush=True)
        return
    others = {}
    for prec in ("fp32", "bf16x3", "bf16x4", "bf16"):
        if prec != args.precision:
            others[prec], _ = run(prec, "batched", 1)
    plain, _ = run("fp64", "plain", max(1, min(args.steps, 2)))
    out["fp64_plain"] = plain
    out["speedup_vs_fp64_plain"] = round(main_res["gridpoints_per_s"] / plain["gridpoints_per_s"], 3)
    out["errors_by_precision"] = {p: r["error_vs_fp64_oracle"] for p, r in
                                  [("fp64_plain", plain), (args.precision, main_res),
                                   *others.items()]}
    out["ms_by_precision"] = {p: r["ms_per_step"] for p, r in others.items()}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--path", default="batched", choices=["plain", "batched"])
    ap.add_argument("--precision", default=None, choices=list(PEAK),
                    help="batched products: bf16x4 by default; bf16x3 for c4, the precision "
                         "that configuration names")
    ap.add_argument("--no-alternatives", action="store_true",
                    help="skip timing the other split-BF16 product (bf16x3 / bf16x4) beside it")
    ap.add_argument("--no-fp64-compare", action="store_true",
                    help="skip timing the plain FP64 path on the same workload")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ingest", action="store_true",
                    help="skip timing the file-order / raw-obs upload and download")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="record the product kernel's per-dispatch HBM traffic from two "
                         "rocprofv3 --pmc counter_collection.csv files in profiles/traffic.json")
    ap.add_argument("--problems-per-dispatch", type=int, default=0)
    ap.add_argument("--product-sweep", default=None, metavar="SPEC",
                    help="run one analysis per launch choice ('tuner': every product-side "
                         "choice the autotuner can reach; or order,bucket_bits,xcd,tile;...) "
                         "for per-dispatch counters; with --summarize-pmc, the same list "
                         "names the dispatches")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch choices instead of letkf_gpu_autotune")
    ap.add_argument("--tuning", default=None,
                    help="launch choices order,bucket_bits[,tridiag_reg,overlap,chunk_log2,xcd,tile] set "
                         "instead of autotuning (as a saved autotune result is reused)")
    ap.add_argument("--cycles", type=int, default=10, help="c5: number of analysis cycles")
    ap.add_argument("--no-graph", action="store_true",
                    help="c5: run each cycle eagerly instead of replaying the captured analysis")
    ap.add_argument("--cycle-sample", type=int, default=1000,
                    help="c5: gridpoints per cycle checked against the FP64 oracle; "
                         "c4: gridpoints checked")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "bf16x3" if args.config == "c4" else "bf16x4"
    tun_fixed = None
    if args.tuning:
        v = [int(x) for x in args.tuning.split(",")]
        v += [2, 0, 20, 4, 1][len(v) - 2:] if len(v) < 7 else []
        tun_fixed = dict(order=v[0], bucket_bits=v[1], tridiag_reg=v[2], overlap=v[3],
                         chunk_log2=v[4], xcd=v[5], tile=v[6])
    args.tun_fixed = tun_fixed
    if args.summarize_pmc:
        summarize_pmc(*args.summarize_pmc, args.config, args.path, args.precision,
                      args.problems_per_dispatch, tun_fixed, args.product_sweep)
        return
    if args.product_sweep:
        product_sweep(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; a multi-GPU "
                          "analysis would shard the gridpoints across GPUs"}))
        sys.exit(2)

    if args.config == "c5":
        bench_cycles(args)
        return
    if args.config == "c4":
        bench_c4(args)
        return

    from letkf_amd.solver import LetkfGpuSolver

    t0 = time.time()
    prob, workload = make_problem(args.config)
    log(f"[bench] problem {args.config}: nij={prob.nij} k={prob.k} nobs={prob.nobs} "
        f"({time.time() - t0:.1f} s)")
    tu = time.perf_counter()
    solver = LetkfGpuSolver(prob, precision=args.precision, path=args.path)
    upload_s = time.perf_counter() - tu
    log(f"[bench] inputs uploaded ({upload_s:.2f} s)")
    tuning = None
    if tun_fixed:
        solver.set_tuning(**tun_fixed)
        tuning = solver.tuning()
        log(f"[bench] tuning set: {tuning}")
    elif not args.no_autotune:
        tuning = solver.autotune(repeats=4)
        log(f"[bench] autotune: {tuning}")
    else:
        tuning = solver.tuning()
    for i in range(args.warmup):
        solver.run()
        log(f"[bench] warmup {i} done")
    ph_acc = None
    t_start = time.perf_counter()
    for i in range(args.steps):
        solver.run()
        t = solver.timing()
        if ph_acc is None:
            ph_acc = {kk: 0.0 for kk in t}
        for kk, v in t.items():
            ph_acc[kk] += v
    t_end = time.perf_counter()
    elapsed = t_end - t_start
    log(f"[bench] {args.steps} steps in {elapsed:.3f} s")
    tavg = {kk: v / args.steps for kk, v in ph_acc.items()}
    ms = 1e3 * elapsed / args.steps
    value = prob.nij * args.steps / elapsed
    G = prob.loc.maxgroups()
    slabs = prob.ns / G if G > 1 else prob.ns     # slabs per (gridpoint, group) problem
    roof = roofline(tavg, prob.k, args.path, args.precision, args.config, tuning, slabs,
                    prob.ns, prob.nij, prob.nobs)

    out = {
        "metric": "analysis gridpoints/sec at k=50, batched mixed-prec vs FP64 GPU; rel err vs FP64",
        "value": round(value, 1), "unit": "gridpoints/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": args.precision if args.precision != "fp64" else "f64",
        "data": "synthetic (seeded Philox)",
        "config": {"workload": workload, "n_gp": prob.nij, "k": prob.k, "n_obs": prob.nobs,
                   "path": args.path, "precision": args.precision,
                   "parallelism": "single GPU"},
        "roofline": roof,
        "tuning": tuning,
        "phases_ms": {"search": round(tavg["search_ms"], 3),
                      "order": round(tavg["order_ms"], 3),
                      "product": round(tavg["product_ms"], 3),
                      "transform": round(tavg["eigen_ms"], 3),
                      "finalize": round(tavg["finalize_ms"], 3)},
        "work": {"solves": int(tavg["n_problems"]), "kept_pairs": int(tavg["n_pairs"]),
                 "candidates": int(tavg["n_cand"]), "chunks": int(tavg["n_launch"])},
        "upload_s": round(upload_s, 3),
        "pcie_inclusive_gridpoints_per_s": round(prob.nij / (upload_s + ms * 1e-3), 1),
    }

    if not args.no_cpu:
        nt = cpu_threads()
        out["cpu_baseline"], out["error_vs_fp64_oracle"], sample = cpu_baseline(
            prob, solver, args.cpu_seconds, nt)
    solver.final()
    if args.path == "batched" and not args.no_alternatives:
        # the other split-BF16 product on the same inputs: time and error
        alt = {}
        for prec in ("bf16x3", "bf16x4"):
            if prec == args.precision:
                continue
            s3 = LetkfGpuSolver(prob, precision=prec, path="batched")
            if tun_fixed:
                s3.set_tuning(**tun_fixed)
            elif not args.no_autotune:
                s3.autotune(repeats=1)
            s3.run()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                s3.run()
            dt = (time.perf_counter() - t0) / args.steps
            rec = {"ms_per_step": round(1e3 * dt, 3), "gridpoints_per_s": round(prob.nij / dt, 1),
                   "product_ms": round(s3.timing()["product_ms"], 3)}
            if not args.no_cpu:
                pts, orc = sample
                st, mn, sp = s3.state_f64()

                def rel(a, b):
                    den = np.abs(b).max()
                    return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
                rec["error_vs_fp64_oracle"] = {
                    "mean_rel": rel(mn[:, pts], orc["mean"]), "sprd_rel": rel(sp[:, pts], orc["sprd"]),
                    "members_rel": rel(st[:, :, pts], orc["state"]),
                    "perturbations_rel": rel(st[:, :, pts] - mn[None, :, pts],
                                             orc["state"] - orc["mean"][None])}
            s3.final()
            alt[prec] = rec
            log(f"[bench] {prec}: {rec}")
        out["other_precisions"] = alt
    if not args.no_ingest:
        # the steps either side of the path on the device: members in file
        # order -> perturbations / mean / spread, members' hx + qc -> hx
        # mean / perturbations / compaction (5% of the obs flagged bad)
        g = np.random.default_rng(99)
        mem = np.asfortranarray((prob.state + prob.mean[None]).astype(np.float32).transpose(2, 1, 0))
        hxm = np.asfortranarray((prob.hx + prob.hxmean[None]).astype(np.float32))
        qc = (g.random(prob.nobs) < 0.05).astype(np.int32)
        tu = time.perf_counter()
        s2 = LetkfGpuSolver(prob, precision=args.precision, path=args.path, members=mem,
                            obs_members=(qc, hxm))
        up2 = time.perf_counter() - tu
        s2.run()
        tr = time.perf_counter()
        s2.state_members()
        down = time.perf_counter() - tr
        s2.final()
        out["ingest"] = {"upload_members_s": round(up2, 3), "upload_s": out["upload_s"],
                         "egress_members_s": round(down, 3),
                         "note": "host wall time incl. PCIe: set_state_members + "
                                 "set_obs_members (device-side mean/perturbation/spread, "
                                 "hx split, QC compaction of 5% bad obs) vs set_state + "
                                 "set_obs; get_state_members (file-order egress)"}
        log(f"[bench] ingest {out['ingest']}")
    if not args.no_fp64_compare and not (args.path == "plain" and args.precision == "fp64"):
        ref = LetkfGpuSolver(prob, precision="fp64", path="plain")
        if not args.no_autotune:
            ref.autotune(repeats=1)
        for _ in range(args.warmup):
            ref.run()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ref.run()
        dt = time.perf_counter() - t0
        rt = ref.timing()
        ref.final()
        out["fp64_plain"] = {"value": round(prob.nij * args.steps / dt, 1),
                             "ms_per_step": round(1e3 * dt / args.steps, 3),
                             "phases_ms": {"search": round(rt["search_ms"], 3),
                                           "product": round(rt["product_ms"], 3),
                                           "transform": round(rt["eigen_ms"], 3),
                                           "finalize": round(rt["finalize_ms"], 3)}}
        out["speedup_vs_fp64_plain"] = round(out["value"] / out["fp64_plain"]["value"], 3)
        log(f"[bench] fp64 plain: {out['fp64_plain']}")
        if prob.k < 64:
            # the plain FP64 path with the batched path's transform (the
            # rational A^(-1/2) instead of the QL eigensolve): the speed-up of
            # the precision alone
            os.environ["LETKF_GPU_RAT"] = "2"
            try:
                ref = LetkfGpuSolver(prob, precision="fp64", path="plain")
            finally:
                del os.environ["LETKF_GPU_RAT"]
            if not args.no_autotune:
                ref.autotune(repeats=1)
            for _ in range(args.warmup):
                ref.run()
            t0 = time.perf_counter()
            for _ in range(args.steps):
                ref.run()
            dt = time.perf_counter() - t0
            rt = ref.timing()
            ref.final()
            out["fp64_plain_rational"] = {
                "value": round(prob.nij * args.steps / dt, 1),
                "ms_per_step": round(1e3 * dt / args.steps, 3),
                "phases_ms": {"search": round(rt["search_ms"], 3),
                              "product": round(rt["product_ms"], 3),
                              "transform": round(rt["eigen_ms"], 3),
                              "finalize": round(rt["finalize_ms"], 3)}}
            out["speedup_vs_fp64_plain_same_transform"] = round(
                out["value"] / out["fp64_plain_rational"]["value"], 3)
            log(f"[bench] fp64 plain, rational transform: {out['fp64_plain_rational']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
