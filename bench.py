#!/usr/bin/env python3
"""Benchmark of the local analysis: analysis gridpoints/sec on one MI355X.

Workload (default): the 0.25-degree 1440x720 surface analysis with k = 50,
2M observations on 30 synthetic swath arcs and a 300 km cut-off radius
(seeded synthetic data, inputs resident in HBM before the timed region).
A step is one full analysis (letkf_gpu_run) over all 1,036,800 gridpoints.

Printed: one JSON line with the metric, the roofline of the dominant kernel
(from the library's HIP events on its own stream), the CPU baseline (the
FP64 oracle on a bounded sample of gridpoints, all host threads up to 16)
and the error of the GPU analysis against that oracle on the same sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "umd-letkf_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

PEAK = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0 / 3.0}
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


def roofline(t, k, precision, nsteps):
    """Algorithmic work of the product and eigen phases per step."""
    prod_flop = (2.0 * k * k + 2.0 * k) * t["n_pairs"]
    eig_flop = t["n_problems"] * (9.0 * k ** 3 + 2.0 * k ** 3 + 4.0 * k * k)
    phases = {}
    for name, flop, peak in (("product", prod_flop, PEAK[precision]),
                             ("eigen", eig_flop, PEAK["fp64"])):
        ms = t[f"{name}_ms"]
        ach = flop / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
        phases[name] = dict(ms=ms, launches=t["n_launch"],
                            avg_launch_ms=ms / max(t["n_launch"], 1),
                            flop=flop, achieved=ach, peak=peak, frac=ach / peak)
    phases["search"] = dict(ms=t["search_ms"], candidates=t["n_cand"])
    phases["finalize"] = dict(ms=t["finalize_ms"])
    dom = max(("product", "eigen"), key=lambda n: phases[n]["ms"])
    p = phases[dom]
    roof = dict(kernel=dom, bound="mfma" if dom == "product" else "fp64-valu",
                achieved=round(p["achieved"], 4), peak=p["peak"], unit="TFLOP/s",
                frac=round(p["frac"], 5), traffic=None,
                work=f"{p['flop']:.4e} flop per step over {p['launches']} launches")
    return roof, phases


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3")
    ap.add_argument("--path", default="batched", choices=["plain", "batched"])
    ap.add_argument("--precision", default="bf16x3", choices=list(PEAK))
    ap.add_argument("--no-fp64-compare", action="store_true",
                    help="skip timing the plain FP64 path on the same workload")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; "
                          "the analysis shards by gridpoint across GPUs (see DESIGN.md)"}))
        sys.exit(2)

    from letkf_amd.solver import LetkfGpuSolver

    t0 = time.time()
    prob, workload = make_problem(args.config)
    log(f"[bench] problem {args.config}: nij={prob.nij} k={prob.k} nobs={prob.nobs} "
        f"({time.time() - t0:.1f} s)")
    solver = LetkfGpuSolver(prob, precision=args.precision, path=args.path)
    log("[bench] inputs uploaded")
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
    roof, phases = roofline(tavg, prob.k, args.precision, args.steps)

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
        "phases_ms": {kk: round(v["ms"], 3) for kk, v in phases.items()},
        "work": {"solves": int(tavg["n_problems"]), "kept_pairs": int(tavg["n_pairs"]),
                 "candidates": int(tavg["n_cand"])},
    }

    if not args.no_cpu:
        import oracle as O
        nt = max(1, min(16, os.cpu_count() or 1))
        g = np.random.default_rng(1234)
        pool = np.nonzero(prob.mask == 0)[0]
        pilot = np.sort(g.choice(pool, min(64, pool.size), replace=False))
        tp = time.perf_counter()
        O.analysis(prob, points=pilot, nthreads=nt)
        per = (time.perf_counter() - tp) / pilot.size
        n = int(min(pool.size, max(64, args.cpu_seconds / max(per, 1e-6))))
        pts = np.sort(g.choice(pool, n, replace=False))
        tc = time.perf_counter()
        orc = O.analysis(prob, points=pts, nthreads=nt)
        dt = time.perf_counter() - tc
        out["cpu_baseline"] = {"value": round(n / dt, 2), "unit": "gridpoints/s", "cores": nt,
                               "kind": "port",
                               "sample": f"{n} random unmasked gridpoints of the same "
                                         f"analysis, FP64 oracle, {dt:.1f} s"}
        st, mn, sp = solver.state_f64()
        d = solver.diag()
        sel_ok = bool(np.array_equal(d["lg_selhash"][pts, :].T, orc["lg_hash"]) and
                      np.array_equal(d["lg_obscount"][pts, :].T.astype(np.int64), orc["lg_m"]))

        def rel(a, b):
            den = np.abs(b).max()
            return float(np.abs(a - b).max() / (den if den > 0 else 1.0))
        inc_o = orc["mean"] - prob.mean[:, pts]
        out["error_vs_fp64_oracle"] = {
            "sample_gridpoints": n, "selection_bit_exact": sel_ok,
            "mean_rel": rel(mn[:, pts], orc["mean"]), "sprd_rel": rel(sp[:, pts], orc["sprd"]),
            "members_rel": rel(st[:, :, pts], orc["state"]),
            "increment_rel": rel(mn[:, pts] - prob.mean[:, pts], inc_o),
            "mean_maxabs": float(np.abs(mn[:, pts] - orc["mean"]).max()),
        }
    solver.final()
    if not args.no_fp64_compare and not (args.path == "plain" and args.precision == "fp64"):
        ref = LetkfGpuSolver(prob, precision="fp64", path="plain")
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
                             "phases_ms": {kk: round(rt[f"{kk}_ms"], 3) for kk in
                                           ("search", "product", "eigen", "finalize")}}
        out["speedup_vs_fp64_plain"] = round(out["value"] / out["fp64_plain"]["value"], 3)
        log(f"[bench] fp64 plain: {out['fp64_plain']}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
