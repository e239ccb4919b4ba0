#!/usr/bin/env python3
"""Throughput of the particle hot path on one MI355X.

A step is one full WCSPH step of the dam-break workload (neighbour rebuild:
cell keys, radix sort, reorder, cell table; density + Tait EOS gather;
momentum + viscosity gather fused with the FP64 symplectic-Euler update).
The workload is 1,000,000 fluid particles (dx = 5 mm, 100^3 block) plus the
799,320 particles of a 3-layer 1.6 x 0.67 x 1.2 m tank.  `value` is
particle-updates/s = fluid particles x steps / wall time of the timed steps,
inputs resident in HBM.  The other precision forms and the 4M mixed-precision
workload are reported under "forms".

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision f64|f32|mixed]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

PEAK_HBM_GBS = 8000.0
PREC = {"f64": 0, "f32": 1, "mixed": 2}
DTYPE = {"f64": "f64", "f32": "f32", "mixed": "f32 gathers + f64 integrator"}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def run_form(pa, scene_args, precision, steps, warmup, timing=True):
    d, arrs = pa.scene(pa.SCENE_DAM_BREAK, **scene_args)
    d.precision = PREC[precision]
    with pa.Context(d) as ctx:
        for k in ("x", "y", "z", "vx", "vy", "vz"):
            ctx.upload(k, arrs[k])
        ctx.step(warmup)
        ctx.sync()
        if timing:
            ctx.timing_enable(True)
        ctx.sync()
        t0 = time.perf_counter()
        ctx.step(steps)
        ctx.sync()
        t1 = time.perf_counter()
        kt = ctx.timing() if timing else []
        ctx.timing_enable(False)
    sec = t1 - t0
    nf = int(d.n_moving)
    res = {
        "value": nf * steps / sec,
        "ms_per_step": 1e3 * sec / steps,
        "n_fluid": nf,
        "n_total": int(d.n_particles),
        "kernels": kt,
    }
    return res


def roofline(kernels):
    if not kernels:
        return None
    dom = max(kernels, key=lambda k: k["total_ms"])
    avg_s = dom["total_ms"] / dom["launches"] / 1e3
    achieved = dom["bytes_per_launch"] / avg_s / 1e9
    return {
        "kernel": dom["kernel"],
        "bound": "hbm",
        "achieved": round(achieved, 2),
        "peak": PEAK_HBM_GBS,
        "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 5),
        "traffic": None,
        "avg_us": round(avg_s * 1e6, 3),
        "bytes_per_launch": dom["bytes_per_launch"],
    }


def cpu_baseline(pa, scene_args, budget_s=12.0):
    """FP64 oracle (OpenMP over particles) on the host cores, bounded sample of the same workload."""
    import oracle

    d, arrs = pa.scene(pa.SCENE_DAM_BREAK, **scene_args)
    st = oracle.State(oracle.params_from_desc(d), arrs)
    st.step(1)   # warm-up (first touch, thread pool)
    t0 = time.perf_counter()
    st.step(1)
    one = time.perf_counter() - t0
    k = max(1, min(50, int(budget_s / max(one, 1e-6))))
    t0 = time.perf_counter()
    st.step(k)
    sec = time.perf_counter() - t0
    cores = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    return {
        "value": float(d.n_moving) * k / sec,
        "unit": "particle-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"{k} FP64 steps of the same {d.n_particles}-particle dam break "
                  f"(after 2 warm-up steps), C oracle, OpenMP {cores} threads",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--precision", choices=list(PREC), default="f64")
    ap.add_argument("--no-forms", action="store_true", help="skip the other precision forms")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only "
                                   "(spatial slabs with halo exchange), so --gpus must be 1"}))
        sys.exit(2)

    import prestige_amd as pa

    pa.load_library()
    scene1 = dict(n_target=1_000_000, dx=0.005, seed=0)
    log(f"[bench] main form {args.precision}: {args.steps} steps after {args.warmup} warm-up")
    main_res = run_form(pa, scene1, args.precision, args.steps, args.warmup)
    forms = {}
    if not args.no_forms:
        for p in PREC:
            if p == args.precision:
                continue
            log(f"[bench] form {p} (1M)")
            r = run_form(pa, scene1, p, min(args.steps, 500), args.warmup)
            forms[p] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "roofline": roofline(r["kernels"])}
        log("[bench] form mixed (4M, dx = 3.15 mm)")
        r = run_form(pa, dict(n_target=4_000_000, dx=0.00315, seed=0), "mixed", min(args.steps, 300), args.warmup)
        forms["mixed_4M"] = {"value": r["value"], "ms_per_step": r["ms_per_step"], "n_fluid": r["n_fluid"],
                             "n_total": r["n_total"], "roofline": roofline(r["kernels"])}
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle)")
        cpu = cpu_baseline(pa, scene1)
    out = {
        "metric": "particle-updates/s/GPU",
        "value": main_res["value"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[args.precision],
        "data": "synthetic (deterministic dam-break lattice)",
        "config": {
            "workload": "WCSPH dam break, 1,000,000 fluid + 799,320 tank particles, dx 5 mm, "
                        "h = 1.2 dx, cubic spline, Tait gamma 7, alpha 0.1, dt = 0.25 h / c0",
            "n_fluid": main_res["n_fluid"],
            "n_total": main_res["n_total"],
            "precision": args.precision,
        },
        "roofline": roofline(main_res["kernels"]),
        "cpu_baseline": cpu,
        "kernels": [{k: (round(v, 4) if isinstance(v, float) else v) for k, v in kk.items()}
                    for kk in main_res["kernels"]],
        "forms": forms,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(main_res["value"] / cpu["value"], 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
