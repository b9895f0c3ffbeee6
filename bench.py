#!/usr/bin/env python3
"""Benchmark: SPH particle-steps/s on the 3-D periodic Taylor-Green box, 128^3 particles.

One "step" is one full CheapSynchronousRK2 step of the whole box (neighbour
build, sum density + dt, mid-step update, grad-h, SPH + smoothing-scale
derivatives, compatible-energy full update, grad-h), with the state resident in
HBM.  Prints ONE JSON line (the contract documented in DESIGN.md).

    python bench.py [--gpus 1] [--steps 50] [--warmup 3] [--n1d 128] [--precision fp64|mixed]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from spheral_amd import abi, engine, problems  # noqa: E402

# FLOPs per unique pair per step by pass (3-D; +,-,*,/,sqrt = 1), and per particle-step.
FLOP_PAIR = {"sph": 472, "smoothing": 79, "density": 63, "gradh": 158, "dt_pair": 12, "energy": 34}
FLOP_PAIR_TOTAL = 818
FLOP_NODE = 450
BYTES_NODE_3D = 520           # compulsory HBM bytes per particle-step, FP64 3-D
PEAK_FP64 = 78.6e12           # MI355X FP64 vector/matrix peak, FLOP/s
PEAK_FP32 = 157.3e12
PEAK_HBM = 8.0e12


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(n1d, steps, precision):
    """The oracle (FP64 C++ restatement, OpenMP) on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    threads = min(os.cpu_count() or 1, int(os.environ.get("SPX_CPU_THREADS", "16")))
    O.set_threads(threads)
    cfg, st = problems.taylor_green_3d(n1d)
    n = len(st["m"])
    r = O.OracleRun(cfg, n)
    r.set_state(**st)
    r.initialize()
    r.step()                                       # warm-up step
    t0 = time.perf_counter()
    for _ in range(steps):
        r.step()
    dt = time.perf_counter() - t0
    return {"value": n*steps/dt, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (FP64 C++/OpenMP restatement) on the same Taylor-Green workload at {n1d}^3="
                      f"{n} particles, {steps} timed steps after 1 warm-up, {threads} threads, {cpu_model()}",
            "seconds": dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n1d", type=int, default=128)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--reorder", type=int, default=1)
    ap.add_argument("--cpu-n1d", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU sharding is design-only "
                                   "(see DESIGN.md), run with --gpus 1"}))
        sys.exit(2)
    prec = abi.SPX_PREC_FP64 if args.precision == "fp64" else abi.SPX_PREC_MIXED
    cfg, st = problems.taylor_green_3d(args.n1d, precision=prec)
    cfg.reorder_interval = args.reorder
    n = len(st["m"])
    print(f"[bench] building Taylor-Green {args.n1d}^3 = {n} particles", file=sys.stderr, flush=True)
    run = engine.make_run(cfg, st)
    run.initialize()
    for _ in range(args.warmup):
        run.step()
    npairs = run.num_pairs()
    # timed region: K steps, state resident in HBM; step() returns after the device finished
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.step()
    elapsed = time.perf_counter() - t0
    t, cycle, last_dt = run.time()
    # per-kernel-class device times (HIP events on the run's stream) over a few extra steps
    run.enable_timing()
    per_class = {}
    nprof = max(1, min(5, args.steps))
    for _ in range(nprof):
        run.step()
        for k, v in run.kernel_times().items():
            per_class[k] = per_class.get(k, 0.0) + v
    per_class = {k: v*args.steps/nprof for k, v in per_class.items()}
    ms_step = 1000.0*elapsed/args.steps
    value = n*args.steps/elapsed
    ppn = npairs/n
    cons = run.conservation()
    # dominant kernel: the SPH derivative pass
    sph_ms = per_class.get("sph", float("nan"))/args.steps
    sph_flops = FLOP_PAIR["sph"]*npairs + FLOP_PAIR["smoothing"]*npairs
    achieved = sph_flops/(sph_ms*1e-3)
    peak = PEAK_FP64 if prec == abi.SPX_PREC_FP64 else PEAK_FP32
    line = {
        "metric": "SPH particle-steps/s at 2M 3D",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == abi.SPX_PREC_FP64 else "f32-pair/f64-accumulate",
        "data": "synthetic",
        "config": {"workload": f"taylor_green_3d_{args.n1d}^3_wendlandC4_nperh2_cheapRK2", "particles": n,
                   "unique_pairs": npairs, "pairs_per_particle": ppn, "mean_neighbors": 2.0*ppn,
                   "kernel": "WendlandC4 TableKernel(1000)", "boundary": "periodic", "reorder_interval": args.reorder,
                   "t_end": t, "cycles": cycle},
        "roofline": {
            "kernel": "k_sph (SPH + SPHSmoothingScale pair pass)",
            "bound": "mfma",
            "pipe": "VALU FP64 (no MFMA: no dense contraction); peak = MI355X FP64 78.6 TFLOP/s",
            "achieved": achieved/1e12,
            "peak": peak/1e12,
            "unit": "TFLOP/s",
            "frac": achieved/peak,
            "traffic": None,
            "algorithmic_flops_per_launch": sph_flops,
            "avg_launch_ms": sph_ms,
        },
        "step_roofline": {
            "flop_per_particle_step": FLOP_PAIR_TOTAL*ppn + FLOP_NODE,
            "bytes_per_particle_step": BYTES_NODE_3D,
            "ceiling_particle_steps_per_s": 1.0/max((FLOP_PAIR_TOTAL*ppn + FLOP_NODE)/peak, BYTES_NODE_3D/PEAK_HBM),
        },
        "kernel_ms_per_step": {k: v/args.steps for k, v in per_class.items()},
        "conservation": {"momentum": cons[1:4].tolist(), "energy": float(cons[4] + cons[5])},
    }
    line["step_roofline"]["frac"] = value/line["step_roofline"]["ceiling_particle_steps_per_s"]
    if not args.no_cpu:
        print("[bench] timing the CPU baseline (oracle) on a bounded sample", file=sys.stderr, flush=True)
        cb = cpu_baseline(args.cpu_n1d, args.cpu_steps, prec)
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample")}
        line["speedup_vs_cpu"] = value/cb["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
