#!/usr/bin/env python3
"""Benchmark: SPH particle-steps/s on the 3-D periodic Taylor-Green box, 128^3 particles.

One "step" is one full CheapSynchronousRK2 step of the whole box (neighbour
build, sum density + dt, mid-step update, grad-h, SPH + smoothing-scale
derivatives, compatible-energy full update, grad-h), with the state resident in
HBM.  Prints ONE JSON line: metric, value, unit, ms_per_step, config, the dominant kernel's
roofline (HIP-event timing, committed rocprofv3 traffic) and the CPU baseline.

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


def sph_kernel(prec):
    """rocprofv3's name of the step's SPH pair pass on this workload (Monaghan-Gingold Q)."""
    return f"k_sph<3, false, {'double' if prec == abi.SPX_PREC_FP64 else 'float'}, false>"


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    passes over this workload, each scaled by the copy-kernel calibration recorded with it
    (None when no such measurement is committed)."""
    path = os.path.join(ROOT, "profiles", "traffic_taylor_green_128.json")
    try:
        with open(path) as fh:
            rec = json.load(fh)["kernels"][kernel]
        return rec["fetch_bytes"] + rec["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_share():
    """Host cores this process may use: the affinity mask, bounded by a cgroup CPU quota when one
    is set (cgroup v2 cpu.max), and the machine's nproc beside them."""
    nproc = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        allowed = nproc
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q, per = fh.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q)//int(per)))
    except (OSError, ValueError):
        pass
    share = min(allowed, quota) if quota else allowed
    omp = os.environ.get("OMP_NUM_THREADS")
    if omp and omp.isdigit() and int(omp) > 0:
        share = min(share, int(omp))
    return {"nproc": nproc, "affinity": allowed, "cgroup_quota": quota, "omp_num_threads": omp, "threads": share}


def cpu_baseline(n1d, steps, precision, integrator=abi.SPX_INTEGRATOR_CHEAP_SYNCHRONOUS_RK2, reps=3, warmup=2):
    """The oracle (FP64 C++ restatement, OpenMP) on the same Taylor-Green workload: `warmup`
    untimed steps, then `reps` timed segments of `steps` steps; the value is the median segment's
    rate.  Threads: every host core this process may use (see cpu_share)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    share = cpu_share()
    threads = share["threads"]
    O.set_threads(threads)
    cfg, st = problems.taylor_green_3d(n1d)
    cfg.integrator = integrator
    n = len(st["m"])
    r = O.OracleRun(cfg, n)
    r.set_state(**st)
    r.initialize()
    for _ in range(warmup):
        r.step()
    seg = []
    for _ in range(reps):
        t0 = time.perf_counter()
        for _ in range(steps):
            r.step()
        seg.append(time.perf_counter() - t0)
        print(f"[bench] cpu segment {len(seg)}/{reps}: {seg[-1]:.2f} s", file=sys.stderr, flush=True)
    med = float(np.median(seg))
    return {"value": n*steps/med, "unit": "particle-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle (FP64 C++/OpenMP restatement) on the same Taylor-Green workload at {n1d}^3="
                      f"{n} particles: {warmup} warm-up steps, then the median of {reps} timed segments of "
                      f"{steps} step(s); {threads} OpenMP threads = this process's CPU share "
                      f"(nproc {share['nproc']}, affinity {share['affinity']}, cgroup quota {share['cgroup_quota']}, "
                      f"OMP_NUM_THREADS {share['omp_num_threads']}); {cpu_model()}",
            "segments_s": seg, "share": share}


def ensemble(args):
    """Kelvin-Helmholtz ensemble: args.runs independent 2-D periodic runs, N log-uniform in
    [4096, 262144] (PCG64 seed args.seed), each to t = args.t_end.  Leg 1: all runs in one
    mixed-precision batch (spx_batch), admitted at once.  Leg 2: the same runs one after
    another as plain FP64 spx_runs.  runs/hr counts completed runs over wall time including
    state upload; accuracy is the L1 distance of each batched mixed-precision result from
    its sequential FP64 result (density, specific energy), relative to the FP64 L1 norm."""
    rng = np.random.Generator(np.random.PCG64(args.seed))
    targets = np.round(np.exp(rng.uniform(np.log(4096.0), np.log(262144.0), args.runs))).astype(int)
    cfg_m, _ = problems.kelvin_helmholtz_2d(int(targets[0]), precision=abi.SPX_PREC_MIXED)
    cfg_d, _ = problems.kelvin_helmholtz_2d(int(targets[0]), precision=abi.SPX_PREC_FP64)
    states = [problems.kelvin_helmholtz_2d(int(n))[1] for n in targets]
    sizes = np.array([len(s["m"]) for s in states])
    total = int(sizes.sum())
    print(f"[bench] ensemble: {args.runs} KH runs, {total} particles, t_end {args.t_end}", file=sys.stderr, flush=True)
    # leg 1: one mixed-precision batch
    t0 = time.perf_counter()
    b = engine.Batch(cfg_m, capacity=total, max_runs=args.runs)
    ids = [b.add_run(s, args.t_end) for s in states]
    t_admit = time.perf_counter() - t0
    if args.kernel_breakdown:
        b.enable_timing()            # per-class events: every step then runs eagerly
    bsteps = 0
    while b.step() > 0:
        bsteps += 1
    t_batch = time.perf_counter() - t0
    per_class = {k: v/max(bsteps, 1) for k, v in b.kernel_times().items()} if args.kernel_breakdown else None
    cycles = np.array([b.status(r)["cycle"] for r in ids])
    psteps = int((cycles*sizes).sum())
    print(f"[bench] batched: {t_batch:.2f} s, {bsteps} batch steps", file=sys.stderr, flush=True)
    line = {
        "metric": "ensemble runs/hr",
        "value": args.runs/t_batch*3600.0,
        "unit": "runs/hr",
        "n_gpus": 1,
        "steps": bsteps,
        "warmup": 0,
        "ms_per_step": 1000.0*t_batch/max(bsteps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair / f64 accumulate",
        "data": "synthetic",
        "config": {"workload": f"kelvin_helmholtz_2d_ensemble_{args.runs}runs_N4096-262144_t{args.t_end}",
                   "runs": args.runs, "particles": total, "seed": args.seed, "t_end": args.t_end,
                   "kernel": "BSpline TableKernel(1000), nPerh 2", "boundary": "periodic unit square"},
        "batched_seconds": t_batch,
        "admission_seconds": t_admit,
        "kernel_ms_per_batch_step": per_class,
        "step_stats": b.step_stats(),
        "neighbor_list_builds": int(sum(b.neighbor_builds(r) for r in ids)),
        "particle_steps": psteps,
        "particle_steps_per_s": psteps/t_batch,
    }
    if not args.no_sequential:
        t0 = time.perf_counter()
        l1 = {"mass density": 0.0, "specific thermal energy": 0.0}
        seq_time = 0.0
        for rid, st in zip(ids, states):
            ts = time.perf_counter()
            r = engine.make_run(cfg_d, st)
            r.initialize()
            r.advance(args.t_end, 10**7)
            ref = {f: r.field(f) for f in l1}
            r.close()
            seq_time += time.perf_counter() - ts
            for f in l1:
                got = b.field(rid, f)
                l1[f] = max(l1[f], float(np.abs(got - ref[f]).sum()/np.abs(ref[f]).sum()))
        line["sequential_fp64_seconds"] = seq_time
        line["sequential_fp64_runs_per_hr"] = args.runs/seq_time*3600.0
        line["speedup_vs_sequential_fp64"] = seq_time/t_batch
        line["max_l1_vs_fp64"] = l1
        print(f"[bench] sequential fp64: {seq_time:.2f} s", file=sys.stderr, flush=True)
    b.close()
    print(json.dumps(line), flush=True)


FLOP_TREE_INTERACTION = 27   # one monopole or point-point term of the walk (+,-,*,/,sqrt = 1)


def evrard(args):
    """Evrard collapse with TreeGravity: args.particles particles to t = args.t_end (default
    0.8) in mixed precision, then the same run in plain FP64; reports the mixed run's
    particle-steps/s and its L1 distance from the FP64 run (density, specific energy), the tree
    walk's interactions per walk and interactions/s (counting copy of the walk on the final
    tree), the per-class kernel times and the walk time from HIP events over 20 steps after the
    end time (per-class timing runs the midpoint gravity pass in order on the main stream; the
    timed run overlaps it with the SPH passes), and the oracle (FP64 CPU restatement) on the
    same problem as the CPU leg (one warm-up step, then the median of three single steps)."""
    t_end = args.t_end if args.t_end != 0.1 else 0.8
    out = {}
    for prec, name in ((abi.SPX_PREC_MIXED, "mixed"), (abi.SPX_PREC_FP64, "fp64")):
        cfg, st = problems.evrard_3d(args.particles, precision=prec)
        n = len(st["m"])
        r = engine.make_run(cfg, st)
        r.initialize()
        t0 = time.perf_counter()
        k, t = 0, 0.0
        while t < t_end*(1.0 - 1e-15) and k < args.max_steps:
            k1, t = r.advance(t_end, 50)
            k += k1
            print(f"[bench] evrard {name}: step {k} t {t:.4f}", file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        fields = {f: r.field(f) for f in ("mass density", "specific thermal energy")}
        builds = r.neighbor_builds()
        # the walk's throughput at the end state: interactions of one walk over the tree of the
        # last step, and the walk's time over 5 more steps
        cells, points = r.gravity_interactions()
        # per-class HIP events (which run the midpoint gravity pass in order on the main stream
        # instead of beside the SPH passes) over 20 steps from the end state
        r.enable_timing()
        for _ in range(20):
            r.step()
        kt = r.kernel_times()
        walk_ms = kt.get("gravity", float("nan"))/20
        out[name] = {"seconds": wall, "steps": k, "t": t, "particle_steps_per_s": n*k/wall, "builds": builds,
                     "kernel_ms_per_step": {c: v/20 for c, v in kt.items()},
                     "walk_interactions": {"cell": cells, "point": points}, "walk_ms_end": walk_ms,
                     "fields": fields}
        r.close()
    m, d = out["mixed"], out["fp64"]
    l1 = {f: float(np.abs(m["fields"][f] - d["fields"][f]).sum()/np.abs(d["fields"][f]).sum()) for f in m["fields"]}
    walk_ms = d["walk_ms_end"]
    inter = d["walk_interactions"]["cell"] + d["walk_interactions"]["point"]
    line = {
        "metric": "SPH + TreeGravity particle-steps/s",
        "value": m["particle_steps_per_s"], "unit": "particle-steps/s", "n_gpus": 1, "steps": m["steps"], "warmup": 0,
        "ms_per_step": 1000.0*m["seconds"]/max(m["steps"], 1), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 pair / f64 accumulate, f64 tree", "data": "synthetic",
        "config": {"workload": f"evrard_3d_{args.particles}_to_t{t_end}", "particles": args.particles, "t_end": t_end,
                   "opening": 0.5, "softening": 0.01, "kernel": "BSpline TableKernel(1000), nPerh 1.51"},
        "t_reached": m["t"], "fp64_particle_steps_per_s": d["particle_steps_per_s"], "fp64_steps": d["steps"],
        "fp64_t_reached": d["t"],
        "l1_mixed_vs_fp64": l1, "neighbor_list_builds": {"mixed": m["builds"], "fp64": d["builds"]},
        "kernel_ms_per_step": m["kernel_ms_per_step"], "fp64_kernel_ms_per_step": d["kernel_ms_per_step"],
        # the walk (FP64 in both modes), at the end state of the FP64 run: one walk per step
        "tree_walk": {"interactions_per_walk": inter, "cell_interactions": d["walk_interactions"]["cell"],
                      "point_interactions": d["walk_interactions"]["point"], "walk_ms": walk_ms,
                      "interactions_per_s": inter/(walk_ms*1e-3),
                      "flop_per_interaction": FLOP_TREE_INTERACTION,
                      "achieved_tflops": inter*FLOP_TREE_INTERACTION/(walk_ms*1e-3)/1e12,
                      "peak_tflops": PEAK_FP64/1e12,
                      "frac": inter*FLOP_TREE_INTERACTION/(walk_ms*1e-3)/PEAK_FP64},
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        share = cpu_share()
        O.set_threads(share["threads"])
        cfg, st = problems.evrard_3d(args.particles, precision=abi.SPX_PREC_FP64)
        n = len(st["m"])
        o = O.OracleRun(cfg, n)
        o.set_state(**st)
        o.initialize()
        o.step()
        seg = []
        for _ in range(3):
            t0 = time.perf_counter()
            o.step()
            seg.append(time.perf_counter() - t0)
            print(f"[bench] evrard cpu step {len(seg)}/3: {seg[-1]:.2f} s", file=sys.stderr, flush=True)
        med = float(np.median(seg))
        line["cpu_baseline"] = {"value": n/med, "unit": "particle-steps/s", "cores": share["threads"], "kind": "port",
                                "sample": f"oracle (FP64 C++/OpenMP restatement, TreeGravity walk per point in "
                                          f"parallel) on the same Evrard problem at {n} particles: one warm-up step "
                                          f"after start-up, then the median of 3 single steps; {share['threads']} "
                                          f"OpenMP threads; {cpu_model()}",
                                "segments_s": seg}
        line["speedup_vs_cpu_fp64"] = d["particle_steps_per_s"]/(n/med)
    print(json.dumps(line), flush=True)


SERVE_KINDS = ("sod2d", "sod3d", "gresho2d")
SERVE_T_END = {"sod2d": 0.15, "sod3d": 0.15, "gresho2d": 1.0}


def serve_problem(kind, n, precision):
    if kind == "gresho2d":
        return problems.gresho_2d(n, precision=precision)
    return problems.sod_tube(2 if kind == "sod2d" else 3, n, precision=precision)


def serve_trace(args):
    """Served stream: args.requests requests with Poisson arrivals at args.rate/s (PCG64 seed 99),
    kind uniform over {Sod 2-D, Sod 3-D, Gresho 2-D}, N log-uniform in [1000, 500000], fixed end
    times (Sod 0.15, Gresho 1.0).  Leg 1: continuous batching in mixed precision (spheral_amd.serve,
    one resident batch per kind, admission by particle budget and run slots).  Leg 2: the same trace
    served one request at a time in arrival order as plain FP64 spx_runs (each request's service
    time measured alone; queueing by the same arrival times)."""
    from spheral_amd import serve
    rng = np.random.Generator(np.random.PCG64(99))
    R = args.requests
    arrivals = np.cumsum(rng.exponential(1.0/args.rate, R))
    kinds = [SERVE_KINDS[k] for k in rng.integers(0, 3, R)]
    sizes = np.round(np.exp(rng.uniform(np.log(1000.0), np.log(500000.0), R))).astype(int)
    reqs = [(k, serve_problem(k, int(n), abi.SPX_PREC_MIXED)[1], SERVE_T_END[k]) for k, n in zip(kinds, sizes)]
    cfgs = {k: serve_problem(k, 4096, abi.SPX_PREC_MIXED)[0] for k in SERVE_KINDS}
    print(f"[bench] serve: {R} requests over {arrivals[-1]:.2f} s, {sum(len(r[1]['m']) for r in reqs)} particles",
          file=sys.stderr, flush=True)
    # one lane per kind, each with an HBM budget (bytes); the lane sizes its batch from its
    # configuration's measured bytes per particle (a 65536-particle probe request run to the
    # kind's end time)
    probes = {k: serve_problem(k, 65536, abi.SPX_PREC_MIXED)[1] for k in SERVE_KINDS}
    srv = serve.Server({k: (cfgs[k], args.lane_gib*2**30, 128, probes[k], SERVE_T_END[k]) for k in SERVE_KINDS})
    lanes = {k: {"hbm_budget_bytes": l.hbm_bytes, "bytes_per_particle": l.bytes_per_particle, "capacity": l.capacity}
             for k, l in srv.lanes.items()}
    print(f"[bench] lanes: {lanes}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    handles = []
    for (k, st, te), at in zip(reqs, arrivals):
        wait = t0 + at - time.perf_counter()
        if wait > 0:
            time.sleep(wait)
        handles.append(srv.submit(k, st, te))
    for i, h in enumerate(handles):
        h.event.wait()
        if h.error is not None:
            raise h.error
        print(f"[bench] served {i + 1}/{R}: {h.kind} n={h.n} tts {h.time_to_solution:.2f} s", file=sys.stderr, flush=True)
    makespan = time.perf_counter() - t0
    for k, l in srv.lanes.items():
        lanes[k]["peak_device_bytes"] = l.peak_bytes
        lanes[k]["peak_within_budget"] = l.peak_bytes <= l.hbm_bytes
        lanes[k]["batch_steps"] = l.steps
    srv.close()
    tts = [h.time_to_solution for h in handles]
    print(f"[bench] served: makespan {makespan:.2f} s, p50 {serve.percentile(tts, 50):.2f} s", file=sys.stderr, flush=True)
    # leg 2: sequential FP64 in arrival order
    cfgs64 = {k: serve_problem(k, 4096, abi.SPX_PREC_FP64)[0] for k in SERVE_KINDS}
    service, l1 = [], {"mass density": 0.0, "specific thermal energy": 0.0}
    for (k, st, te), h in zip(reqs, handles):
        ts = time.perf_counter()
        res, _ = serve.solve_alone(cfgs64[k], st, te)
        service.append(time.perf_counter() - ts)
        for f in l1:
            l1[f] = max(l1[f], float(np.abs(h.result[f] - res[f]).sum()/np.abs(res[f]).sum()))
        print(f"[bench] sequential {len(service)}/{R}", file=sys.stderr, flush=True)
    end, tts_seq = 0.0, []
    for at, sv in zip(arrivals, service):
        end = max(end, at) + sv
        tts_seq.append(end - at)
    psteps = int(sum(h.steps*h.n for h in handles))
    line = {
        "metric": "served runs/hr and time-to-solution",
        "value": R/makespan*3600.0, "unit": "runs/hr", "n_gpus": 1, "steps": R, "warmup": 0,
        "ms_per_step": 1000.0*makespan/R, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 pair / f64 accumulate", "data": "synthetic",
        "config": {"workload": f"serve_sod2d_sod3d_gresho2d_{R}req_poisson{args.rate}", "requests": R,
                   "rate_per_s": args.rate, "seed": 99, "N_range": [1000, 500000], "t_end": SERVE_T_END},
        "tts_p50_s": serve.percentile(tts, 50), "tts_p99_s": serve.percentile(tts, 99),
        "makespan_s": makespan, "particle_steps_per_s": psteps/makespan,
        "sequential_fp64_runs_per_hr": R/end*3600.0, "sequential_fp64_tts_p50_s": serve.percentile(tts_seq, 50),
        "sequential_fp64_tts_p99_s": serve.percentile(tts_seq, 99), "speedup_runs_per_hr": end/makespan,
        "max_l1_vs_fp64": l1, "lanes": lanes,
    }
    print(json.dumps(line), flush=True)


def accuracy(args):
    """Mixed-precision error next to the roundoff floor.  For each problem, three runs to the same
    end time: plain FP64, FP64 with every initial position moved by one ulp, and mixed precision.
    L1(perturbed FP64 - FP64) is what any roundoff-level difference (summation order, thread count,
    compiler contraction) grows to over the run; L1(mixed - FP64) is read against it."""
    cases = [
        ("taylor_green_3d_32^3", lambda p: problems.taylor_green_3d(32, precision=p), 0.02),
        ("kelvin_helmholtz_2d_16k", lambda p: problems.kelvin_helmholtz_2d(16384, precision=p), 0.1),
        ("sod_tube_2d_20k", lambda p: problems.sod_tube(2, 20000, precision=p), 0.15),
        ("gresho_2d_20k", lambda p: problems.gresho_2d(20000, precision=p), 1.0),
        ("evrard_3d_20k", lambda p: problems.evrard_3d(20000, precision=p), 0.2),
    ]
    fields = ("mass density", "specific thermal energy")
    rows = []
    for name, make, t_end in cases:
        out = {}
        for label, prec, perturb in (("fp64", abi.SPX_PREC_FP64, False), ("fp64_ulp", abi.SPX_PREC_FP64, True),
                                     ("mixed", abi.SPX_PREC_MIXED, False)):
            cfg, st = make(prec)
            if perturb:
                st = dict(st)
                st["x"] = np.nextafter(st["x"], np.inf)
            r = engine.make_run(cfg, st)
            r.initialize()
            steps, t = r.advance(t_end, 10**7)
            out[label] = {f: r.field(f) for f in fields}
            out[label]["steps"] = steps
            r.close()
        ref = out["fp64"]
        l1 = lambda a: {f: float(np.abs(a[f] - ref[f]).sum()/np.abs(ref[f]).sum()) for f in fields}
        rows.append({"case": name, "t_end": t_end, "steps": ref["steps"], "l1_mixed_vs_fp64": l1(out["mixed"]),
                     "l1_fp64_ulp_vs_fp64": l1(out["fp64_ulp"])})
        print(f"[bench] accuracy {name}: {rows[-1]}", file=sys.stderr, flush=True)
    print(json.dumps({"metric": "mixed-precision L1 vs FP64 and the roundoff floor", "cases": rows}), flush=True)


def rebuild_amortised(run, n, cycles=2, max_steps=600):
    """particle-steps/s over `cycles` whole candidate-list rebuild cycles (None when fewer rebuilds
    happen within max_steps)."""
    b = run.neighbor_builds()
    for _ in range(max_steps):
        run.step()
        if run.neighbor_builds() > b:
            break
    else:
        return None
    b = run.neighbor_builds()
    steps = 0
    t0 = time.perf_counter()
    while run.neighbor_builds() < b + cycles and steps < max_steps:
        run.step()
        steps += 1
    dt = time.perf_counter() - t0
    if run.neighbor_builds() < b + cycles:
        return None
    return {"value": n*steps/dt, "steps": steps, "builds": cycles, "ms_per_step": 1000.0*dt/steps}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n1d", type=int, default=128)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--cpu-n1d", type=int, default=128)
    ap.add_argument("--cpu-steps", type=int, default=1, help="steps per timed CPU segment (median of 3)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-rebuild-window", action="store_true",
                    help="skip the rate over whole candidate-rebuild cycles (profiling runs)")
    ap.add_argument("--no-eager-omega", action="store_true",
                    help="skip the rate with the end-of-step grad-h Omega computed every step (profiling runs)")
    ap.add_argument("--integrator", choices=["cheap", "synchronous"], default="cheap",
                    help="taylor_green: CheapSynchronousRK2 (the headline) or SynchronousRK2")
    ap.add_argument("--workload", choices=["taylor_green", "ensemble", "evrard", "serve", "accuracy"],
                    default="taylor_green")
    ap.add_argument("--requests", type=int, default=500, help="serve: number of requests in the trace")
    ap.add_argument("--lane-gib", type=float, default=24.0, help="serve: HBM budget of each lane (GiB)")
    ap.add_argument("--rate", type=float, default=20.0, help="serve: Poisson arrival rate (requests/s)")
    ap.add_argument("--particles", type=int, default=250_000, help="evrard: particle count")
    ap.add_argument("--max-steps", type=int, default=100_000, help="evrard: step limit")
    ap.add_argument("--runs", type=int, default=256, help="ensemble: number of Kelvin-Helmholtz runs")
    ap.add_argument("--t-end", type=float, default=0.1, help="ensemble: end time of every run")
    ap.add_argument("--seed", type=int, default=7, help="ensemble: seed of the run sizes")
    ap.add_argument("--no-sequential", action="store_true", help="ensemble: skip the sequential FP64 leg")
    ap.add_argument("--kernel-breakdown", action="store_true",
                    help="ensemble: per-kernel-class HIP-event times (the batch's steps then run eagerly)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU: sharding a run across GPUs is not "
                                   "built; run with --gpus 1"}))
        sys.exit(2)
    if args.workload == "ensemble":
        ensemble(args)
        return
    if args.workload == "evrard":
        evrard(args)
        return
    if args.workload == "serve":
        serve_trace(args)
        return
    if args.workload == "accuracy":
        accuracy(args)
        return
    prec = abi.SPX_PREC_FP64 if args.precision == "fp64" else abi.SPX_PREC_MIXED
    cfg, st = problems.taylor_green_3d(args.n1d, precision=prec)
    sync = args.integrator == "synchronous"
    if sync:
        cfg.integrator = abi.SPX_INTEGRATOR_SYNCHRONOUS_RK2
    n = len(st["m"])
    print(f"[bench] building Taylor-Green {args.n1d}^3 = {n} particles", file=sys.stderr, flush=True)
    run = engine.make_run(cfg, st)
    run.initialize()
    for _ in range(args.warmup):
        run.step()
    npairs = run.num_pairs()
    builds0 = run.neighbor_builds()
    # per-kernel-class device times: HIP events recorded on the run's stream around each
    # kernel class of every timed step (a reused event pool; summed natively)
    run.enable_timing()
    # timed region: K steps, state resident in HBM; step() returns after the device finished
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run.step()
    elapsed = time.perf_counter() - t0
    builds = run.neighbor_builds() - builds0
    t, cycle, last_dt = run.time()
    per_class = run.kernel_times()
    ms_step = 1000.0*elapsed/args.steps
    value = n*args.steps/elapsed
    ppn = npairs/n
    cons = run.conservation()
    # dominant kernel: the SPH derivative pass
    # SynchronousRK2 evaluates the derivatives twice per step (its "sph" class also holds the
    # start-of-step half-step velocity records, a small copy kernel)
    sph_ms = per_class.get("sph", float("nan"))/(args.steps*(2 if sync else 1))
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
        "dtype": "f64" if prec == abi.SPX_PREC_FP64 else "f32 pair / f64 accumulate",
        "data": "synthetic",
        "config": {"workload": f"taylor_green_3d_{args.n1d}^3_wendlandC4_nperh2_{'synchronousRK2' if sync else 'cheapRK2'}",
                   "particles": n,
                   "unique_pairs": npairs, "pairs_per_particle": ppn, "mean_neighbors": 2.0*ppn,
                   "kernel": "WendlandC4 TableKernel(1000)", "boundary": "periodic",
                   "t_end": t, "cycles": cycle},
        "roofline": {
            "kernel": "k_sph (SPH + SPHSmoothingScale pair pass)",
            "bound": "valu",
            "pipe": ("VALU FP64 (no MFMA: no dense contraction); peak = MI355X FP64 78.6 TFLOP/s"
                     if prec == abi.SPX_PREC_FP64 else
                     "VALU FP32 pair arithmetic, FP64 accumulation (no MFMA); peak = MI355X FP32 157.3 TFLOP/s"),
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
        # candidate-list rebuilds (re-sort + cell search) inside the timed steps; every other
        # step filters the exact neighbour sets from the lists
        "neighbor_list_builds": builds,
        "conservation": {"momentum": cons[1:4].tolist(), "energy": float(cons[4] + cons[5])},
    }
    line["step_roofline"]["frac"] = value/line["step_roofline"]["ceiling_particle_steps_per_s"]
    # graph replays (none here: timed steps run eagerly) and the measured neighbour-group sizes
    line["step_stats"] = run.step_stats()
    line["device_bytes"] = run.device_bytes()
    line["device_bytes_per_particle"] = run.device_bytes()/n
    # The end-of-step grad-h Omega (SPHBase::postStateUpdate) is computed when first read: the
    # next step never reads it, and the state gives the same bits either way.  The rate with it
    # computed at every step, as the reference's integrator does (timing on, as above):
    if not args.no_eager_omega:
        k2 = max(10, args.steps//4)
        run.set_end_of_step_omega(True)
        t0 = time.perf_counter()
        for _ in range(k2):
            run.step()
        line["value_eager_end_of_step_omega"] = n*k2/(time.perf_counter() - t0)
        run.set_end_of_step_omega(False)
    # The timed window may hold few candidate-list rebuilds (none at 20 steps); the rate over whole
    # rebuild cycles: from the step after one rebuild through the step of the second rebuild after it.
    amort = None if args.no_rebuild_window else rebuild_amortised(run, n)
    if amort is not None:
        line["value_rebuild_amortised"] = amort["value"]
        line["rebuild_window"] = {k: amort[k] for k in ("steps", "builds", "ms_per_step")}
    # attainable HBM rate on this box: plain streaming copy kernel, 1 GiB read + 1 GiB write per launch
    line["hbm_copy_gbps"] = engine.bandwidth_probe(1 << 30, 10)
    line["roofline"]["traffic"] = measured_traffic(sph_kernel(prec))
    if not args.no_cpu:
        print("[bench] timing the CPU baseline (oracle) on a bounded sample", file=sys.stderr, flush=True)
        cb = cpu_baseline(args.cpu_n1d, args.cpu_steps, prec, cfg.integrator)
        line["cpu_baseline"] = {k: cb[k] for k in ("value", "unit", "cores", "kind", "sample", "segments_s")}
        line["speedup_vs_cpu"] = value/cb["value"]
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
