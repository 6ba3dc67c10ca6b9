"""Throughput benchmark of the MI355X LES step (one GPU).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                [--precision fp64] [--no-cpu-baseline]

Prints one JSON line: Mcell-steps/s of the resident-state step (inputs in HBM
before the timed region; the step as a caller runs it, with its replayed step
graphs), the roofline figures of the dominant kernel group (HIP-event
durations of every launch in a second timed run of the same steps, which
also gives `ms_per_step_with_kernel_events`) and the CPU baseline (the
oracle restatement of the reference step, timed on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-elmm_amd"))

from elmm_amd import capi, cases, serving, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0


# Canonical algorithmic bytes per interior cell for one RK stage and kernel
# group (each field a group must read or write, counted once; halos, masks and
# surface lists excluded), per precision form: (stage 1, stages 2-3).
CANONICAL = {
    "fp64": {
        "sgs": (32, 32),                 # R U,V,W; W Visc
        "momentum": (88, 112),           # R U,V,W,Visc,Pr,[Ustar x3]; W Ustar x3, U2 x3
        "scalar": (56, 64),              # R U,V,W,S,TDiff,[Sadv]; W Sadv,S   (per scalar)
        "poisson_x_fwd": (48, 48),       # divergence of U2,V2,W2 -> RHS (32) + first transform pass (16)
        "poisson_y": (16, 16),           # one transform pass (R + W)
        "poisson_z": (16, 16),           # tridiagonal solve pass (R + W)
        "poisson_x_inv": (16, 16),       # last transform pass (R + W)
        "projection": (72, 72),          # R U2,V2,W2,Phi,Pr; W U,V,W,Pr
    },
    "fp32r": {
        "sgs": (16, 16),
        "momentum": (44, 56),
        "scalar": (28, 32),
        "poisson_x_fwd": (20, 20),       # FP32 U2,V2,W2 -> FP64 spectral write (default FP64 solve)
        "poisson_y": (16, 16),
        "poisson_z": (16, 16),
        "poisson_x_inv": (16, 16),
        "projection": (36, 36),
    },
}
PRECISION = {"fp64": capi.FP64, "fp32r": capi.FP32_REFINED}


def algorithmic_bytes(group, cfg, steps, prec):
    """Algorithmic HBM bytes of one kernel group over the timed steps."""
    table = CANONICAL[prec]
    if group not in table:
        return None
    cells = cfg.nx * cfg.ny * cfg.nz
    b1, b23 = table[group]
    per_step = b1 + 2 * b23
    mult = cfg.nscalars if group == "scalar" else (2 if group == "poisson_y" else 1)
    return cells * per_step * mult * steps


def run_one(wl, cfg, case, fields, prec, steps, warmup):
    """Time `steps` steps of one precision form with the state resident in
    HBM; returns throughput, per-group kernel times and diagnostics.

    Two timed runs of `steps` steps each: the first as a caller runs the step
    (replayed step graphs, no per-kernel events) gives the throughput; the
    second records HIP events around every launch on the engine's stream
    (which makes the launches eager) and gives the per-kernel durations of the
    roofline and the diagnostics of its last step."""
    c = capi.with_precision(cfg, PRECISION[prec])
    t0 = time.perf_counter()
    solver = capi.Solver(c, case)
    cases.upload_initial(solver, c, fields)
    solver.init()
    setup_s = time.perf_counter() - t0
    solver.step(warmup, diag=False)
    t0 = time.perf_counter()
    solver.step(steps, diag=False)
    dt_s = time.perf_counter() - t0
    solver.timers(enable=True)
    solver.timers(reset=True)
    t0 = time.perf_counter()
    diag = solver.step(steps)
    dt_ev = time.perf_counter() - t0
    timers = solver.timers()
    solver.timers(enable=False)
    solver.close()
    cells = cfg.nx * cfg.ny * cfg.nz
    return {"value": cells * steps / dt_s / 1e6, "ms_per_step": dt_s * 1e3 / steps,
            "ms_per_step_with_kernel_events": dt_ev * 1e3 / steps, "timers": timers,
            "max_div": diag.max_div, "stored_div": diag.stored_div, "poisson_res": diag.poisson_res,
            "setup_s": setup_s}


def roofline_of(run, cfg, steps, prec, wl):
    dom_name, (dom_ms, dom_launches) = max(run["timers"].items(), key=lambda kv: kv[1][0])
    nbytes = algorithmic_bytes(dom_name, cfg, steps, prec)
    if nbytes is None:
        return None
    avg_s = dom_ms / 1e3 / dom_launches
    per_launch = nbytes / dom_launches
    achieved = per_launch / avg_s / 1e9
    traffic = workloads.measured_traffic(dom_name, f"{wl.name}_{prec}")
    return {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "bytes_per_launch": int(per_launch), "avg_launch_us": round(avg_s * 1e6, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=workloads.DEFAULT)
    ap.add_argument("--precision", default="both", choices=["both", "fp32r", "fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble", type=int, default=48, help="batched-ensemble cases (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=20)
    ap.add_argument("--serve", type=int, default=8, help="requests of the serving leg (0: skip)")
    ap.add_argument("--serve-steps-scale", type=float, default=0.1,
                    help="run lengths U{200..2000} scaled by this factor to bound the run")
    ap.add_argument("--cpu-threads", type=int, default=cpu_threads_default())
    ap.add_argument("--cpu-steps", type=int, default=5, help="timed oracle steps of the CPU baseline")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="BYTES",
                    help="only run the plain copy kernel on BYTES and print its bandwidth")
    args = ap.parse_args()
    if args.copy_probe:
        print(json.dumps({"copy_probe_bytes": args.copy_probe, "GBs": round(capi.copy_probe(args.copy_probe, 5), 1)}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the step has no multi-GPU path in "
              "this build", file=sys.stderr)
        sys.exit(2)

    wl = workloads.get(args.workload)
    cfg, cuboids, sources = wl.build()
    t0 = time.perf_counter()
    case = capi.Case(cfg, cuboids=cuboids, sources=sources)
    case_s = time.perf_counter() - t0
    fields = wl.fields(cfg)
    precs = ["fp64", "fp32r"] if args.precision == "both" else [args.precision]
    runs = {p: run_one(wl, cfg, case, fields, p, args.steps, args.warmup) for p in precs}
    prim = "fp32r" if "fp32r" in runs else "fp64"
    r = runs[prim]

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, cfg, case, fields, args.cpu_threads, steps=args.cpu_steps)

    cells = cfg.nx * cfg.ny * cfg.nz
    out = {
        "metric": "Mcell-steps/s, FP32+refine vs FP64 on 1 GPU; p50/p99 scenario latency",
        "value": round(r["value"], 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "ms_per_step_with_kernel_events": round(r["ms_per_step_with_kernel_events"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 stencils, f64 pressure solve" if prim == "fp32r" else "f64",
        "data": "synthetic (seeded perturbations, float32-representable)",
        "config": {"workload": wl.name, "grid": [cfg.nx, cfg.ny, cfg.nz], "cells": cells, "dt": cfg.dt,
                   "precision": prim, "obstacles": len(cuboids), "scalars": cfg.nscalars},
        "roofline": roofline_of(r, cfg, args.steps, prim, wl),
        "cpu_baseline": cpu,
        "max_div": r["max_div"],
        "stored_div": r["stored_div"],
        "poisson_res": r["poisson_res"],
        "case_build_s": round(case_s, 2),
        "hbm_copy_GBs": round(capi.copy_probe(1 << 30, 5), 1),
    }
    if "fp64" in runs and prim != "fp64":
        f = runs["fp64"]
        out["fp64"] = {"value": round(f["value"], 2), "ms_per_step": round(f["ms_per_step"], 4),
                       "ms_per_step_with_kernel_events": round(f["ms_per_step_with_kernel_events"], 4),
                       "max_div": f["max_div"], "roofline": roofline_of(f, cfg, args.steps, "fp64", wl),
                       "kernel_ms": {k: [round(v[0], 3), v[1]] for k, v in sorted(f["timers"].items())}}
        out["speedup_vs_fp64"] = round(r["value"] / f["value"], 3)
    out["kernel_ms"] = {k: [round(v[0], 3), v[1]] for k, v in sorted(r["timers"].items())}
    pr = capi.FP32_REFINED if prim == "fp32r" else capi.FP64
    if args.ensemble > 0:
        out["ensemble"] = serving.run_ensemble(n_cases=args.ensemble, steps=args.ensemble_steps, precision=pr)
    if args.serve > 0:
        sv = serving.run_serving(n_requests=args.serve, steps_scale=args.serve_steps_scale, precision=pr)
        sv["capacity"] = serving.run_capacity(n_requests=args.serve, steps_scale=args.serve_steps_scale,
                                              precision=pr)
        sv["serial_fp64"] = serving.run_serial(n_requests=args.serve, steps_scale=args.serve_steps_scale)
        sv["capacity_vs_serial_fp64"] = round(sv["capacity"]["value"] / sv["serial_fp64"]["value"], 3)
        out["serving"] = sv
    print(json.dumps(out))


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads_default():
    """Host cores this process may use: OMP_NUM_THREADS when set (the GPU
    box sets its CPU share there), else the CPUs of the process's affinity
    mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(wl, cfg, case, fields, threads, warmup=1, steps=5):
    """The oracle (C restatement of the reference step, OpenMP FP64) timed on
    the host cores: after the initial projection, `warmup` untimed steps,
    then `steps` timed steps (reduced-step sample of the workload), with the
    OpenMP threads bound to cores (OMP_PROC_BIND=close, OMP_PLACES=cores; set
    before the oracle library and its OpenMP runtime are loaded).

    Binding pins the calling thread to the first core when the first parallel
    region starts; the process's CPU mask is put back afterwards, since every
    thread started later (the engine's set-up threads, the server's workers)
    inherits the mask of the thread that starts it."""
    allowed = os.sched_getaffinity(0)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_py import Oracle
    try:
        o = Oracle(cfg, case)
        cases.upload_initial(o, cfg, fields)
        o.init()
        if warmup:
            o.step(warmup)
        t0 = time.perf_counter()
        o.step(steps)
        dt = time.perf_counter() - t0
        o.close()
    finally:
        os.sched_setaffinity(0, allowed)
        for k in ("OMP_PROC_BIND", "OMP_PLACES"):
            os.environ.pop(k, None)
    cells = cfg.nx * cfg.ny * cfg.nz
    return {"value": round(cells * steps / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"reduced-step: {steps} timed steps after {warmup} warm-up step(s) of {wl.name} "
                      f"({cfg.nx}x{cfg.ny}x{cfg.nz}) following the initial projection; oracle C/OpenMP FP64, "
                      f"{dt:.1f} s, {dt / steps:.2f} s per step",
            "nproc": os.cpu_count(), "cpus_allowed": len(allowed), "cpu_model": _cpu_model(),
            "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": "close", "OMP_PLACES": "cores"}}


if __name__ == "__main__":
    main()
