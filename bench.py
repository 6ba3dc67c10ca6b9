"""Throughput benchmark of the MI355X LES step (one GPU).

python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                [--precision fp64] [--no-cpu-baseline]

Prints one JSON line: Mcell-steps/s of the resident-state step (inputs in HBM
before the timed region), the roofline figures of the dominant kernel group
(HIP-event durations measured in the timed region) and the CPU baseline (the
oracle restatement of the reference step, timed on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-elmm_amd"))

from elmm_amd import capi, cases, workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0


# Canonical algorithmic bytes per interior cell for one RK stage and kernel
# group (each field a group must read or write, counted once; halos, masks and
# surface lists excluded).  FP64 element = 8 B.  (stage 1, stages 2-3)
CANONICAL_FP64 = {
    "sgs": (32, 32),                 # R U,V,W; W Visc
    "momentum": (88, 112),           # R U,V,W,Visc,Pr,[Ustar x3]; W Ustar x3, U2 x3
    "scalar": (56, 64),              # R U,V,W,S,TDiff,[Sadv]; W Sadv,S   (per scalar)
    "poisson_x_fwd": (48, 48),       # divergence of U2,V2,W2 -> RHS (32) + 1st transform pass (16)
    "poisson_y": (16, 16),           # one transform pass (R + W)
    "poisson_z": (16, 16),           # tridiagonal solve pass (R + W)
    "poisson_x_inv": (16, 16),       # last transform pass (R + W)
    "projection": (72, 72),          # R U2,V2,W2,Phi,Pr; W U,V,W,Pr
}


def algorithmic_bytes(group, cfg, steps):
    """Algorithmic HBM bytes of one kernel group over the timed steps."""
    if group not in CANONICAL_FP64:
        return None
    cells = cfg.nx * cfg.ny * cfg.nz
    b1, b23 = CANONICAL_FP64[group]
    per_step = b1 + 2 * b23
    mult = cfg.nscalars if group == "scalar" else (2 if group == "poisson_y" else 1)
    return cells * per_step * mult * steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=workloads.DEFAULT)
    ap.add_argument("--precision", default="fp64", choices=["fp64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16") or 16))
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the step has no multi-GPU path in "
              "this build", file=sys.stderr)
        sys.exit(2)

    wl = workloads.get(args.workload)
    cfg, cuboids, sources = wl.build()
    t0 = time.perf_counter()
    case = capi.Case(cfg, cuboids=cuboids, sources=sources)
    fields = wl.fields(cfg)
    solver = capi.Solver(cfg, case)
    cases.upload_initial(solver, cfg, fields)
    solver.init()
    setup_s = time.perf_counter() - t0

    solver.step(args.warmup, diag=False)
    solver.timers(enable=True)
    solver.timers(reset=True)
    t0 = time.perf_counter()
    diag = solver.step(args.steps)
    dt_s = time.perf_counter() - t0
    timers = solver.timers()
    solver.timers(enable=False)

    cells = cfg.nx * cfg.ny * cfg.nz
    value = cells * args.steps / dt_s / 1e6
    ms_per_step = dt_s * 1e3 / args.steps

    # dominant kernel group by summed event time
    dom_name, (dom_ms, dom_launches) = max(timers.items(), key=lambda kv: kv[1][0])
    nbytes = algorithmic_bytes(dom_name, cfg, args.steps)
    roofline = None
    if nbytes is not None:
        avg_s = dom_ms / 1e3 / dom_launches
        per_launch = nbytes / dom_launches
        achieved = per_launch / avg_s / 1e9
        traffic = workloads.measured_traffic(dom_name, wl.name)
        roofline = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "bytes_per_launch": int(per_launch), "avg_launch_us": round(avg_s * 1e6, 2)}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(wl, cfg, cuboids, sources, args.cpu_threads)

    step_model = workloads.step_bytes_model(cfg, has_ib=bool(cuboids), precision=args.precision)
    out = {
        "metric": "Mcell-steps/s, FP32+refine vs FP64 on 1 GPU; p50/p99 scenario latency",
        "value": round(value, 2),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded perturbations, float32-representable)",
        "config": {"workload": wl.name, "grid": [cfg.nx, cfg.ny, cfg.nz], "cells": cells, "dt": cfg.dt,
                   "precision": args.precision, "obstacles": len(cuboids), "scalars": cfg.nscalars},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "step_model": {"bytes_per_cell_step": step_model,
                       "effective_GBs": round(step_model * cells * args.steps / dt_s / 1e9, 1)},
        "max_div": diag.max_div,
        "setup_s": round(setup_s, 2),
        "kernel_ms": {k: [round(v[0], 3), v[1]] for k, v in sorted(timers.items())},
    }
    print(json.dumps(out))


def cpu_baseline(wl, cfg, cuboids, sources, threads):
    """The oracle (C restatement of the reference step, OpenMP) timed on a
    bounded sample: one step of the same case after its initial projection."""
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_py import Oracle
    case = capi.Case(cfg, cuboids=cuboids, sources=sources)
    o = Oracle(cfg, case)
    cases.upload_initial(o, cfg, wl.fields(cfg))
    o.init()
    nsteps = 1
    t0 = time.perf_counter()
    o.step(nsteps)
    dt = time.perf_counter() - t0
    cells = cfg.nx * cfg.ny * cfg.nz
    o.close()
    return {"value": round(cells * nsteps / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} step of {wl.name} ({cfg.nx}x{cfg.ny}x{cfg.nz}) after the initial projection, "
                      f"oracle C/OpenMP FP64, {dt:.1f} s"}


if __name__ == "__main__":
    main()
