"""Benchmark workloads (the configurations the throughput is quoted on) and
the byte model of a step.

urban512 : 512x512x128, dx=1 m, periodic x/y, rough log-law bottom
           (z0=0.01 m), free-slip top, 64 cuboid obstacles (16x16 cells,
           heights 8-32 cells, seeded) on an 8x8 staggered array, inflow
           u=3 m/s + seeded perturbations, one passive scalar with a 1 unit/s
           point source at street level, dt=0.05 s.
flat512  : the same grid and flow without obstacles.
uniform64: 64^3 uniform case (dx=1 m, u=1 m/s +- 0.1, dt=0.05 s).
channel256: turbulent channel 256x128x128, 2 pi x pi x 2, tanh-stretched z
           (gamma 1.5), smooth-wall wall model at both walls, nu = 1e-3,
           pr_gradient_x = -1 (Re_tau = 1000), law-of-the-wall start with 5 %
           noise, dt = 2e-4 s, variable-z CD2 advection.
"""
import json
import os

import numpy as np

from . import capi, cases

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DEFAULT = "urban512"


class Workload:
    def __init__(self, name, build, fields):
        self.name = name
        self.build = build
        self.fields = fields


def _flat512():
    cfg = cases.uniform_config(512, 512, 128, dx=1.0, dt=0.05, nscalars=1, autotune=1)
    return cfg, [], [(101.3, 257.7, 0.5, 1.0, 0)]


def urban_obstacles(seed=3, n=512, pitch=64, foot=16):
    """64 cuboids on an 8x8 array with pitch 64 cells, odd rows shifted by
    32 cells in x, footprint 16x16 cells, heights U{8..32} cells."""
    rng = np.random.Generator(np.random.Philox(key=[seed, 11]))
    heights = rng.integers(8, 33, size=(8, 8))
    out = []
    for r in range(8):
        for c in range(8):
            x0 = c * pitch + 8 + (32 if r % 2 else 0)
            y0 = r * pitch + 24
            out.append((float(x0), float(x0 + foot), float(y0), float(y0 + foot), 0.0, float(heights[r, c])))
    return out


def _urban512():
    cfg, _, src = _flat512()
    return cfg, urban_obstacles(), [(12.5, 12.5, 0.5, 1.0, 0)]


def _fields_u3(cfg):
    return cases.initial_fields(cfg, u0=3.0, amp=0.3, seed=3)


def _uniform64():
    return cases.config1(64), [], []


def _channel256():
    return cases.channel_config(256, 128, 128, autotune=1), [], []


_WL = {
    "channel256": Workload("channel256", _channel256, lambda cfg: cases.channel_fields(cfg)),
    "flat512": Workload("flat512", _flat512, _fields_u3),
    "urban512": Workload("urban512", _urban512, _fields_u3),
    "uniform64": Workload("uniform64", _uniform64, lambda cfg: cases.initial_fields(cfg)),
}


def get(name):
    if name not in _WL:
        raise SystemExit(f"unknown workload {name}; choose from {sorted(_WL)}")
    return _WL[name]


def step_bytes_model(cfg, has_ib, precision):
    """Canonical algorithmic bytes per cell-step: each field touched once per
    kernel group per stage (K1 SGS, K2 momentum, K3 divergence, K4 five
    transform/solve passes, K5 projection, K6 scalar)."""
    if precision == "fp64":
        return 1168 if has_ib else 1144
    return 800 if has_ib else 788


def measured_traffic(kernel, workload):
    """HBM bytes per launch from the committed rocprofv3 PMC summary of this
    kernel and workload (profiles/traffic_<kernel>_<workload>.json), if any."""
    p = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload}.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")
