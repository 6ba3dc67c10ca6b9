"""Dispersion scenarios and the drivers of the batched-ensemble and serving
forms (both on top of the library's continuous-batching server,
elmm_server_*).

A scenario is a case of N cells (x/y rounded to multiples of 16 of the form
2^a 3^b for the spectral solver, z to a multiple of 8, Nx = Ny ~ 2 Nz), random cuboid
obstacles up to a plan-area density, a wind of random direction and speed,
and a point source at a free street-level cell; dt = 0.05 s.
"""
import math
import queue
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from . import capi, cases


def _rng(seed, salt):
    return np.random.Generator(np.random.Philox(key=[int(seed), int(salt)]))


# x/y extents the spectral solver takes (2^a 3^b) that are multiples of 16
XY_SIZES = sorted({16 * (2 ** a) * (3 ** b) for a in range(7) for b in range(3) if 16 * (2 ** a) * (3 ** b) <= 1024})


def grid_for_cells(n):
    """(nx, ny, nz) with nx = ny ~ 2 nz: nx the multiple of 16 of the form
    2^a 3^b nearest (in ratio) to (2n)^(1/3), nz a multiple of 8 in
    [8, 256] so that nx*ny*nz ~ n (64^3 -> 96x96x32, 192x192x96 -> itself)."""
    target = (2.0 * n) ** (1.0 / 3.0)
    nx = min((s for s in XY_SIZES if s >= 32), key=lambda s: abs(math.log(s / target)))
    nz = int(round(n / (nx * nx) / 8.0)) * 8
    return nx, nx, max(8, min(256, nz))


def scenario(seed, ncells, precision=capi.FP32_REFINED, density=None, speed=None, theta=None):
    """Config, cuboids, point sources and initial fields of one scenario."""
    rng = _rng(seed, 17)
    nx, ny, nz = grid_for_cells(ncells)
    cfg = capi.with_precision(cases.uniform_config(nx, ny, nz, dx=1.0, dt=0.05, nscalars=1, autotune=1), precision)
    density = rng.uniform(0.10, 0.30) if density is None else density
    # one cuboid per 16x16 tile, in random tile order, until the plan-area
    # density is reached; footprints 6-12 cells with a free margin
    tiles = [(a, b) for a in range(nx // 16) for b in range(ny // 16)]
    order = rng.permutation(len(tiles))
    cub, area = [], 0.0
    for t in order:
        if area >= density * nx * ny:
            break
        a, b = tiles[t]
        fx, fy = int(rng.integers(6, 13)), int(rng.integers(6, 13))
        x0 = 16 * a + 2 + int(rng.integers(0, 13 - fx + 1))
        y0 = 16 * b + 2 + int(rng.integers(0, 13 - fy + 1))
        h = int(rng.integers(4, max(5, nz // 3)))
        cub.append((float(x0), float(x0 + fx), float(y0), float(y0 + fy), 0.0, float(h)))
        area += fx * fy
    theta = rng.uniform(0, 2 * math.pi) if theta is None else theta
    speed = rng.uniform(1.0, 5.0) if speed is None else speed
    # source: a free cell at k = 1
    for _ in range(1000):
        sx, sy = rng.uniform(0, nx), rng.uniform(0, ny)
        if not any(c[0] - 1 <= sx <= c[1] + 1 and c[2] - 1 <= sy <= c[3] + 1 for c in cub):
            break
    src = [(float(sx), float(sy), 0.5, 1.0, 0)]
    U, V, W, Pr, S = cases.initial_fields(cfg, u0=speed * math.cos(theta), amp=0.05 * speed, seed=seed,
                                          v0=speed * math.sin(theta))
    return cfg, cub, src, (U, V, W, Pr, S)


def _prepare(seed, ncells, precision):
    """Scenario and case; the solid-body masks and lists are built on the GPU
    (elmm_case_build_device, the host builder's result byte for byte)."""
    cfg, cub, src, fields = scenario(seed, ncells, precision)
    case = capi.Case(cfg, cuboids=cub, sources=src, device=True)
    return cfg, case, fields


def _pct(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(math.ceil(p / 100.0 * len(xs))) - 1))
    return xs[k]


def run_ensemble(n_cases=48, steps=20, seed=4, precision=capi.FP32_REFINED, warmup=3, workers=8):
    """Batched ensemble: n_cases scenarios with cell counts log-uniform in
    [64^3, 192x192x96], each resident on its own stream, advance together.
    The case set-up (scenario and initial fields on the host, masks and
    lists on the GPU) and the solver set-up (create, upload, initial
    projection, `warmup` steps, which also capture the step graphs) are timed
    apart from the device-resident run of `steps` steps of every case; value
    = aggregate Mcell-steps/s of that run."""
    rng = _rng(seed, 1)
    sizes = np.exp(rng.uniform(math.log(64 ** 3), math.log(192 * 192 * 96), n_cases))
    t0 = time.perf_counter()
    with ThreadPoolExecutor(workers) as ex:
        prepared = list(ex.map(lambda a: _prepare(seed * 1000 + a[0], a[1], precision), enumerate(sizes)))
    t1 = time.perf_counter()
    solvers = []
    for cfg, case, fields in prepared:
        sv = capi.Solver(cfg, case)
        cases.upload_initial(sv, cfg, fields)
        sv.init()
        sv.step_async(warmup)
        solvers.append(sv)
    for sv in solvers:
        sv.wait()
    t2 = time.perf_counter()
    for sv in solvers:
        sv.step_async(steps)
    for sv in solvers:
        sv.wait()
    wall = time.perf_counter() - t2
    cells = [c.nx * c.ny * c.nz for c, _, _ in prepared]
    # after the timed run: one more step per case with the projection's FP64
    # divergence diagnostic (check_divergence, src/pressure.f90:884-989)
    divs = [sv.step(1).max_div for sv in solvers]
    for sv in solvers:
        sv.close()
    return {"cases": n_cases, "steps": steps, "cells_min": int(min(cells)), "cells_max": int(max(cells)),
            "cells_total": int(sum(cells)), "wall_s": round(wall, 4),
            "value": round(sum(cells) * steps / wall / 1e6, 2), "unit": "Mcell-steps/s",
            "case_setup_s": round(t1 - t0, 3), "solver_setup_s": round(t2 - t1, 3),
            "max_div": float(max(divs)),
            "note": "device-resident run of every case (one stream each, step graphs replayed); set-up timed apart"}


def run_serving(n_requests=8, rate=2.0, seed=5, steps_scale=0.1, precision=capi.FP32_REFINED, budget=200e9,
                chunk=10, workers=4):
    """Serving loop: Poisson arrivals (rate per second), cell counts
    lognormal (median 1.4 M, sigma 0.8) truncated to [0.25 M, 8 M], run
    lengths U{200..2000} x steps_scale.  A request's latency runs from its
    arrival (case set-up included) to the poll that sees it finished."""
    rng = _rng(seed, 2)
    gaps = rng.exponential(1.0 / rate, n_requests)
    arrivals = np.cumsum(gaps) - gaps[0]
    cells = np.clip(np.exp(rng.normal(math.log(1.4e6), 0.8, n_requests)), 0.25e6, 8e6)
    lengths = np.maximum(1, np.round(rng.integers(200, 2001, n_requests) * steps_scale)).astype(int)
    srv = capi.Server(budget, max_active=64, chunk_steps=chunk)
    ready = queue.Queue()
    t0 = time.perf_counter()

    def build(k):
        # set-up of request k starts at its arrival time
        delay = arrivals[k] - (time.perf_counter() - t0)
        if delay > 0:
            time.sleep(delay)
        ready.put((k, _prepare(seed * 1000 + k, cells[k], precision)))

    ex = ThreadPoolExecutor(workers)
    for k in range(n_requests):
        ex.submit(build, k)
    ids, done = {}, {}
    while len(done) < n_requests:
        while not ready.empty():
            k, (cfg, case, fields) = ready.get()
            ids[k] = (srv.submit(cfg, case, fields, int(lengths[k])), cfg)
        srv.progress()
        now = time.perf_counter() - t0
        for k, (rid, cfg) in list(ids.items()):
            if k not in done and srv.status(rid).state in (2, -1):
                done[k] = now - arrivals[k]
                srv.release(rid)
        time.sleep(0.0005)
    wall = time.perf_counter() - t0
    ex.shutdown()
    cell_steps, _ = srv.stats()
    srv.close()
    lat = list(done.values())
    return {"requests": n_requests, "rate_per_s": rate, "steps_scale": steps_scale,
            "cells_median": int(np.median(cells)), "steps_total": int(lengths.sum()),
            "p50_latency_s": round(_pct(lat, 50), 3), "p99_latency_s": round(_pct(lat, 99), 3),
            "wall_s": round(wall, 3), "value": round(cell_steps / wall / 1e6, 2), "unit": "Mcell-steps/s"}


def _requests(n_requests, seed, steps_scale):
    rng = _rng(seed, 2)
    gaps = rng.exponential(1.0 / 2.0, n_requests)
    cells = np.clip(np.exp(rng.normal(math.log(1.4e6), 0.8, n_requests)), 0.25e6, 8e6)
    lengths = np.maximum(1, np.round(rng.integers(200, 2001, n_requests) * steps_scale)).astype(int)
    del gaps
    return cells, lengths


def run_capacity(n_requests=8, seed=5, steps_scale=0.1, precision=capi.FP32_REFINED, budget=200e9, chunk=10,
                 workers=4):
    """The serving request stream at saturation: every request (cases built
    beforehand) is submitted at once; wall time from the first submit to the
    last completion, which includes each case's upload and initial
    projection.  Aggregate Mcell-steps/s."""
    cells, lengths = _requests(n_requests, seed, steps_scale)
    with ThreadPoolExecutor(workers) as ex:
        prepared = list(ex.map(lambda k: _prepare(seed * 1000 + k, cells[k], precision), range(n_requests)))
    srv = capi.Server(budget, max_active=64, chunk_steps=chunk)
    t0 = time.perf_counter()
    for k, (cfg, case, fields) in enumerate(prepared):
        srv.submit(cfg, case, fields, int(lengths[k]))
    while True:
        a, q = srv.progress()
        if a == 0 and q == 0:
            break
        time.sleep(0.0002)
    wall = time.perf_counter() - t0
    cell_steps, _ = srv.stats()
    srv.close()
    return {"value": round(cell_steps / wall / 1e6, 2), "unit": "Mcell-steps/s", "wall_s": round(wall, 3),
            "note": "all requests queued at once; upload and initial projection included"}


def run_serial(n_requests=8, seed=5, steps_scale=0.1, precision=capi.FP64):
    """The same request stream solved one after another on the same GPU (no
    overlap), with the accounting of run_capacity: wall time of upload,
    initial projection and steps of each request in turn."""
    cells, lengths = _requests(n_requests, seed, steps_scale)
    prepared = [_prepare(seed * 1000 + k, cells[k], precision) for k in range(n_requests)]
    total, busy = 0, 0.0
    for k, (cfg, case, fields) in enumerate(prepared):
        t = time.perf_counter()
        s = capi.Solver(cfg, case)
        cases.upload_initial(s, cfg, fields)
        s.init()
        s.step(int(lengths[k]), diag=False)
        busy += time.perf_counter() - t
        total += cfg.nx * cfg.ny * cfg.nz * int(lengths[k])
        s.close()
    return {"value": round(total / busy / 1e6, 2), "unit": "Mcell-steps/s",
            "precision": "fp64" if precision == capi.FP64 else "fp32r",
            "note": "one request at a time: upload, initial projection and steps"}
