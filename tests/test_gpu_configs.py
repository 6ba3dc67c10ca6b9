"""The benchmark configurations at their stated sizes on the GPU.

C2  turbulent channel 256x128x128 (stretched z, smooth-wall model at both
    walls): FP64 bit-identical to the oracle after the initial projection and
    one step; 100 steps bit-identical at 128x64x64 (the oracle's time bounds
    the size: about 1 s per step at full size on the box's cores); the
    precision form for 10 steps at full size (FP64 divergence <= 1e-10 every
    step, finite fields, velocity within 1e-5 of the FP64 path).
C3  urban 512x512x128 with 64 cuboids and a point source: FP64 bit-identical
    to the oracle after the initial projection, 1 step and 10 steps; the precision
    form for 10 steps against the oracle (relative L2 velocity <= 1e-5,
    divergence <= 1e-10 every step) -- the north-star acceptance test.
C4  48 scenarios of the batched ensemble advancing together (one stream
    each): divergence <= 1e-10 in every case, and three cases (smallest,
    median, largest) bit-equal to the same case run alone.
C5  a serving stream of requests from 0.25 M to 8 M cells: every result
    equals the same request run alone, bit for bit.
Reference step: src/tsteps.f90:24-243.
"""
import math

import numpy as np
import pytest

from elmm_amd import capi, cases, serving, workloads
from helpers import compare_state, velocity_rel_l2
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def _state(s):
    return [s.download(c) for c in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR)]


def _finite(s, cfg):
    for c in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR):
        a = s.download(c)
        assert np.all(np.isfinite(a[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz])), c


# ------------------------------------------------------------------ C2
def _channel(nx, ny, nz):
    cfg = cases.channel_config(nx, ny, nz)
    case = capi.Case(cfg)
    return cfg, case, cases.channel_fields(cfg, seed=2)


def test_c2_channel256_fp64_matches_oracle(require_gpu):
    cfg, case, f = _channel(256, 128, 128)
    o, s = Oracle(cfg, case), capi.Solver(cfg, case)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg, f)
    o.init()
    s.init()
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc"])
    print(f"channel256 init: {err}", flush=True)
    assert max(err.values()) == 0.0, err
    do, ds = o.step(1), s.step(1)
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc"])
    print(f"channel256 1 step: {err}, max_div gpu {ds.max_div:.3e} oracle {do.max_div:.3e}", flush=True)
    assert max(err.values()) == 0.0, err
    assert ds.max_div <= max(do.max_div, 1e-15) * 1.5


def test_c2_channel_fp64_hundred_steps(require_gpu):
    cfg, case, f = _channel(128, 64, 64)
    o, s = Oracle(cfg, case), capi.Solver(cfg, case)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg, f)
    o.init()
    s.init()
    o.step(100)
    s.step(100)
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc"])
    print(f"channel 128x64x64, 100 steps: {err}", flush=True)
    assert max(err.values()) == 0.0, err


def test_c2_channel256_fp32_refined(require_gpu):
    cfg, case, f = _channel(256, 128, 128)
    s32 = capi.Solver(capi.with_precision(cfg, capi.FP32_REFINED), case)
    s64 = capi.Solver(cfg, case)
    for s in (s32, s64):
        cases.upload_initial(s, cfg, f)
        s.init()
    for k in range(10):
        d = s32.step(1)
        assert d.max_div <= 1e-10 and np.isfinite(d.stored_div), (k, d.max_div)
    s64.step(10)
    _finite(s32, cfg)
    l2 = velocity_rel_l2(s32, s64, cfg)
    print(f"channel256 precision form vs FP64 after 10 steps: rel L2 {l2:.3e}", flush=True)
    assert l2 <= 1e-5


# ------------------------------------------------------------------ C3
@pytest.fixture(scope="module")
def urban():
    wl = workloads.get("urban512")
    cfg, cub, src = wl.build()
    case = capi.Case(cfg, cuboids=cub, sources=src)
    fields = wl.fields(cfg)
    o = Oracle(cfg, case)
    cases.upload_initial(o, cfg, fields)
    o.init()
    st = {"cfg": cfg, "case": case, "fields": fields, "oracle": o, "steps": 0}
    yield st
    o.close()


def _urban_solver(st, precision):
    cfg = capi.with_precision(st["cfg"], precision)
    s = capi.Solver(cfg, st["case"])
    cases.upload_initial(s, cfg, st["fields"])
    s.init()
    return s


def test_c3_urban512_fp64_one_step_matches_oracle(require_gpu, urban):
    cfg, o = urban["cfg"], urban["oracle"]
    s = _urban_solver(urban, capi.FP64)
    err = compare_state(s, o, cfg)
    print(f"urban512 init: {err}", flush=True)
    assert max(err.values()) == 0.0, err
    assert urban["steps"] == 0
    do, ds = o.step(1), s.step(1)
    urban["steps"] = 1
    err = compare_state(s, o, cfg)
    print(f"urban512 1 step: {err}; max_div gpu {ds.max_div:.3e} oracle {do.max_div:.3e}", flush=True)
    assert max(err.values()) == 0.0, err
    s.close()


def test_c3_urban512_precision_form_ten_steps_vs_oracle(require_gpu, urban):
    """North-star acceptance: the FP32 + FP64-refined step on the
    512x512x128 immersed-boundary case, 10 steps: divergence <= 1e-10 every
    step and relative L2 velocity error <= 1e-5 against the FP64 oracle
    (FP64 GPU path checked against the same oracle state)."""
    cfg, o = urban["cfg"], urban["oracle"]
    o.step(10 - urban["steps"])
    urban["steps"] = 10
    s32 = _urban_solver(urban, capi.FP32_REFINED)
    for k in range(10):
        d = s32.step(1)
        assert d.max_div <= 1e-10, (k, d.max_div)
    _finite(s32, cfg)
    l2 = velocity_rel_l2(s32, o, cfg)
    err32 = compare_state(s32, o, cfg)
    s64 = _urban_solver(urban, capi.FP64)
    s64.step(10)
    err64 = compare_state(s64, o, cfg)
    print(f"urban512 10 steps: precision form rel L2 {l2:.3e}, per field {err32}; FP64 {err64}", flush=True)
    assert l2 <= 1e-5
    assert max(err64.values()) == 0.0, err64
    s32.close()
    s64.close()


# ------------------------------------------------------------------ C4
def test_c4_ensemble_48_cases(require_gpu):
    n_cases, steps, seed = 48, 5, 4
    rng = np.random.Generator(np.random.Philox(key=[seed, 1]))
    sizes = np.exp(rng.uniform(math.log(64 ** 3), math.log(192 * 192 * 96), n_cases))
    prepared = [serving.scenario(seed * 1000 + a, n, capi.FP32_REFINED) for a, n in enumerate(sizes)]
    built = []
    for cfg, cub, src, fields in prepared:
        case = capi.Case(cfg, cuboids=cub, sources=src)
        sv = capi.Solver(cfg, case)
        cases.upload_initial(sv, cfg, fields)
        sv.init()
        built.append((cfg, case, fields, sv))
    for _, _, _, sv in built:
        sv.step_async(steps)
    for _, _, _, sv in built:
        sv.wait()
    divs = [sv.step(1).max_div for _, _, _, sv in built]
    cells = [c.nx * c.ny * c.nz for c, _, _, _ in built]
    print(f"ensemble: {n_cases} cases, {min(cells)}..{max(cells)} cells, max divergence {max(divs):.3e}",
          flush=True)
    assert max(divs) <= 1e-10
    order = np.argsort(cells)
    for q in (order[0], order[len(order) // 2], order[-1]):
        cfg, case, fields, sv = built[q]
        alone = capi.Solver(cfg, case)
        cases.upload_initial(alone, cfg, fields)
        alone.init()
        alone.step(steps + 1)
        for a, b in zip(_state(sv), _state(alone)):
            assert np.array_equal(a, b), (q, cells[q])
        alone.close()
    for _, _, _, sv in built:
        sv.close()


# ------------------------------------------------------------------ C5
def test_c5_serving_stream_mix_equals_alone(require_gpu):
    rng = np.random.Generator(np.random.Philox(key=[5, 2]))
    n = 6
    cells = list(np.clip(np.exp(rng.normal(math.log(1.4e6), 0.8, n)), 0.25e6, 8e6))
    cells[0], cells[1] = 8e6, 0.25e6          # both ends of the stated range
    lengths = [int(x) for x in rng.integers(8, 25, n)]
    reqs = []
    for k in range(n):
        cfg, cub, src, fields = serving.scenario(5000 + k, cells[k], capi.FP32_REFINED)
        reqs.append((cfg, capi.Case(cfg, cuboids=cub, sources=src), fields, lengths[k]))
    srv = capi.Server(200e9, max_active=64, chunk_steps=4)
    ids = []
    for cfg, case, fields, nsteps in reqs:
        ids.append(srv.submit(cfg, case, fields, nsteps))
        srv.progress()                                    # arrivals while others run
    while True:
        a, q = srv.progress()
        if a == 0 and q == 0:
            break
    print("serving stream: cells " + ", ".join(str(c.nx * c.ny * c.nz) for c, _, _, _ in reqs), flush=True)
    for rid, (cfg, case, fields, nsteps) in zip(ids, reqs):
        st = srv.status(rid)
        assert st.state == 2 and st.steps_done == nsteps and np.isfinite(st.max_div), \
            (st.state, st.status, capi._lib.elmm_last_error().decode())
        U, V, W, Pr, S = srv.result(rid)
        s = capi.Solver(cfg, case)
        cases.upload_initial(s, cfg, fields)
        s.init()
        s.step(nsteps)
        for got, want in zip([U, V, W, Pr, S[0]], _state(s) + [s.download(capi.F_SCALAR)]):
            assert np.array_equal(got, want)
        s.close()
        srv.release(rid)
    srv.close()
