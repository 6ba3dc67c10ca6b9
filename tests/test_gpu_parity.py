"""GPU parity: the HIP FP64 step against the CPU oracle through the C ABI.

Bars (FP64 form): <= 1e-12 relative (max-norm, per field) after one step and
<= 1e-9 after 100 steps; post-projection divergence no larger than the
oracle's.  On the uniform 64^3 case the two are bit-identical.
"""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def pair(cfg, scalar_amp=None, seed=1, cuboids=(), sources=(), u0=None):
    case = capi.Case(cfg, cuboids=cuboids, sources=sources)
    f = cases.initial_fields(cfg, seed=seed) if u0 is None else cases.initial_fields(cfg, u0=u0, amp=0.1 * u0,
                                                                                      seed=seed)
    if scalar_amp:
        n = (cfg.nx, cfg.ny, cfg.nz)
        f[4][0][3:3 + n[0], 3:3 + n[1], 3:3 + n[2]] = 1.0 + cases.perturbation(n, scalar_amp, seed, 7)
    o = Oracle(cfg, case)
    s = capi.Solver(cfg, case)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg, f)
    o.init()
    s.init()
    return o, s


def test_init_tail_matches(require_gpu):
    cfg = cases.config1(32)
    o, s = pair(cfg, scalar_amp=0.5)
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc", "TDiff", "S0"])
    assert max(err.values()) <= 1e-12, err


def test_one_step_and_ten_steps_32(require_gpu):
    cfg = cases.config1(32)
    o, s = pair(cfg, scalar_amp=0.5)
    do, ds = o.step(1), s.step(1)
    err = compare_state(s, o, cfg)
    assert max(err.values()) <= 1e-12, err
    assert ds.max_div <= max(do.max_div, 1e-15) * 1.5 and ds.max_div < 1e-13
    o.step(9)
    s.step(9)
    err = compare_state(s, o, cfg)
    assert max(err.values()) <= 1e-11, err


def test_config1_64_one_and_hundred_steps(require_gpu):
    """Uniform 64^3 case: 1 step <= 1e-12, 100 steps <= 1e-9 (observed: 0)."""
    cfg = cases.config1(64)
    o, s = pair(cfg, scalar_amp=0.5)
    o.step(1)
    s.step(1)
    err1 = compare_state(s, o, cfg)
    assert max(err1.values()) <= 1e-12, err1
    o.step(99)
    d = s.step(99)
    err100 = compare_state(s, o, cfg)
    assert max(err100.values()) <= 1e-9, err100
    assert d.max_div < 1e-13


def test_bit_identical_to_oracle(require_gpu):
    """Same operation order as the oracle: results are identical, not only
    close (regression guard for the restated arithmetic)."""
    cfg = cases.config1(32)
    o, s = pair(cfg, scalar_amp=0.5, seed=5)
    o.step(3)
    s.step(3)
    err = compare_state(s, o, cfg)
    assert max(err.values()) == 0.0, err


@pytest.mark.parametrize("dx", [0.09, 0.11])
def test_bit_identical_without_qualifying_reciprocals(require_gpu, dx):
    """Grid spacings whose rounded reciprocal fails the correction step's
    condition (|RN(1/d) d - 1| > 2^-54 for d = 0.09 and 0.11) select the
    kernel instantiations that divide: still the oracle's bits, with an
    obstacle (IB mass sources in the x pass)."""
    cfg = cases.uniform_config(32, 32, 16, dx=dx, dt=0.004, nscalars=1)
    cub = [(dx * 8, dx * 16, dx * 8, dx * 16, 0.0, dx * 6)]
    o, s = pair(cfg, seed=6, cuboids=cub, u0=3.0 * dx)
    o.step(3)
    s.step(3)
    err = compare_state(s, o, cfg)
    assert max(err.values()) == 0.0, err


def test_deterministic_across_runs(require_gpu):
    cfg = cases.config1(32)
    _, s1 = pair(cfg, scalar_amp=0.5)
    _, s2 = pair(cfg, scalar_amp=0.5)
    s1.step(5)
    s2.step(5)
    for f in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR, capi.F_SCALAR):
        assert np.array_equal(s1.download(f), s2.download(f))


def test_async_step_equals_sync(require_gpu):
    cfg = cases.config1(32)
    _, s1 = pair(cfg)
    _, s2 = pair(cfg)
    s1.step(4, diag=False)
    s2.step_async(4)
    s2.wait()
    assert np.array_equal(s1.download(capi.F_U), s2.download(capi.F_U))


def test_step_before_init_is_an_error(require_gpu):
    cfg = cases.config1(16)
    s = capi.Solver(cfg, capi.Case(cfg))
    with pytest.raises(capi.ElmmError) as e:
        s.step(1)
    assert e.value.code == -5


OBSTACLES = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (20.0, 26.0, 18.0, 30.0, 0.0, 10.0, 0.01)]


def test_immersed_boundary_case_parity(require_gpu):
    """Two cuboids (one smooth, one rough, one wrapping the periodic y edge),
    a street-level point source: IB forcing, IB mass sources, solid-body wall
    models, nullified points and the scalar wall faces, 1 and 10 steps."""
    cfg = cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05, nscalars=1)
    o, s = pair(cfg, seed=3, cuboids=OBSTACLES, sources=[(4.5, 4.5, 0.5, 1.0, 0)], u0=3.0)
    err = compare_state(s, o, cfg)
    assert max(err.values()) <= 1e-12, err
    do, ds = o.step(1), s.step(1)
    err = compare_state(s, o, cfg)
    assert max(err.values()) <= 1e-12, err
    assert ds.max_div <= max(do.max_div, 1e-15) * 1.5
    o.step(9)
    s.step(9)
    err = compare_state(s, o, cfg)
    assert max(err.values()) <= 1e-11, err


EDGE_OBSTACLES = [(0.0, 3.0, 10.0, 14.0, 0.0, 5.0), (60.0, 70.0, 20.0, 26.0, 0.0, 8.0),
                  (100.0, 110.0, 28.0, 32.0, 0.0, 6.0, 0.01)]


def test_obstacles_at_domain_edges_bit_identical(require_gpu):
    """Cuboids against the periodic x and y edges (IB points at i = 1..3 and
    j = 30..32, sources of the wrapped ghost cells) and one across the
    64-cell boundary of the x pass's Q row segments (nx = 128): the ghost
    cells rewritten after IBMomentum from the set-up lists and the segment
    masks of Q must give the oracle's bits, 1 and 3 steps."""
    cfg = cases.uniform_config(128, 32, 16, dx=1.0, dt=0.05, nscalars=1)
    o, s = pair(cfg, seed=5, cuboids=EDGE_OBSTACLES, sources=[(64.5, 4.5, 0.5, 1.0, 0)], u0=3.0)
    for n in (1, 2):
        o.step(n)
        s.step(n)
        err = compare_state(s, o, cfg)
        assert max(err.values()) == 0.0, (n, err)


@pytest.mark.parametrize("env", [{"ELMM_MOMENTUM": "percomp"}, {"ELMM_ZCHUNK": "5"}, {"ELMM_SCALAR": "cellwise"},
                                 {"ELMM_SGS": "cellwise"}, {"ELMM_GHOST_DELTA": "0"}, {"ELMM_ZPIPE": "1"},
                                 {"ELMM_SGS_ZCHUNK": "5"}])
def test_momentum_kernel_variants_bit_identical(require_gpu, monkeypatch, env):
    """The per-component momentum kernels, the one-thread-per-cell scalar and
    SGS kernels, the z-marching kernels with short z chunks (chunk boundaries
    inside the domain), the full ghost fill after IBMomentum (instead of the
    listed ghost cells) and the z solve with pipelined loads (the variant
    large grids use) give the same bits as the oracle on the obstacle case."""
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    cfg = cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05, nscalars=1)
    o, s = pair(cfg, seed=4, cuboids=OBSTACLES, sources=[(4.5, 4.5, 0.5, 1.0, 0)], u0=3.0)
    o.step(2)
    s.step(2)
    err = compare_state(s, o, cfg)
    assert max(err.values()) == 0.0, err


@pytest.mark.parametrize("env", [{"ELMM_ZPIPE": "1"}, {"ELMM_MOM_TILE": "32x8"}, {"ELMM_MOM_TILE": "64x8"},
                                 {"ELMM_MOM_TILE": "64x4", "ELMM_ZCHUNK": "17"}])
@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32_REFINED])
def test_variants_across_z_segments(require_gpu, monkeypatch, env, precision):
    """nz = 40: two full 16-level segments of the pipelined z solve and a
    partial one (its cross-segment prefetch and register rotation), and the
    other momentum tile shapes: the same bits as the default launch (FP64:
    and as the oracle) after 2 steps."""
    cfg = capi.with_precision(cases.uniform_config(64, 32, 40, dx=1.0, dt=0.05, nscalars=1), precision)
    cub = [(8.0, 16.0, 8.0, 16.0, 0.0, 22.0), (40.0, 52.0, 18.0, 30.0, 0.0, 33.0, 0.01)]
    case = capi.Case(cfg, cuboids=cub, sources=[(4.5, 4.5, 0.5, 1.0, 0)])
    f = cases.initial_fields(cfg, u0=3.0, amp=0.3, seed=6)
    ref = capi.Solver(cfg, case)
    cases.upload_initial(ref, cfg, f)
    ref.init()
    ref.step(2)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    s = capi.Solver(cfg, case)
    cases.upload_initial(s, cfg, f)
    s.init()
    s.step(2)
    err = compare_state(s, ref, cfg)
    assert max(err.values()) == 0.0, (env, err)
    if precision == capi.FP64:
        o = Oracle(cfg, case)
        cases.upload_initial(o, cfg, f)
        o.init()
        o.step(2)
        err = compare_state(s, o, cfg)
        assert max(err.values()) == 0.0, (env, err)


@pytest.mark.parametrize("zpipe", ["0", "1"])
def test_deep_z_beyond_256_levels(require_gpu, monkeypatch, zpipe):
    """nz = 300: 19 segments of the z solve, the last one partial, with its
    checkpoints and the mean mode's sums in LDS sized at launch (the limit
    was nz <= 256); FP64 bit-identical to the oracle after 2 steps, with and
    without the pipelined z solve."""
    monkeypatch.setenv("ELMM_ZPIPE", zpipe)
    cfg = cases.uniform_config(32, 32, 300, dx=1.0, dt=0.05, nscalars=1)
    o, s = pair(cfg, seed=9, cuboids=OBSTACLES, sources=[(4.5, 4.5, 0.5, 1.0, 0)], u0=3.0)
    o.step(2)
    s.step(2)
    err = compare_state(s, o, cfg)
    assert max(err.values()) == 0.0, err


def test_autotuned_tiles_give_the_same_bits(require_gpu):
    """The autotuner times the candidate tiles on the first step and keeps the
    fastest; the results equal the default launch bit for bit, and the chosen
    shapes are reported (elmm_get_tiles)."""
    for precision in (capi.FP64, capi.FP32_REFINED):
        base = capi.with_precision(cases.uniform_config(128, 64, 48, dx=1.0, dt=0.05, nscalars=1), precision)
        tuned = capi.with_precision(cases.uniform_config(128, 64, 48, dx=1.0, dt=0.05, nscalars=1, autotune=1),
                                    precision)
        case = capi.Case(base, cuboids=[(20.0, 36.0, 20.0, 36.0, 0.0, 14.0)])
        f = cases.initial_fields(base, u0=2.0, amp=0.2, seed=8)
        out = []
        for cfg in (base, tuned):
            s = capi.Solver(cfg, case)
            cases.upload_initial(s, cfg, f)
            s.init()
            s.step(3)
            out.append(s)
        t = out[1].tiles()
        print(f"precision {precision}: tuned tiles {t}; default {out[0].tiles()}", flush=True)
        assert t["tuned"] == 1 and out[0].tiles()["tuned"] == 0
        # 64x8 (512 threads) is a candidate for 4-byte fields only
        sizes = (256, 512) if precision == capi.FP32_REFINED else (256,)
        assert t["mom_tx"] * t["mom_ty"] in sizes and t["mom_zchunk"] in (32, 64)
        err = compare_state(out[1], out[0], base)
        assert max(err.values()) == 0.0, err
