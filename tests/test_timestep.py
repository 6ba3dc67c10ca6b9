"""Time-step control of TMarchRK3 and the main loop: TimeStepLength's CFL
scan and variable dt (src/dynamics.f90:511-600), TMarchRK3's delta
(src/tsteps.f90:188-207), the time update and divergence tests
(src/main.f90:112-169).

CPU tests pin the oracle: the CFL maximum against a numpy brute force of the
formula, the variable dt against min(CFL / max, dt_max) and the end-time
cut, and delta against its closed form on a uniform flow driven by a mean
pressure gradient.  GPU tests compare the HIP library with the oracle
through the C ABI.
"""
import numpy as np
import pytest

from elmm_amd import capi, cases
from oracle_py import Oracle


def cfl_max_numpy(U, V, W, cfg):
    """max over Pr cells of sum_dir max(|u_l|, |u_r|) / delta (dynamics.f90:574-588)."""
    n = (cfg.nx, cfg.ny, cfg.nz)
    s = np.s_[3:3 + n[0], 3:3 + n[1], 3:3 + n[2]]
    sx = np.s_[2:2 + n[0], 3:3 + n[1], 3:3 + n[2]]
    sy = np.s_[3:3 + n[0], 2:2 + n[1], 3:3 + n[2]]
    sz = np.s_[3:3 + n[0], 3:3 + n[1], 2:2 + n[2]]
    p = np.maximum(np.abs(U[s]), np.abs(U[sx])) / cfg.dx
    p = p + np.maximum(np.abs(V[s]), np.abs(V[sy])) / cfg.dy
    p = p + np.maximum(np.abs(W[s]), np.abs(W[sz])) / cfg.dz
    return float(np.max(p))


def _oracle(cfg, seed=1, u0=1.0, fields=None):
    case = capi.Case(cfg)
    f = fields if fields is not None else cases.initial_fields(cfg, u0=u0, amp=0.1 * u0, seed=seed)
    o = Oracle(cfg, case)
    cases.upload_initial(o, cfg, f)
    o.init()
    return o, case, f


def test_oracle_cfl_scan_matches_brute_force():
    cfg = cases.config1(16, nscalars=0)
    o, _, _ = _oracle(cfg, u0=2.0)
    o.step(1)
    # the CFL of the next step is evaluated on the state the step starts from
    U, V, W = (o.download(c) for c in range(3))
    m = cfl_max_numpy(U, V, W, cfg)
    d = o.step(1)
    assert d.cfl == m * cfg.dt


def test_oracle_variable_dt_is_min_of_cfl_and_dt_max():
    cfg = cases.config1(16, nscalars=0, variable_dt=1, cfl_target=0.1, dt_max=0.05)
    o, _, _ = _oracle(cfg, u0=3.0)
    U, V, W = (o.download(c) for c in range(3))
    m = cfl_max_numpy(U, V, W, cfg)
    d = o.step(1)
    assert d.dt == min(0.1 / m, 0.05) and d.dt < 0.05
    assert d.time == d.dt
    # dt_max binds on a slow flow
    cfg2 = cases.config1(16, nscalars=0, variable_dt=1, cfl_target=0.3, dt_max=0.01)
    o2, _, _ = _oracle(cfg2, u0=1.0)
    d2 = o2.step(2)
    assert d2.dt == 0.01 and d2.time == 0.02


def test_oracle_variable_dt_is_cut_at_end_time():
    cfg = cases.config1(16, nscalars=0, variable_dt=1, cfl_target=0.3, dt_max=0.05, end_time=0.07)
    o, _, _ = _oracle(cfg, u0=1.0)
    d1 = o.step(1)
    assert d1.dt == 0.05
    d2 = o.step(1)
    assert d2.dt == 0.07 - 0.05 and d2.time == 0.07


def test_oracle_delta_closed_form_uniform_forced_flow():
    """Uniform U, V between free-slip walls, no SGS, driven by a mean pressure
    gradient (gx, gy): every stage adds -coef (gx, gy) uniformly and the
    projection leaves it, so delta = sum_stages (2 alpha dt |g|) / dt summed
    over U and V = |gx| + |gy| (sum of 2 alpha = 1)."""
    gx, gy = -0.75, 0.5
    cfg = cases.uniform_config(16, 16, 16, nscalars=0, top=capi.FREESLIP, sgs_model=0, wall_model=0,
                               delta_diag=1, pr_gradient_x=gx, pr_gradient_y=gy,
                               btype=[capi.PERIODIC] * 4 + [capi.FREESLIP, capi.FREESLIP])
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    U, V, W, Pr = (np.zeros(shape, order="F") for _ in range(4))
    U[3:3 + 16, 3:3 + 16, 3:3 + 16] = 2.0
    V[3:3 + 16, 3:3 + 16, 3:3 + 16] = -1.0
    o, _, _ = _oracle(cfg, fields=(U, V, W, Pr, []))
    d = o.step(1)
    assert abs(d.delta - (abs(gx) + abs(gy))) <= 1e-12 * (abs(gx) + abs(gy)), d.delta
    Un = o.download(capi.F_U)
    assert np.allclose(Un[3:19, 3:19, 3:19], 2.0 - cfg.dt * gx, rtol=0, atol=1e-13)


def test_oracle_cfl_divergence_from_step_three():
    """CFL above CFL_max is diagnosed from time_step 3 on (main.f90:161-168)."""
    cfg = cases.config1(16, nscalars=0, cfl_max=0.01)
    o, _, _ = _oracle(cfg)
    d = o.step(2)
    assert d.steps_done == 2
    with pytest.raises(RuntimeError, match="diverged. 2"):
        o.step(3)


# ---------------------------------------------------------------- GPU parity

def _pair(cfg, seed=1, u0=1.0):
    case = capi.Case(cfg)
    f = cases.initial_fields(cfg, u0=u0, amp=0.1 * u0, seed=seed)
    o = Oracle(cfg, case)
    s = capi.Solver(cfg, case)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg, f)
    o.init()
    s.init()
    return o, s


def _fields_equal(o, s):
    for c in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR):
        a, b = o.download(c), s.download(c)
        assert np.array_equal(a[3:-3, 3:-3, 3:-3], b[3:-3, 3:-3, 3:-3]), c


@pytest.mark.gpu
def test_gpu_delta_cfl_time_match_oracle(require_gpu):
    cfg = cases.config1(32, delta_diag=1, start_time=1.5)
    o, s = _pair(cfg)
    for n in (1, 4):
        do, ds = o.step(n), s.step(n)
        print(f"delta oracle {do.delta!r} gpu {ds.delta!r}; cfl {do.cfl!r} {ds.cfl!r}; time {do.time!r} {ds.time!r}",
              flush=True)
        assert ds.delta == do.delta and ds.delta > 0
        assert ds.cfl == do.cfl and ds.time == do.time and ds.steps_done == n
    _fields_equal(o, s)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32_REFINED])
def test_gpu_variable_dt_matches_oracle(require_gpu, precision):
    cfg = capi.with_precision(cases.config1(32, variable_dt=1, cfl_target=0.1, dt_max=0.05, end_time=0.4),
                              precision)
    o, s = _pair(cfg, u0=3.0)
    for n in (1, 3, 5):
        do, ds = o.step(n), s.step(n)
        print(f"precision {precision}: dt oracle {do.dt!r} gpu {ds.dt!r} time {do.time!r} {ds.time!r}", flush=True)
        if precision == capi.FP64:
            assert ds.dt == do.dt and ds.time == do.time
        else:
            assert abs(ds.dt - do.dt) <= 1e-5 * do.dt
        assert ds.dt < 0.05
    if precision == capi.FP64:
        _fields_equal(o, s)


@pytest.mark.gpu
def test_gpu_cfl_divergence_reported_from_step_three(require_gpu):
    cfg = cases.config1(32, cfl_max=0.01)
    _, s = _pair(cfg)
    d = s.step(2)
    assert d.steps_done == 2
    with pytest.raises(capi.ElmmError) as e:
        s.step(3)
    print(f"divergence: {e.value}", flush=True)
    assert e.value.code == -4 and "step 3" in str(e.value)


@pytest.mark.gpu
def test_gpu_async_wait_reports_cfl_and_divergence(require_gpu):
    cfg = cases.config1(32)
    o, s = _pair(cfg)
    s.step_async(4)
    d = s.wait_diag()
    do = o.step(4)
    assert d.steps_done == 4 and d.cfl == do.cfl and d.time == do.time
    # NaN velocity: the CFL monitor makes it a divergence at the next join
    cfg2 = cases.config1(32, cfl_max=1.0)
    _, s2 = _pair(cfg2)
    s2.step(3)
    u = s2.download(capi.F_U)
    u[10, 11, 12] = np.nan
    s2.upload(capi.F_U, u)
    s2.step_async(2)
    with pytest.raises(capi.ElmmError) as e:
        s2.wait_diag()
    assert e.value.code == -4
