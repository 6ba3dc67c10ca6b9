"""Geometric multigrid solve of the pressure equation (poisson_solver =
ELMM_POISSON_MG; the role of PoissMG, src/multigrid.f90:1047-1172).

CPU: the oracle's multigrid restatement is pinned against the direct solve
(itself pinned by the residual, eigenfunction and dense-solve checks of
test_oracle_pins.py): on random compatible right-hand sides of the stretched
channel and of uniform grids (including one whose coarsest level is 12 x 6),
the max-norm residual of Residue_variable_z (src/poisson.f90:23-60, the
oracle's independent residual routine) is below 1e-10 and the gradient of
Phi, what the projection uses, equals the direct solve's; a full step with
the multigrid solve stays within 1e-10 of the step with the direct solve.
GPU: the device multigrid is bit-identical to the oracle's (FP64), on the
stretched channel and on an obstacle case, after the initial projection and
1 and 3 steps; the 256 x 128 x 128 channel (C2 as stated: multigrid to a
1e-10 residual) agrees with the spectral solve after one step; the precision
form stays within 1e-5 of the oracle.
"""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state, velocity_rel_l2
from oracle_py import Oracle


def _compatible_rhs(cfg, seed):
    rng = np.random.default_rng(seed)
    rhs = rng.uniform(-1, 1, (cfg.nx, cfg.ny, cfg.nz))
    if cfg.stretched_z:
        dz = np.diff(np.array(cfg.zw[0:cfg.nz + 1]))
    else:
        dz = np.full(cfg.nz, cfg.dz)
    rhs -= (rhs * dz[None, None, :]).sum() / (dz.sum() * cfg.nx * cfg.ny)
    return np.asfortranarray(rhs)


def _ghosted(a):
    """Interior array (nx, ny, nz) into the ghosted layout (lower bound -2)
    with the Bound_Phi ghosts the residual reads: periodic x, y and the
    Neumann mirror in z."""
    nx, ny, nz = a.shape
    g = np.zeros((nx + 6, ny + 6, nz + 6), order="F")
    g[3:3 + nx, 3:3 + ny, 3:3 + nz] = a
    g[2, 3:3 + ny, 3:3 + nz] = a[-1]
    g[3 + nx, 3:3 + ny, 3:3 + nz] = a[0]
    g[:, 2, 3:3 + nz] = g[:, 2 + ny, 3:3 + nz]
    g[:, 3 + ny, 3:3 + nz] = g[:, 3, 3:3 + nz]
    g[:, :, 2] = g[:, :, 3]
    g[:, :, 3 + nz] = g[:, :, 2 + nz]
    return g


@pytest.mark.parametrize("grid", ["channel", "uniform", "uniform_12x6_coarsest"])
def test_oracle_multigrid_matches_direct_solve(grid):
    if grid == "channel":
        cfg = cases.channel_config(64, 32, 32)
    elif grid == "uniform":
        cfg = cases.uniform_config(32, 32, 16)
    else:
        cfg = cases.uniform_config(96, 48, 24)
    o = Oracle(cfg, capi.Case(cfg))
    rhs = _compatible_rhs(cfg, 1)
    direct = o.op_poisson(rhs)
    phi, cycles, res = o.op_poisson_mg(rhs)
    assert 1 <= cycles <= 12 and res < 1e-10
    r_indep = o.op_residual(_ghosted(phi), _ghosted(rhs))
    assert r_indep < 2e-10, r_indep
    scale = np.abs(direct).max()
    for ax in range(3):
        d = np.abs(np.diff(phi, axis=ax) - np.diff(direct, axis=ax)).max()
        assert d <= 1e-8 * scale, (grid, ax, d, scale)
    o.close()


def test_oracle_multigrid_step_equals_direct_step():
    base = cases.channel_config(64, 32, 32)
    mgc = cases.channel_config(64, 32, 32, poisson_solver=capi.POISSON_MG)
    f = cases.channel_fields(base, seed=2)
    out = []
    for cfg in (base, mgc):
        o = Oracle(cfg, capi.Case(cfg))
        cases.upload_initial(o, cfg, f)
        o.init()
        d = o.step(2)
        out.append((o, d))
    (o1, d1), (o2, d2) = out
    assert d2.poisson_res < 1e-10
    err = compare_state(o2, o1, base, ["U", "V", "W"])
    assert max(err.values()) <= 1e-10, err
    o1.close()
    o2.close()


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("case_kind", ["channel", "obstacles"])
def test_gpu_multigrid_bit_identical_to_oracle(require_gpu, case_kind):
    if case_kind == "channel":
        cfg = cases.channel_config(64, 32, 32, poisson_solver=capi.POISSON_MG)
        case = capi.Case(cfg)
        f = cases.channel_fields(cfg, seed=4)
        fields = ["U", "V", "W", "Pr", "Visc"]
    else:
        cfg = cases.uniform_config(48, 32, 16, dx=1.0, dt=0.05, nscalars=1, poisson_solver=capi.POISSON_MG)
        case = capi.Case(cfg, cuboids=[(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (28.0, 36.0, 16.0, 26.0, 0.0, 9.0, 0.02)],
                         sources=[(4.5, 4.5, 0.5, 1.0, 0)])
        f = cases.initial_fields(cfg, u0=2.0, amp=0.3, seed=13)
        fields = None
    o, s = Oracle(cfg, case), capi.Solver(cfg, case)
    for x in (o, s):
        cases.upload_initial(x, cfg, f)
        x.init()
    err = compare_state(s, o, cfg, fields)
    assert max(err.values()) == 0.0, err
    for n in (1, 2):
        do, ds = o.step(n), s.step(n)
        err = compare_state(s, o, cfg, fields)
        print(f"multigrid {case_kind}: {ds.refinements} cycles, residual {ds.poisson_res:.2e}, diff {err}",
              flush=True)
        assert max(err.values()) == 0.0, err
        assert ds.poisson_res < 1e-10 and ds.refinements == do.refinements
    o.close()
    s.close()


@pytest.mark.gpu
def test_gpu_multigrid_channel256_matches_spectral_solve(require_gpu):
    """C2 as stated: 256 x 128 x 128 stretched channel with the multigrid
    pressure solve to a 1e-10 residual, against the spectral solve after the
    initial projection and one step, and the precision form with multigrid
    against the FP64 multigrid path."""
    base = cases.channel_config(256, 128, 128)
    mgc = cases.channel_config(256, 128, 128, poisson_solver=capi.POISSON_MG)
    case = capi.Case(base)
    f = cases.channel_fields(base, seed=2)
    s_fft, s_mg = capi.Solver(base, case), capi.Solver(mgc, case)
    s_mg32 = capi.Solver(capi.with_precision(mgc, capi.FP32_REFINED), case)
    for s in (s_fft, s_mg, s_mg32):
        cases.upload_initial(s, base, f)
        s.init()
    d_fft, d_mg, d32 = s_fft.step(1), s_mg.step(1), s_mg32.step(1)
    err = compare_state(s_mg, s_fft, base, ["U", "V", "W"])
    print(f"channel256 multigrid: {d_mg.refinements} cycles, residual {d_mg.poisson_res:.2e}, "
          f"max_div {d_mg.max_div:.2e} (spectral {d_fft.max_div:.2e}); vs spectral {err}", flush=True)
    assert d_mg.poisson_res < 1e-10 and d_mg.max_div < 1e-10
    assert max(err.values()) <= 1e-9, err
    assert d32.poisson_res < 1e-10 and d32.max_div <= 1e-10
    l2 = velocity_rel_l2(s_mg32, s_mg, base)
    assert l2 <= 1e-5, l2
    for s in (s_fft, s_mg, s_mg32):
        s.close()
