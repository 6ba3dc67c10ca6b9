"""Buoyant temperature (enable_buoyancy): the temperature's RK3 stage
(src/scalars.f90:124-211, with its solid-cell values 131-207), the buoyancy
force on W (BuoyancyForce, 2nd order, src/dynamics.f90:608-700) and TDiff
recomputed every stage (ComputeTDiff, src/scalar_diffusion.f90:983-1009).

CPU pins of the oracle: the force against its closed form A/2 (T(k+1) +
T(k)) - A T_ref on a random temperature; the solid-cell values against a
numpy brute force; hydrostatic balance (fluid at rest with a linear
temperature profile stays at rest, the pressure gradient balances the
buoyancy face by face, the temperature profile is unchanged); a planted
sign error in the force is caught by the balance check.
GPU: bit-identical to the oracle (FP64) with obstacles, after the initial
tail and 1 and 3 steps; the precision form within 1e-5.
"""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state, velocity_rel_l2
from oracle_py import F_WSTAR, Oracle

OBST = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (22.0, 30.0, 18.0, 28.0, 0.0, 9.0, 0.02)]


T_BOT, T_TOP = 293.0, 299.0


def temperature_bcs(nz, dz=1.0):
    """Periodic x/y; the bottom (a wall with wall-model points, neutral wall
    laws with a Neumann temperature) with the gradient of the linear profile
    from T_BOT at z = 0 to T_TOP at z = nz dz; the free-slip top Dirichlet
    T_TOP."""
    return dict(temp_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.DIRICHLET],
                side_temp=[0, 0, 0, 0, (T_TOP - T_BOT) / (nz * dz), T_TOP])


def buoyant_config(nx=32, ny=32, nz=16, **kw):
    args = dict(dx=1.0, dt=0.05, nscalars=1, enable_buoyancy=1, grav_acc=9.81, temperature_ref=295.0,
                **temperature_bcs(nz))
    args.update(kw)
    return cases.uniform_config(nx, ny, nz, **args)


def temperature_field(cfg, seed=3, amp=0.2, t_bot=T_BOT, t_top=T_TOP):
    """Linear profile between the Dirichlet values at the walls (z = 0 and
    z = nz dz) plus seeded noise; ghosted, lower bound -2."""
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    z = (np.arange(-2, cfg.nz + 4) - 0.5) * cfg.dz
    prof = t_bot + (t_top - t_bot) * z / (cfg.nz * cfg.dz)
    T = np.asfortranarray(np.broadcast_to(prof[None, None, :], shape).copy())
    if amp:
        g = np.random.Generator(np.random.Philox(key=[seed, 5]))
        T[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz] += g.uniform(-amp, amp, (cfg.nx, cfg.ny, cfg.nz))
    return T


def G(a, i, j, k):
    return a[i + 2, j + 2, k + 2]


def test_oracle_buoyancy_force_closed_form():
    cfg = buoyant_config()
    o = Oracle(cfg, capi.Case(cfg))
    T = temperature_field(cfg, amp=0.5)
    o.upload(capi.F_TEMPERATURE, T)
    ws = np.asfortranarray(np.random.default_rng(2).uniform(-1, 1, T.shape))
    o.upload(F_WSTAR, ws)
    o.op_buoyancy()
    got = o.download(F_WSTAR)
    A = cfg.grav_acc / cfg.temperature_ref
    want = ws.copy()
    s = (slice(3, 3 + cfg.nx), slice(3, 3 + cfg.ny), slice(3, 3 + cfg.nz - 1))
    want[s] = ws[s] + A / 2 * (T[3:3 + cfg.nx, 3:3 + cfg.ny, 4:3 + cfg.nz] + T[s]) - A * cfg.temperature_ref
    assert np.array_equal(got, want)
    o.close()


def test_oracle_temperature_in_solids_brute_force():
    cfg = buoyant_config()
    case = capi.Case(cfg, cuboids=OBST)
    o = Oracle(cfg, case)
    T = np.asfortranarray(np.random.default_rng(4).uniform(290, 300, (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)))
    o.upload(capi.F_TEMPERATURE, T)
    o.op_temp_fix()
    got = o.download(capi.F_TEMPERATURE)
    pt = case.masks()[0]
    want = T.copy()
    n = (cfg.nx, cfg.ny, cfg.nz)
    changed = 0
    for k in range(-1, n[2] + 3):
        for j in range(-1, n[1] + 3):
            for i in range(-1, n[0] + 3):
                if G(pt, i, j, k) <= 0:
                    continue
                vals = []
                for (di, dj, dk) in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    a, b, c = i + di, j + dj, k + dk
                    if not (-1 <= a <= n[0] + 2 and -1 <= b <= n[1] + 2 and -1 <= c <= n[2] + 2):
                        continue
                    if G(pt, a, b, c) <= 0:
                        vals.append(G(T, a, b, c))
                if vals:
                    s = 0.0
                    for v in vals:
                        s = s + v
                    want[i + 2, j + 2, k + 2] = s / len(vals)
                    changed += 1
    assert changed > 50
    assert np.array_equal(got, want)
    o.close()


def _hydrostatic(cfg):
    o = Oracle(cfg, capi.Case(cfg))
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    zero = np.zeros(shape, order="F")
    for f in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR):
        o.upload(f, zero)
    o.upload(capi.F_SCALAR, zero)
    T = temperature_field(cfg, amp=0.0)
    o.upload(capi.F_TEMPERATURE, T)
    o.init()
    o.step(3)
    return o, T


def test_oracle_hydrostatic_balance():
    """At rest with a linear temperature: velocity stays zero, Pr(k+1) -
    Pr(k) = dz A ((T(k+1) + T(k))/2 - T_ref) on every face, T unchanged."""
    cfg = buoyant_config()
    o, T0 = _hydrostatic(cfg)
    for c in (capi.F_U, capi.F_V, capi.F_W):
        assert np.abs(o.download(c)).max() <= 1e-12
    Pr = o.download(capi.F_PR)[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]
    T = o.download(capi.F_TEMPERATURE)
    A = cfg.grav_acc / cfg.temperature_ref
    tin = T[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]
    want = cfg.dz * A * ((tin[:, :, 1:] + tin[:, :, :-1]) / 2 - cfg.temperature_ref)
    assert np.abs(np.diff(Pr, axis=2) - want).max() <= 1e-9 * np.abs(want).max()
    assert np.abs(tin - T0[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]).max() <= 1e-10
    o.close()


def test_hydrostatic_check_catches_a_sign_error():
    """The same balance with the gravity sign flipped in the configuration
    (the force then points the other way): the pressure gradient flips, so
    the check above would fail on a sign mistake in the force."""
    cfg = buoyant_config(grav_acc=-9.81)
    o, _ = _hydrostatic(cfg)
    Pr = o.download(capi.F_PR)[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]
    T = o.download(capi.F_TEMPERATURE)[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]
    A = 9.81 / cfg.temperature_ref
    want = cfg.dz * A * ((T[:, :, 1:] + T[:, :, :-1]) / 2 - cfg.temperature_ref)
    assert np.abs(np.diff(Pr, axis=2) - want).max() > 1e-3
    o.close()


def test_temperature_flux_needs_a_wall_model_wall():
    """A temperature flux (ELMM_BC_CONSTFLUX) exists only as the surface heat
    flux of wall-model points; on the free-slip top it is rejected."""
    cfg = buoyant_config(temp_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.CONSTFLUX],
                         side_temp=[0, 0, 0, 0, 0.0, 0.1])
    with pytest.raises(capi.ElmmError, match="needs a no-slip wall with the wall model"):
        capi.Case(cfg)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_gpu_buoyancy_bit_identical_to_oracle(require_gpu):
    cfg = buoyant_config(48, 32, 16)
    case = capi.Case(cfg, cuboids=OBST, sources=[(4.5, 4.5, 0.5, 1.0, 0)])
    f = cases.initial_fields(cfg, u0=1.0, amp=0.2, seed=21)
    T = temperature_field(cfg, seed=22)
    o, s = Oracle(cfg, case), capi.Solver(cfg, case)
    s32 = capi.Solver(capi.with_precision(cfg, capi.FP32_REFINED), case)
    for x in (o, s, s32):
        cases.upload_initial(x, cfg, f)
        x.upload(capi.F_TEMPERATURE, T)
        x.init()
    fields = ["U", "V", "W", "Pr", "Visc", "TDiff", "S0"]
    err = compare_state(s, o, cfg, fields)
    assert max(err.values()) == 0.0, err
    assert np.array_equal(s.download(capi.F_TEMPERATURE), o.download(capi.F_TEMPERATURE))
    for n in (1, 2):
        o.step(n)
        s.step(n)
        s32.step(n)
        err = compare_state(s, o, cfg, fields)
        tdiff = np.abs(s.download(capi.F_TEMPERATURE) - o.download(capi.F_TEMPERATURE)).max()
        print(f"buoyancy after {n} more steps: {err}, temperature {tdiff:.2e}", flush=True)
        assert max(err.values()) == 0.0, err
        assert np.array_equal(s.download(capi.F_TEMPERATURE), o.download(capi.F_TEMPERATURE))
        l2 = velocity_rel_l2(s32, o, cfg)
        t32 = np.abs(s32.download(capi.F_TEMPERATURE) - o.download(capi.F_TEMPERATURE)).max() / 299.0
        assert l2 <= 1e-5 and t32 <= 1e-6, (l2, t32)
    for x in (o, s, s32):
        x.close()
