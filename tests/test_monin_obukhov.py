"""Monin-Obukhov wall laws (src/wallmodels.f90:1546-1716, 1792-1938) on the
domain walls: a no-slip wall with the wall model takes a Dirichlet
temperature as its points' prescribed surface temperature (WM_MO_Dirichlet,
WM_MO_DirichletStress: u* and the surface heat flux by the Zilitinkevich-Esau
iteration) and a flux temperature (ELMM_BC_CONSTFLUX) as their surface heat
flux (WM_MO_FLUX, WM_MO_FluxStress when positive); the temperature then gets
the flux times the point's area factor (TemperatureExtra, src/scalars.f90:
322-326) and the mirror ghost rule on that wall (TempBtype = 0,
src/wallmodels.f90:618-620; scalarboundaries.f90:233-239, 271-277).  With
moisture, the Dirichlet law takes the virtual heat flux and returns the
moisture flux too (WM_MO_Dirichlet(Stress)_moist,
WM_MO_Dirichlet_ustar_tfl_mfl, src/wallmodels.f90:1721-1784, 1887-1963); a
Dirichlet moisture is the points' surface moisture, a flux moisture their
moisture flux (and the mirror ghost rule), added like the heat
(MoistureExtra, src/scalars.f90:336-356).

CPU pins of the oracle: its fixed-operation exp, atan and pow within a few
ulps of libm; the stability functions against their textbook formulas (libm);
the flux law's u* a fixed point of u* = 0.4 v / ln(z/z0 - Psi_M(z/L(u*)))
within its stopping tolerance, and the neutral log law in the limit of a
vanishing flux; the Dirichlet law's (u*, flux) consistent with the bulk
relations at the Obukhov length they imply, unstable for a warm surface and
stable for a cold one, and the molecular flux in the laminar branch; a
heated floor under fluid at rest warms its first layer by q dt / dz per step
and nothing else (a lost area factor, a sign error of the top wall or a
wrong ghost rule fails it); the moist law's three fluxes against the bulk
relations at the virtual flux's Obukhov length; a moisture-flux floor wets
its first layer by m dt / dz.  GPU: bit-identical to the oracle (FP64) with a
warm rough floor, a heat-flux floor, a cooled rough top, and with moisture (a
wetter surface, a moisture-flux floor), with obstacles, after the initial
tail and 1 and 3 steps; the precision form within 1e-5.
"""
import math

import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state, velocity_rel_l2
from oracle_py import Oracle, det_fn, mo_dirichlet, mo_dirichlet_moist
from test_buoyancy import OBST, T_BOT, T_TOP, buoyant_config, temperature_field

NU = cases.NU_DEFAULT
KAPPA = cases.NU_SCALAR_DEFAULT


def _ulps(a, b):
    return abs(a - b) / math.ulp(b) if b != 0 else abs(a)


def test_det_exp_atan_pow_within_ulps_of_libm():
    xs = np.concatenate([np.linspace(-40, 40, 4001), [-1e-300, 0.0, 1e-12, 0.34657359, -0.34657359]])
    assert max(_ulps(det_fn("det_exp", x), math.exp(x)) for x in xs) <= 2
    ts = np.concatenate([np.linspace(-3, 3, 2001), np.logspace(-8, 6, 500), -np.logspace(-8, 6, 500)])
    assert max(_ulps(det_fn("det_atan", t), math.atan(t)) for t in ts) <= 2
    # exp(y log x): the rounding of y log x (|y log x| up to 12 here) moves
    # the result by |y log x| ulps at most
    for y in (0.25, 5.0 / 6.0, 0.8):
        bs = np.concatenate([np.linspace(1, 100, 1000), np.logspace(-6, 1, 300)])
        worst = max(_ulps(det_fn("det_pow", b, y), b ** y) / (2 + abs(y * math.log(b))) for b in bs)
        assert worst <= 2, worst
    assert det_fn("det_pow", 0.0, 5.0 / 6.0) == 0.0


def _psim(zeta):
    if zeta < 0:
        x = (1 - 15 * zeta) ** 0.25
        return math.log((1 + x * x) / 2 * ((1 + x) / 2) ** 2) - 2 * math.atan(x) + math.pi / 2
    return -4.8 * zeta


def _psim_mod(zeta):
    if zeta < 0:
        x = (1 - 6 * zeta * 0.4) ** 0.25
        return math.log((1 + x * x) / 2 * ((1 + x) / 2) ** 2) - 2 * math.atan(x) + math.pi / 2
    return -3 * (zeta * 0.4) ** (5 / 6)


def _psih_mod(zeta):
    if zeta < 0:
        x = (1 - 6 * zeta * 0.4) ** 0.25
        return 2 * math.log((1 + x * x) / 2)
    return -2.5 * (zeta * 0.4) ** 0.8


def test_stability_functions_match_their_formulas():
    for z in list(np.linspace(-50, -1e-6, 400)) + [1e-8, 0.01, 0.3, 1.0, 7.0]:
        for name, ref in (("psim_mo", _psim), ("psim_mo_mod", _psim_mod), ("psih_mo_mod", _psih_mod)):
            got, want = det_fn(name, z), ref(z)
            assert abs(got - want) <= 1e-13 * max(1.0, abs(want)), (name, z, got, want)
    assert det_fn("psim_mo", 0.0) == 0.0 and det_fn("psim_mo_mod", 0.0) == 0.0
    # unstable: Psi_M > 0 and growing with -zeta; stable: negative
    assert 0 < det_fn("psim_mo", -0.1) < det_fn("psim_mo", -1.0)
    assert det_fn("psih_mo_mod", 0.5) < 0


def _obukhov(ustar, tflux, g, tref, z):
    return z * (0.4 * (g / tref) * tflux) / (-ustar ** 3)


@pytest.mark.parametrize("vel,dist,z0,tflux", [(2.0, 0.5, 0.01, 0.05), (5.0, 2.0, 0.1, 0.2), (1.0, 0.5, 0.02, 0.5)])
def test_flux_law_is_a_fixed_point(vel, dist, z0, tflux):
    g, tref = 9.81, 295.0
    u = det_fn("mo_flux_ustar", vel, dist, 1.0, z0, tflux, tref, g, NU)
    zl = _obukhov(u, tflux, g, tref, dist)
    assert zl < 0  # a positive (upward) heat flux is unstable
    want = 0.4 * vel / math.log(dist / z0 - _psim(zl))
    assert abs(u - want) <= 3e-3 * want, (u, want)
    neutral = 0.4 * vel / math.log(dist / z0)
    assert u > neutral  # convective mixing raises u*
    u0 = det_fn("mo_flux_ustar", vel, dist, 1.0, z0, 1e-14, tref, g, NU)
    assert abs(u0 - neutral) <= 2e-3 * neutral


@pytest.mark.parametrize("tempdif", [-3.0, -0.5, 0.5, 2.0])
def test_dirichlet_law_bulk_relations(tempdif):
    """tempdif = T(air) - T(surface): negative for a warm surface."""
    g, tref, vel, dist, z0 = 9.81, 295.0, 2.0, 0.5, 0.01
    u, tf = 1.0, 0.0
    for _ in range(3):  # the state persists from stage to stage
        u, tf = mo_dirichlet(u, tf, dist, z0, vel, tempdif, NU, KAPPA, g, tref)
    L = math.log(dist / z0)
    zl = -dist * g * tf / (0.4 * tref * u ** 3)
    assert (tf > 0) == (tempdif < 0) and (zl < 0) == (tempdif < 0)
    u_want = vel * 0.4 / (L - _psim_mod(zl))
    tf_want = -tempdif * u_want * 0.47 / (L - _psih_mod(zl))
    assert abs(u - u_want) <= 5e-3 * u_want and abs(tf - tf_want) <= 5e-3 * abs(tf_want), (u, u_want, tf, tf_want)
    neutral = 0.4 * vel / L
    assert (u > neutral) == (tempdif < 0)
    # laminar branch: the molecular flux
    u2, tf2 = mo_dirichlet(1.0, 0.0, dist, z0, 1e-9, tempdif, NU, KAPPA, g, tref)
    assert tf2 == -((KAPPA * tempdif) / dist)


def mo_config(nx=32, ny=32, nz=16, bottom="temp", top="slip", **kw):
    """Rough no-slip floor (z0 0.01) with a prescribed surface temperature
    (warmer than the air) or heat flux; top free-slip Dirichlet, or a rough
    no-slip top with a prescribed (cooler) surface temperature."""
    bt = capi.DIRICHLET if bottom == "temp" else capi.CONSTFLUX
    bval = T_BOT + 2.0 if bottom == "temp" else 0.1
    args = dict(temp_btype=[capi.PERIODIC] * 4 + [bt, capi.DIRICHLET], side_temp=[0, 0, 0, 0, bval, T_TOP])
    if top == "wall":
        args.update(top=capi.NOSLIP)
        args["side_temp"][5] = T_TOP - 2.0
    args.update(kw)
    cfg = buoyant_config(nx, ny, nz, **args)
    if top == "wall":
        cfg.z0_wall[5] = 0.02
    return cfg


def test_wall_points_take_the_thermal_class():
    cfg = mo_config(bottom="flux", top="wall")
    case = capi.Case(cfg, cuboids=OBST)
    ijk, _, _ = case.wmp_list()
    tc, tv = case.wm_thermal(-1)
    floor, roof = ijk[:, 2] == 1, ijk[:, 2] == cfg.nz
    assert np.all(tc[floor & (tc != 0)] == capi.WM_SURFACE_FLUX) and np.all(tv[tc == capi.WM_SURFACE_FLUX] == 0.1)
    assert np.all(tc[roof] == capi.WM_SURFACE_TEMP) and np.all(tv[roof] == T_TOP - 2.0)
    assert (tc == capi.WM_NEUTRAL).sum() > 0  # the obstacles' points
    assert (tc == capi.WM_SURFACE_FLUX).sum() > 500
    tcu, _ = case.wm_thermal(0 * 6 + 4)  # U points over the floor
    assert (tcu == capi.WM_SURFACE_FLUX).sum() > 500


HEATED = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (22.0, 30.0, 18.0, 28.0, 0.0, 9.0, 0.02, 0.2, 3e-5)]


def test_rough_body_carries_its_surface_fluxes():
    """A rough cuboid's heat and moisture flux reach its wall-model points
    (the flux class); a smooth one's points stay neutral."""
    cfg = buoyant_config(enable_moisture=1, moisture_ref=0.01, moist_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN] * 2,
                         side_moist=[0] * 6)
    case = capi.Case(cfg, cuboids=HEATED + [(2.0, 5.0, 24.0, 28.0, 0.0, 4.0, 0.0, 0.5, 1e-4)])
    _, _, z0 = case.wmp_list()
    tc, tv = case.wm_thermal(-1)
    _, qf = case.wm_moist(-1)
    body = z0 == 0.02  # the heated body's points (the floor's z0 is 0.01)
    assert body.sum() > 50
    assert np.all(tc[body] == capi.WM_SURFACE_FLUX) and np.all(tv[body] == 0.2) and np.all(qf[body] == 3e-5)
    assert np.all(tc[~body] == capi.WM_NEUTRAL) and np.all(qf[~body] == 0)


@pytest.mark.parametrize("where", ["bottom", "top"])
def test_oracle_heat_flux_warms_the_wall_layer(where):
    """Fluid at rest, uniform temperature, a surface heat flux q on one wall
    (the other Neumann, zero gradient): after one step the wall layer has
    changed by q dt / dz (bottom: +, top: the flux leaves through the top
    wall, -) and every other cell is unchanged, up to the molecular
    diffusion within the step, and the wall's ghost cells mirror the
    interior."""
    q = 0.1
    if where == "bottom":
        bc = dict(temp_btype=[capi.PERIODIC] * 4 + [capi.CONSTFLUX, capi.NEUMANN], side_temp=[0, 0, 0, 0, q, 0.0])
        cfg = buoyant_config(16, 16, 8, **bc)
    else:
        bc = dict(temp_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.CONSTFLUX], side_temp=[0, 0, 0, 0, 0.0, q])
        cfg = buoyant_config(16, 16, 8, top=capi.NOSLIP, **bc)
        cfg.z0_wall[5] = 0.02
    o = Oracle(cfg, capi.Case(cfg))
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    zero = np.zeros(shape, order="F")
    for f in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR, capi.F_SCALAR):
        o.upload(f, zero)
    T0 = 295.0
    o.upload(capi.F_TEMPERATURE, np.full(shape, T0, order="F"))
    o.init()
    o.step(1)
    T = o.download(capi.F_TEMPERATURE)
    k_wall = 1 if where == "bottom" else cfg.nz
    dT = (q if where == "bottom" else -q) * cfg.dt / cfg.dz
    inner = T[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]
    layer = inner[:, :, k_wall - 1]
    # within the step the molecular diffusivity moves about kappa dT dt / dz^2
    # (5e-9 K here) into the next layer
    tol = 10 * KAPPA * abs(dT) * cfg.dt / cfg.dz ** 2
    assert np.abs(layer - (T0 + dT)).max() <= tol, np.abs(layer - (T0 + dT)).max()
    others = np.delete(inner, k_wall - 1, axis=2)
    assert np.abs(others - T0).max() <= tol
    if where == "bottom":
        assert np.array_equal(T[3:3 + cfg.nx, 3:3 + cfg.ny, 2], T[3:3 + cfg.nx, 3:3 + cfg.ny, 3])
        assert np.array_equal(T[3:3 + cfg.nx, 3:3 + cfg.ny, 0], T[3:3 + cfg.nx, 3:3 + cfg.ny, 5])
    else:
        n = cfg.nz
        assert np.array_equal(T[3:3 + cfg.nx, 3:3 + cfg.ny, n + 3], T[3:3 + cfg.nx, 3:3 + cfg.ny, n + 2])
        assert np.array_equal(T[3:3 + cfg.nx, 3:3 + cfg.ny, n + 5], T[3:3 + cfg.nx, 3:3 + cfg.ny, n])
    o.close()


@pytest.mark.parametrize("floor", ["warm", "neumann"])
def test_oracle_warm_floor_heats_the_air(floor):
    """Wind over a rough floor 2 K warmer than the air (uniform 295 K): the
    Dirichlet law's upward heat flux warms the first layer within a step by
    about flux dt / dz (the flux from the bulk relation, 0.47 u* 2 K /
    (ln(z/z0) - Psi_H) ~ 0.03 K m/s); a zero-gradient floor leaves it."""
    if floor == "warm":
        bc = dict(temp_btype=[capi.PERIODIC] * 4 + [capi.DIRICHLET, capi.NEUMANN], side_temp=[0, 0, 0, 0, 297.0, 0])
    else:
        bc = dict(temp_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.NEUMANN], side_temp=[0] * 6)
    cfg = buoyant_config(16, 16, 8, **bc)
    o = Oracle(cfg, capi.Case(cfg))
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    U = np.full(shape, 2.0, order="F")
    zero = np.zeros(shape, order="F")
    o.upload(capi.F_U, U)
    for f in (capi.F_V, capi.F_W, capi.F_PR, capi.F_SCALAR):
        o.upload(f, zero)
    o.upload(capi.F_TEMPERATURE, np.full(shape, 295.0, order="F"))
    o.init()
    o.step(1)
    T = o.download(capi.F_TEMPERATURE)[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]
    rise = T[:, :, 0].mean() - 295.0
    if floor == "warm":
        assert 0.2 * 0.03 * cfg.dt < rise < 5 * 0.03 * cfg.dt, rise
        assert abs(T[:, :, 1:].mean() - 295.0) < 0.1 * rise
    else:
        assert abs(rise) < 1e-9


@pytest.mark.parametrize("tempdif,moistdif", [(-2.0, -0.004), (-0.5, 0.003), (1.0, -0.002)])
def test_moist_dirichlet_law_bulk_relations(tempdif, moistdif):
    """WM_MO_Dirichlet_ustar_tfl_mfl: (u*, heat flux, moisture flux)
    consistent with the bulk relations at the Obukhov length of the virtual
    heat flux tflux (1 + 0.61 q) + 0.61 T qflux; a drier surface (moistdif >
    0) gives a downward moisture flux."""
    g, tref, vel, dist, z0, T, q = 9.81, 295.0, 2.0, 0.5, 0.01, 295.0, 0.01
    u, tf, qf = 1.0, 0.0, 0.0
    for _ in range(3):
        u, tf, qf = mo_dirichlet_moist(u, tf, qf, dist, z0, vel, T, T - tempdif, q, q - moistdif, NU, KAPPA, g,
                                       tref)
    L = math.log(dist / z0)
    virt = tf * (1 + 0.61 * q) + 0.61 * T * qf
    zl = -dist * g * virt / (0.4 * tref * u ** 3)
    u_want = vel * 0.4 / (L - _psim_mod(zl))
    tf_want = -tempdif * u_want * 0.47 / (L - _psih_mod(zl))
    qf_want = -moistdif * u_want * 0.47 / (L - _psih_mod(zl))
    for got, want in ((u, u_want), (tf, tf_want), (qf, qf_want)):
        assert abs(got - want) <= 5e-3 * abs(want), (got, want)
    assert (qf > 0) == (moistdif < 0)


def test_oracle_moisture_flux_wets_the_wall_layer():
    """Fluid at rest, uniform temperature and moisture, a moisture flux m on
    the floor: after one step the first layer's moisture has risen by
    m dt / dz, up to molecular diffusion within the step."""
    m = 1e-4
    cfg = buoyant_config(16, 16, 8, temp_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.NEUMANN],
                         side_temp=[0] * 6, enable_moisture=1, moisture_ref=0.01,
                         moist_btype=[capi.PERIODIC] * 4 + [capi.CONSTFLUX, capi.NEUMANN],
                         side_moist=[0, 0, 0, 0, m, 0])
    o = Oracle(cfg, capi.Case(cfg))
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    zero = np.zeros(shape, order="F")
    for f in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR, capi.F_SCALAR):
        o.upload(f, zero)
    o.upload(capi.F_TEMPERATURE, np.full(shape, 295.0, order="F"))
    o.upload(capi.F_MOISTURE, np.full(shape, 0.01, order="F"))
    o.init()
    o.step(1)
    q = o.download(capi.F_MOISTURE)[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]
    dq = m * cfg.dt / cfg.dz
    tol = 10 * KAPPA * dq * cfg.dt / cfg.dz ** 2 + 1e-15
    assert np.abs(q[:, :, 0] - (0.01 + dq)).max() <= tol
    assert np.abs(q[:, :, 1:] - 0.01).max() <= tol
    o.close()


def test_moisture_flux_needs_a_wall_model_wall():
    cfg = mo_config(enable_moisture=1, moisture_ref=0.01,
                    moist_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.CONSTFLUX], side_moist=[0, 0, 0, 0, 0, 1e-4])
    with pytest.raises(capi.ElmmError, match="moisture flux"):
        capi.Case(cfg)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("bottom,top,moist", [("temp", "slip", None), ("flux", "slip", None), ("flux", "wall", None),
                                              ("temp", "slip", "dirichlet"), ("temp", "wall", "flux"),
                                              ("body", "slip", "flux")])
def test_gpu_monin_obukhov_bit_identical_to_oracle(require_gpu, bottom, top, moist):
    kw = {}
    if moist:
        mb = capi.DIRICHLET if moist == "dirichlet" else capi.CONSTFLUX
        mv = 0.014 if moist == "dirichlet" else 2e-4
        kw = dict(enable_moisture=1, moisture_ref=0.01, moist_btype=[capi.PERIODIC] * 4 + [mb, capi.NEUMANN],
                  side_moist=[0, 0, 0, 0, mv, 0])
    if bottom == "body":  # a neutral floor, a heated wet rough cuboid
        kw.update(temp_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.DIRICHLET], side_temp=[0, 0, 0, 0, 0.0, T_TOP])
    cfg = mo_config(48, 32, 16, bottom="temp", top=top, **kw)
    case = capi.Case(cfg, cuboids=HEATED if bottom == "body" else OBST, sources=[(4.5, 4.5, 0.5, 1.0, 0)])
    f = cases.initial_fields(cfg, u0=2.0, amp=0.3, seed=41)
    T = temperature_field(cfg, seed=42)
    o, s = Oracle(cfg, case), capi.Solver(cfg, case)
    s32 = capi.Solver(capi.with_precision(cfg, capi.FP32_REFINED), case)
    qf = None
    if moist:
        g = np.random.default_rng(43)
        qf = np.asfortranarray(0.01 + g.uniform(-0.001, 0.001, T.shape))
    for x in (o, s, s32):
        cases.upload_initial(x, cfg, f)
        x.upload(capi.F_TEMPERATURE, T)
        if moist:
            x.upload(capi.F_MOISTURE, qf)
        x.init()
    fields = ["U", "V", "W", "Pr", "Visc", "TDiff", "S0"]
    fids = [capi.F_TEMPERATURE] + ([capi.F_MOISTURE] if moist else [])

    def same():
        err = compare_state(s, o, cfg, fields)
        assert max(err.values()) == 0.0, err
        for fid in fids:
            assert np.array_equal(s.download(fid), o.download(fid)), fid
        assert np.array_equal(s.wm_ustar(), o.wm_ustar())

    same()
    for n in (1, 2):
        o.step(n)
        s.step(n)
        s32.step(n)
        same()
        l2 = velocity_rel_l2(s32, o, cfg)
        t32 = np.abs(s32.download(capi.F_TEMPERATURE) - o.download(capi.F_TEMPERATURE)).max() / T_TOP
        print(f"MO {bottom}/{top}/{moist} after {n} more steps: precision form velocity L2 {l2:.2e}, temperature {t32:.2e}",
              flush=True)
        assert l2 <= 1e-5 and t32 <= 1e-5, (l2, t32)
    for x in (o, s, s32):
        x.close()
