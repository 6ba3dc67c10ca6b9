"""Stretched-z channel (GRID_VARIABLE_Z, advection_method 2).

CPU pins of the oracle's variable-z operators: second-order convergence of
CDU/V/W_variable_z (src/advection_variable_z.f90:15-321) and of
MomentumDiffusion_variable_z_nobranch_2ord (src/momentum_diffusion.f90:
298-420) to the continuous operators on a tanh grid; the uniform-spacing
limit of the variable-z diffusion and wall-model replacement against the
uniform-grid path; the momentum budget of the wall-model face replacement
(wmfluxes-variable_z-nobranch-{U,V,W}-inc.f90) and the fluxp/fluxm scaling
(src/wallmodels.f90:2285-2310).

GPU: the HIP step against the oracle on a stretched channel (FP64: <= 1e-12
after one step, <= 1e-11 after ten; FP32-refined: relative L2 <= 1e-5 after
ten steps, divergence <= 1e-10).
"""
import numpy as np
import pytest

from elmm_amd import capi, cases
from oracle_py import F_USTAR, F_VSTAR, F_WSTAR, Oracle


def stretched_cfg(n, nz, wall_model=0, **kw):
    return cases.channel_config(n, n, nz, lx=2 * np.pi, ly=2 * np.pi, lz=2.0, gamma=1.5, nu=1e-2,
                                wall_model=wall_model, **kw)


def z_coords(zw):
    """zW(k) and zPr(k) for k = -2..nz+3 (array index k+2), the ghost levels
    past non-periodic walls as src/initial.f90:4418-4440 extends them:
    zW(-j) = zW(0) - (zW(j) - zW(0)), zW(nz+j) = 2 zW(nz-1) - zW(nz-2-j)."""
    nz = len(zw) - 1
    lo = [zw[0] - (zw[j] - zw[0]) for j in (3, 2, 1)]
    hi = [zw[nz - 1] + (zw[nz - 1] - zw[nz - 1 - (j + 1)]) for j in (1, 2, 3)]
    zW = np.concatenate([lo, zw, hi])             # k = -3..nz+3
    zPr = (zW[:-1] + zW[1:]) / 2                  # k = -2..nz+3
    return zW[1:], zPr


def staggered(cfg, fu, fv, fw, fnu=None):
    """Fields sampled at the U, V, W (and Pr) points, ghosted (lower bound -2)."""
    nx, ny = cfg.nx, cfg.ny
    i = np.arange(-2, nx + 4, dtype=np.float64)
    j = np.arange(-2, ny + 4, dtype=np.float64)
    xU, xPr = i * cfg.dx, (i - 0.5) * cfg.dx
    yV, yPr = j * cfg.dy, (j - 0.5) * cfg.dy
    zW, zPr = z_coords(np.asarray(cfg._zw_keep))

    def grid(f, x, y, z):
        return np.asfortranarray(f(x[:, None, None], y[None, :, None], z[None, None, :]))

    out = [grid(fu, xU, yPr, zPr), grid(fv, xPr, yV, zPr), grid(fw, xPr, yPr, zW)]
    if fnu is not None:
        out.append(grid(fnu, xPr, yPr, zPr))
    pts = {"U": (xU, yPr, zPr), "V": (xPr, yV, zPr), "W": (xPr, yPr, zW)}
    return out, pts


def dd(f, axis, h=1e-3):
    """Fourth-order central derivative of f(x, y, z) along axis."""
    def g(x, y, z):
        def at(s):
            a = [x, y, z]
            a[axis] = a[axis] + s
            return f(*a)
        return (8 * (at(h) - at(-h)) - (at(2 * h) - at(-2 * h))) / (12 * h)
    return g


# manufactured fields: smooth, periodic in x and y (period 2 pi)
def fu(x, y, z):
    return 1.0 + 0.5 * np.sin(x) * np.cos(y) * (1 + z) + 0.3 * z**2


def fv(x, y, z):
    return 0.4 * np.cos(x) * np.sin(y) * (1 - 0.5 * z) + 0.2 * z


def fw(x, y, z):
    return 0.3 * np.sin(x + y) * z * (2 - z) + 0.1


def fnu(x, y, z):
    return 0.01 * (1 + 0.5 * np.sin(x) * np.cos(y) + 0.2 * z)


def advection_exact(name, x, y, z):
    """-1/2 [div(u_j f) + u_j d_j f] for f the component `name`."""
    f = {"U": fu, "V": fv, "W": fw}[name]
    vel = (fu, fv, fw)
    out = 0.0
    for a in range(3):
        def flux(X, Y, Z, a=a):
            return vel[a](X, Y, Z) * f(X, Y, Z)
        out = out + dd(flux, a)(x, y, z) + vel[a](x, y, z) * dd(f, a)(x, y, z)
    return -0.5 * out


def diffusion_exact(name, x, y, z):
    """sum_j d_j (nu d_j f)."""
    f = {"U": fu, "V": fv, "W": fw}[name]
    out = 0.0
    for a in range(3):
        def flux(X, Y, Z, a=a):
            return fnu(X, Y, Z) * dd(f, a)(X, Y, Z)
        out = out + dd(flux, a, h=2e-3)(x, y, z)
    return out


def operator_errors(n, op):
    cfg = stretched_cfg(n, n)
    o = Oracle(cfg, capi.Case(cfg))
    (U, V, W, NU), pts = staggered(cfg, fu, fv, fw, fnu)
    for fid, a in ((capi.F_U, U), (capi.F_V, V), (capi.F_W, W), (capi.F_VISC, NU)):
        o.upload(fid, a)
    getattr(o, "op_" + op)()
    exact = advection_exact if op == "advection" else diffusion_exact
    errs = {}
    for name, fid in (("U", F_USTAR), ("V", F_VSTAR), ("W", F_WSTAR)):
        got = o.download(fid)
        x, y, z = pts[name]
        # interior levels whose stencils stay inside the walls: k = 2..nz-2
        ks = slice(4, 3 + n - 2)
        ex = exact(name, x[3:3 + n, None, None], y[None, 3:3 + n, None], z[None, None, ks])
        errs[name] = np.abs(got[3:3 + n, 3:3 + n, ks] - ex).max() / np.abs(ex).max()
    o.close()
    return errs


@pytest.mark.parametrize("op", ["advection", "diffusion"])
def test_variable_z_operators_converge_second_order(op):
    """Halving every spacing of the tanh grid divides the error against the
    continuous operator by about 4 (a dropped term, a wrong sign, index or
    spacing leaves an O(1) error that does not shrink)."""
    e1 = operator_errors(16, op)
    e2 = operator_errors(32, op)
    for c in "UVW":
        assert e2[c] < 0.02, (op, c, e1, e2)
        assert e1[c] / e2[c] > 3.0, (op, c, e1, e2)


def random_fields(cfg, seed):
    rng = np.random.default_rng(seed)
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    U, V, W = (np.asfortranarray(rng.uniform(-1, 1, shape)) for _ in range(3))
    U += 2.0
    NU = np.asfortranarray(0.01 * (1 + rng.uniform(0, 1, shape)))
    return U, V, W, NU


def load(o, fields):
    for fid, a in zip((capi.F_U, capi.F_V, capi.F_W, capi.F_VISC), fields):
        o.upload(fid, a)


def test_uniform_spacing_limit_matches_uniform_path():
    """With zW_k = k (dz = 1) the variable-z diffusion and wall-model
    replacement equal the uniform-grid ones; fluxp = fluxm = the uniform flux."""
    n, nz = 12, 10
    kw = dict(z0_wall=[0, 0, 0, 0, 0.01, 0], btype=[capi.PERIODIC] * 4 + [capi.NOSLIP, capi.FREESLIP])
    uni = cases.uniform_config(n, n, nz, dx=1.0, nscalars=0, **kw)
    vz = capi.make_config(**{**dict(nx=n, ny=n, nz=nz, dx=1.0, dy=1.0, dz=1.0, nu=cases.NU_DEFAULT,
                                    nu_scalar=cases.NU_DEFAULT, dt=0.05, nscalars=0, advection_method=2,
                                    scal_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN] * 2, side_scal=[0] * 6,
                                    zw=np.arange(nz + 1, dtype=np.float64)), **kw})
    f = random_fields(uni, 3)
    res = []
    for cfg in (uni, vz):
        o = Oracle(cfg, capi.Case(cfg))
        load(o, f)
        o.op_wm_fluxes()
        o.op_diffusion()
        res.append(([o.download(fid) for fid in (F_USTAR, F_VSTAR, F_WSTAR)],
                    [o.wm_list(c, 4) for c in range(3)]))
        o.close()
    (a, la), (b, lb) = res
    for c in range(3):
        sl = (slice(3, 3 + n), slice(3, 3 + n), slice(3, 3 + (nz if c < 2 else nz - 1)))
        scale = np.abs(a[c][sl]).max()
        assert np.abs(a[c][sl] - b[c][sl]).max() <= 1e-14 * scale, c
        assert len(la[c][0]) == n * n or c == 2
        assert np.array_equal(la[c][0], lb[c][0]) and np.array_equal(lb[c][0], lb[c][1])


def test_wall_replacement_budget_on_tanh_grid():
    """Summed over the domain (periodic x/y), the interior face fluxes cancel
    and the replaced wall faces leave the wall stresses: sum Us dzPr =
    sum(fluxp_bottom) dzPr(1) - sum(fluxm_top) dzPr(nz); sum Ws dzW = 0 (no
    tangential W at a z wall).  fluxp and fluxm carry one stress, scaled by the
    heights of the cells beyond and before the face."""
    n, nz = 12, 16
    cfg = stretched_cfg(n, nz, wall_model=1)
    o = Oracle(cfg, capi.Case(cfg))
    f = random_fields(cfg, 4)
    o.upload(capi.F_U, f[0])
    o.upload(capi.F_V, f[1])
    o.upload(capi.F_W, f[2])
    o.op_bound_uvw()            # periodic x/y ghosts (and wall ghosts)
    nu = f[3].copy()
    I = nu[3:3 + n, 3:3 + n, :]
    nu[0:3, 3:3 + n, :], nu[3 + n:, 3:3 + n, :] = I[n - 3:], I[:3]
    nu[:, 0:3, :], nu[:, 3 + n:, :] = nu[:, n:n + 3, :], nu[:, 3:6, :]
    o.upload(capi.F_VISC, nu)
    o.op_wm_fluxes()
    o.op_diffusion()
    zW, zPr = z_coords(np.asarray(cfg._zw_keep))
    dzPr = zW[2:nz + 4] - zW[1:nz + 3]              # dzPr(k), k = 0..nz+1
    dzW = zPr[3:3 + nz] - zPr[2:2 + nz]             # dzW(k), k = 0..nz-1
    for c, fid in ((0, F_USTAR), (1, F_VSTAR)):
        s = o.download(fid)[3:3 + n, 3:3 + n, 3:3 + nz]
        total = (s * dzPr[None, None, 1:nz + 1]).sum()
        bp, bm = o.wm_list(c, 4)
        tp, tm = o.wm_list(c, 5)
        assert len(bp) == n * n and len(tp) == n * n
        walls = bp.sum() * dzPr[1] - tm.sum() * dzPr[nz]
        assert abs(total - walls) <= 1e-12 * (np.abs(s * dzPr[None, None, 1:nz + 1]).sum()), (c, total, walls)
        # one stress per point: fluxp dzPr(beyond) == fluxm dzPr(before)
        np.testing.assert_allclose(bp * dzPr[1], bm * dzPr[0], rtol=1e-14)
        np.testing.assert_allclose(tp * dzPr[nz + 1], tm * dzPr[nz], rtol=1e-14)
    # mean flow U ~ 2 > 0 is decelerated by both walls
    assert o.wm_list(0, 4)[0].sum() < 0 and o.wm_list(0, 5)[1].sum() > 0
    s = o.download(F_WSTAR)[3:3 + n, 3:3 + n, 3:3 + nz - 1]
    assert abs((s * dzW[None, None, 1:]).sum()) <= 1e-12 * np.abs(s * dzW[None, None, 1:]).sum()
    o.close()


def test_stretched_config_validation():
    cfg = stretched_cfg(16, 16)
    with pytest.raises(RuntimeError, match="advection_method 2"):
        capi.Case(capi.make_config(**{**dict(nx=16, ny=16, nz=16, dx=1.0, dy=1.0, dz=1.0, nscalars=0,
                                             btype=[capi.PERIODIC] * 4 + [capi.NOSLIP] * 2, dt=0.01, nu=1e-3,
                                             zw=cases.channel_zw(16))}))
    with pytest.raises(RuntimeError, match="solid bodies"):
        capi.Case(cfg, cuboids=[(2.0, 4.0, 2.0, 4.0, 0.0, 0.5)])
    bad = cases.channel_zw(16)
    bad[5] = bad[4]
    with pytest.raises(RuntimeError, match="increase"):
        capi.Case(capi.make_config(**{**dict(nx=16, ny=16, nz=16, dx=1.0, dy=1.0, dz=1.0, nscalars=0,
                                             btype=[capi.PERIODIC] * 4 + [capi.NOSLIP] * 2, dt=0.01, nu=1e-3,
                                             advection_method=2, zw=bad)}))


# --------------------------------------------------------------------- GPU
def channel_pair(cfg, precision=capi.FP64, seed=2):
    case = capi.Case(cfg)
    f = cases.channel_fields(cfg, seed=seed)
    o = Oracle(cfg, case)
    s = capi.Solver(capi.with_precision(cfg, precision), case)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg, f)
    o.init()
    s.init()
    return o, s


@pytest.mark.gpu
def test_gpu_stretched_channel_fp64(require_gpu):
    from helpers import compare_state
    cfg = cases.channel_config(64, 32, 32)
    o, s = channel_pair(cfg)
    do, ds = o.step(1), s.step(1)
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc"])
    assert max(err.values()) == 0.0, err
    assert ds.max_div <= max(do.max_div, 1e-15) * 1.5 and ds.max_div < 1e-11
    o.step(9)
    s.step(9)
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc"])
    assert max(err.values()) == 0.0, err
    o.close()
    s.close()


@pytest.mark.gpu
def test_gpu_stretched_channel_bit_identical(require_gpu):
    from helpers import compare_state
    cfg = cases.channel_config(32, 16, 16)
    o, s = channel_pair(cfg, seed=7)
    o.step(3)
    s.step(3)
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc"])
    assert max(err.values()) == 0.0, err
    o.close()
    s.close()


@pytest.mark.gpu
def test_gpu_stretched_channel_fp32_refined(require_gpu):
    from helpers import velocity_rel_l2
    cfg = cases.channel_config(64, 32, 32)
    o, s = channel_pair(cfg, precision=capi.FP32_REFINED)
    for _ in range(10):
        d = s.step(1)
        assert d.max_div <= 1e-10, d.max_div
    o.step(10)
    l2 = velocity_rel_l2(s, o, cfg)
    assert l2 <= 1e-5, l2
    o.close()
    s.close()
