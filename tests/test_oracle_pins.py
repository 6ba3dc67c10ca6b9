"""Pins of the CPU oracle: closed forms, invariants and independent
formulations (the reference ships no tests or golden vectors, and its FFT
Poisson library is not in the tree).

Each test names the reference procedure whose restatement it pins.
"""
import numpy as np
import pytest

from elmm_amd import capi, cases
from oracle_py import F_SADV, F_USTAR, F_VSTAR, Oracle, det_log

RNG = np.random.default_rng(12345)


def small_cfg(n=16, nz=12, **kw):
    base = dict(nx=n, ny=n, nz=nz, dx=1.0, dt=0.05)
    base.update(kw)
    return cases.uniform_config(base.pop("nx"), base.pop("ny"), base.pop("nz"), **base)


def make(cfg):
    case = capi.Case(cfg)
    return Oracle(cfg, case), case


def ghosted(cfg):
    return np.zeros((cfg.nx + 6, cfg.ny + 6, cfg.nz + 6), order="F")


def periodic_xy(a, n0, n1):
    """Fill x/y ghosts of a ghosted array by periodic wrap (lower bound -2)."""
    I = a[3:3 + n0, 3:3 + n1, :]
    a[0:3, 3:3 + n1, :] = I[n0 - 3:n0, :, :]
    a[3 + n0:6 + n0, 3:3 + n1, :] = I[0:3, :, :]
    a[:, 0:3, :] = a[:, n1:n1 + 3, :]
    a[:, 3 + n1:6 + n1, :] = a[:, 3:6, :]
    return a


# ----------------------------------------------------------- Poisson solver
def laplacian_fd2(phi, dx, dy, dz):
    """7-point operator of Residue_variable_z (src/poisson.f90:23-60) on a
    uniform grid: periodic x/y, Neumann (mirror ghost) in z."""
    lap = (np.roll(phi, 1, 0) - 2 * phi + np.roll(phi, -1, 0)) / dx**2
    lap += (np.roll(phi, 1, 1) - 2 * phi + np.roll(phi, -1, 1)) / dy**2
    pz = np.concatenate([phi[:, :, :1], phi, phi[:, :, -1:]], axis=2)
    lap += (pz[:, :, :-2] - 2 * phi + pz[:, :, 2:]) / dz**2
    return lap


def test_poisson_residual_random_rhs():
    cfg = small_cfg(16, 12, dx=0.5)
    o, _ = make(cfg)
    rhs = RNG.standard_normal((16, 16, 12))
    rhs -= rhs.mean()
    phi = o.op_poisson(rhs)
    res = laplacian_fd2(phi, 0.5, 0.5, 0.5) - rhs
    assert np.abs(res).max() < 1e-12 * np.abs(rhs).max()


def test_poisson_eigenfunction():
    """Discrete eigenfunction cos(2 pi m x/nx) cos(pi q (k-1/2)/nz) with
    eigenvalue -4 sin^2(pi m/nx)/dx^2 - 4 sin^2(pi q/(2 nz))/dz^2."""
    cfg = small_cfg(16, 12)
    o, _ = make(cfg)
    i = np.arange(16)[:, None, None]
    k = np.arange(1, 13)[None, None, :]
    m, q = 3, 5
    f = np.cos(2 * np.pi * m * i / 16) * np.cos(np.pi * q * (k - 0.5) / 12) * np.ones((1, 16, 1))
    lam = -4 * np.sin(np.pi * m / 16) ** 2 - 4 * np.sin(np.pi * q / 24) ** 2
    phi = o.op_poisson(f)
    assert np.abs(phi - f / lam).max() < 1e-13


def test_poisson_matches_dense_solve():
    """Independent dense least-squares solve on an 8x8x6 grid: same gradient
    (the solution is unique up to a constant)."""
    cfg = small_cfg(8, 6)
    o, _ = make(cfg)
    rhs = RNG.standard_normal((8, 8, 6))
    rhs -= rhs.mean()
    phi = o.op_poisson(rhs)
    n = rhs.size
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1
        A[:, c] = laplacian_fd2(e.reshape(8, 8, 6, order="F"), 1, 1, 1).ravel(order="F")
    x = np.linalg.lstsq(A, rhs.ravel(order="F"), rcond=None)[0].reshape(8, 8, 6, order="F")
    d1, d2 = phi - phi.mean(), x - x.mean()
    assert np.abs(d1 - d2).max() < 1e-12 * np.abs(d2).max()


@pytest.mark.parametrize("nx,ny,nz", [(24, 18, 10), (96, 36, 8), (48, 64, 6), (54, 12, 5)])
def test_poisson_mixed_radix_residual(nx, ny, nz):
    """The mixed-radix (2^a 3^b) transform path: residual of the FD2
    operator on a random right-hand side, non-square grids."""
    cfg = cases.uniform_config(nx, ny, nz, dx=0.5)
    o, _ = make(cfg)
    rhs = RNG.standard_normal((nx, ny, nz))
    rhs -= rhs.mean()
    phi = o.op_poisson(rhs)
    res = laplacian_fd2(phi, 0.5, 0.5, 0.5) - rhs
    assert np.abs(res).max() < 1e-12 * np.abs(rhs).max()


def test_poisson_mixed_radix_eigenfunction():
    """sin/cos mode on a 48 x 36 x 10 grid (radix-3 stages in both
    directions): the solution is the mode over its eigenvalue."""
    nx, ny, nz = 48, 36, 10
    cfg = cases.uniform_config(nx, ny, nz, dx=1.0)
    o, _ = make(cfg)
    i = np.arange(nx)[:, None, None]
    j = np.arange(ny)[None, :, None]
    k = np.arange(1, nz + 1)[None, None, :]
    m, p, q = 5, 7, 3
    f = np.sin(2 * np.pi * m * i / nx) * np.cos(2 * np.pi * p * j / ny) * np.cos(np.pi * q * (k - 0.5) / nz)
    lam = -4 * np.sin(np.pi * m / nx) ** 2 - 4 * np.sin(np.pi * p / ny) ** 2 - 4 * np.sin(np.pi * q / (2 * nz)) ** 2
    phi = o.op_poisson(f)
    assert np.abs(phi - f / lam).max() < 1e-13


def test_poisson_mixed_radix_matches_dense_solve():
    cfg = cases.uniform_config(6, 12, 4, dx=1.0)
    o, _ = make(cfg)
    rhs = RNG.standard_normal((6, 12, 4))
    rhs -= rhs.mean()
    phi = o.op_poisson(rhs)
    n = rhs.size
    A = np.zeros((n, n))
    for c in range(n):
        e = np.zeros(n)
        e[c] = 1
        A[:, c] = laplacian_fd2(e.reshape(6, 12, 4, order="F"), 1, 1, 1).ravel(order="F")
    x = np.linalg.lstsq(A, rhs.ravel(order="F"), rcond=None)[0].reshape(6, 12, 4, order="F")
    d1, d2 = phi - phi.mean(), x - x.mean()
    assert np.abs(d1 - d2).max() < 1e-12 * np.abs(d2).max()


# ------------------------------------------------------------- Smagorinsky
def test_smagorinsky_linear_shear():
    """U = a*z gives S13 = a/2, |S| = |a|, nu_t = (Cs*Delta)^2 |a| + nu
    (NuSmag/StrainIJ/Strainu, src/subgrid.f90:55-67, 170-209)."""
    cfg = small_cfg(8, 10, nscalars=0, wall_model=0)
    o, _ = make(cfg)
    a = 0.37
    U = ghosted(cfg)
    zpr = (np.arange(-2, cfg.nz + 4) - 0.5) * 1.0
    U[:, :, :] = a * zpr[None, None, :]
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    o.op_sgs()
    visc = o.download(capi.F_VISC)[3:-3, 3:-3, 3:-3]
    expect = (0.122 * 1.0) ** 2 * a + cfg.nu
    assert np.abs(visc - expect).max() < 1e-15


def test_smagorinsky_constant_field_is_molecular():
    cfg = small_cfg(8, 10, nscalars=0, wall_model=0)
    o, _ = make(cfg)
    U = ghosted(cfg) + 1.3
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg) - 0.4)
    o.upload(capi.F_W, ghosted(cfg))
    o.op_sgs()
    visc = o.download(capi.F_VISC)[3:-3, 3:-3, 3:-3]
    assert np.all(visc == cfg.nu)


# ------------------------------------------------------- CD4 advection
def morinishi_x(u, dx):
    """-[(9/8) d1(U4 u1) - (1/8) d3(U4 u3)] for u(x) on a periodic line
    (Morinishi et al. 1998 divergence form), u at faces i."""
    def r(a, s):
        return np.roll(a, -s)
    # face values at cell centres c = i+1/2 between u(i) and u(i+1)
    U4 = (9 * (r(u, 0) + r(u, 1)) - (r(u, -1) + r(u, 2))) / 16  # 4th order advecting velocity at i+1/2
    u1 = (r(u, 0) + r(u, 1)) / 2
    U4b = (9 * (r(u, 0) + r(u, 1)) - (r(u, -1) + r(u, 2))) / 16
    # at i+3/2 ... flux for the 3-dx stencil uses u3 = (u(i-1)+u(i+2))/2 at centre i+1/2
    u3 = (r(u, -1) + r(u, 2)) / 2
    F1 = U4 * u1            # located at i+1/2
    F3 = U4b * u3           # located at i+1/2
    d1 = (F1 - r(F1, -1)) / dx                    # F1(i+1/2) - F1(i-1/2)
    d3 = (r(F3, 1) - r(F3, -2)) / (3 * dx)        # F3(i+3/2) - F3(i-3/2)
    return -((9 / 8) * d1 - (1 / 8) * d3)


def test_cd4_self_advection_matches_morinishi():
    """CD4divU x term (src/advection.f90:357-394) against the textbook form."""
    cfg = small_cfg(16, 8, nscalars=0, wall_model=0)
    o, _ = make(cfg)
    u = RNG.standard_normal(16)
    U = ghosted(cfg)
    U[3:19, :, :] = u[:, None, None]
    U = periodic_xy(U, 16, 16)
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    o.op_advection()
    us = o.download(F_USTAR)[3:19, 3:19, 3:11]
    ref = morinishi_x(u, 1.0)
    assert np.abs(us - ref[:, None, None]).max() < 1e-14 * max(1, np.abs(ref).max())


def test_cd4_cross_advection_matches_morinishi():
    """U advected by V in y: 4th-order interpolation of V to the U point in x,
    then the divergence form in y (CD4divU y pass, src/advection.f90:398-436)."""
    cfg = small_cfg(16, 8, nscalars=0, wall_model=0)
    o, _ = make(cfg)
    u = RNG.standard_normal((16, 16))
    v = RNG.standard_normal((16, 16))
    U = ghosted(cfg)
    V = ghosted(cfg)
    U[3:19, 3:19, :] = u[:, :, None]
    V[3:19, 3:19, :] = v[:, :, None]
    U, V = periodic_xy(U, 16, 16), periodic_xy(V, 16, 16)
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, V)
    o.upload(capi.F_W, ghosted(cfg))
    o.op_advection()
    us = o.download(F_USTAR)[3:19, 3:19, 5]

    def r(a, si, sj):
        return np.roll(np.roll(a, -si, 0), -sj, 1)
    # V interpolated to x = xU(i) at (i, j+1/2)
    V4 = (9 * (v + r(v, 1, 0)) - (r(v, -1, 0) + r(v, 2, 0))) / 16
    u1 = (u + r(u, 0, 1)) / 2       # at j+1/2
    u3 = (r(u, 0, -1) + r(u, 0, 2)) / 2
    F1, F3 = V4 * u1, V4 * u3       # at (i, j+1/2)
    d1 = F1 - r(F1, 0, -1)
    d3 = (r(F3, 0, 1) - r(F3, 0, -2)) / 3
    ref = -((9 / 8) * d1 - (1 / 8) * d3) + morinishi_x_2d(u)
    assert np.abs(us - ref).max() < 1e-13


def morinishi_x_2d(u):
    return np.stack([morinishi_x(u[:, j], 1.0) for j in range(u.shape[1])], axis=1)


def test_cd4_constant_field_is_zero():
    cfg = small_cfg(8, 8, nscalars=0, wall_model=0)
    o, _ = make(cfg)
    o.upload(capi.F_U, ghosted(cfg) + 2.5)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    o.op_advection()
    assert np.all(o.download(F_USTAR)[3:-3, 3:-3, 3:-3] == 0)
    assert np.all(o.download(F_VSTAR)[3:-3, 3:-3, 3:-3] == 0)


# ------------------------------------------------------------ diffusion
def test_diffusion_fourier_mode():
    """Constant nu and U = sin(2 pi i/n): nu (2 cos(2 pi/n) - 2) U / dx^2
    (MomentumDiffusion_nobranch_2ord, src/momentum_diffusion.f90:169-207)."""
    n = 16
    cfg = small_cfg(n, 8, nscalars=0, wall_model=0, dx=0.5)
    o, _ = make(cfg)
    x = np.arange(-2, n + 4)
    U = ghosted(cfg) + np.sin(2 * np.pi * x / n)[:, None, None]
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    nu = 0.013
    o.upload(capi.F_VISC, ghosted(cfg) + nu)
    o.op_diffusion()
    us = o.download(F_USTAR)[3:3 + n, 3:3 + n, 3:-3]
    expect = nu * (2 * np.cos(2 * np.pi / n) - 2) / 0.25 * np.sin(2 * np.pi * np.arange(1, n + 1) / n)
    assert np.abs(us - expect[:, None, None]).max() < 1e-15


def test_wall_model_log_law_flux():
    """Uniform U=1 over a rough bottom: the bottom face flux becomes
    -(0.41/ln(d/z0))^2 / dz (WMRoughUstar, src/wallmodels.f90:1330;
    flux sign and replacement, src/wallmodels.f90:2311-2321 and
    wmfluxes-nobranch-inc.f90)."""
    cfg = small_cfg(8, 8, nscalars=0)
    o, _ = make(cfg)
    U = ghosted(cfg)
    U[3:11, 3:11, 3:11] = 1.0
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    o.op_bound_uvw()
    o.op_sgs()                       # computes the wall-model fluxes
    nu = 1e-3
    o.upload(capi.F_VISC, ghosted(cfg) + nu)
    o.op_diffusion()
    us = o.download(F_USTAR)
    ustar = 0.41 / np.log(0.5 / 0.01)
    assert np.allclose(us[3:11, 3:11, 3], -ustar**2, rtol=0, atol=1e-15)
    # away from the wall the uniform field does not diffuse
    assert np.abs(us[3:11, 3:11, 4:10]).max() < 1e-18


# --------------------------------------------------------------- scalars
def test_kappa_constant_scalar_no_flux_divergence():
    cfg = small_cfg(8, 8)
    o, _ = make(cfg)
    U, V, W, Pr, S = cases.initial_fields(cfg)
    S[0][:] = 3.0
    for f, a in ((capi.F_U, U), (capi.F_V, V), (capi.F_W, W), (capi.F_SCALAR, S[0])):
        o.upload(f, a)
    o.op_bound_uvw()
    # divergence-free? not required for a constant scalar: the x/y/z flux
    # differences of a constant field are u(i)-u(i-1) times the constant
    o.op_kappa(0)
    sadv = o.download(F_SADV)[3:-3, 3:-3, 3:-3]
    Ui, Vi, Wi = (o.download(c) for c in (0, 1, 2))
    div = (Ui[3:-3, 3:-3, 3:-3] - Ui[2:-4, 3:-3, 3:-3]) + (Vi[3:-3, 3:-3, 3:-3] - Vi[3:-3, 2:-4, 3:-3]) + \
          (Wi[3:-3, 3:-3, 3:-3] - Wi[3:-3, 3:-3, 2:-4])
    assert np.abs(sadv + 3.0 * div).max() < 1e-13


def flux_limiter(r):
    return np.maximum(0, np.minimum(2 * r, np.minimum(2, (1 + 2 * r) / 3)))


def test_kappa_limiter_closed_form():
    """1-D monotone profile, u>0: flux = u (S_i + (S_i - S_{i-1}) phi(r)/2)
    with r = (S_{i+1}-S_i)/(S_i-S_{i-1}) (src/scalar_advection.f90:128-157)."""
    cfg = small_cfg(16, 8)
    o, _ = make(cfg)
    U = ghosted(cfg) + 0.7
    x = np.arange(-2, 20)
    s = np.exp(0.3 * np.sin(2 * np.pi * x / 16))
    S = ghosted(cfg) + s[:, None, None]
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    o.upload(capi.F_SCALAR, S)
    o.op_kappa(0)
    sadv = o.download(F_SADV)[3:19, 5, 5]
    eps = float(np.float32(1e-8))

    def flux(f):  # face f between cells f and f+1 (indices into s with offset 2)
        sr = s[f + 1 + 2] - s[f + 2]
        sl = s[f + 2] - s[f - 1 + 2]
        ph = flux_limiter((sr + eps * np.sign(sl)) / (sl + eps * np.sign(sl)))
        return 0.7 * (s[f + 2] + (s[f + 2] - s[f - 1 + 2]) * ph / 2)
    ref = np.array([flux(i - 1) - flux(i) for i in range(1, 17)])
    assert np.abs(sadv - ref).max() < 1e-14


# --------------------------------------------------------- whole-step pins
def run_steps(cfg, nsteps, scalar=None, fields=None):
    o, case = make(cfg)
    f = fields or cases.initial_fields(cfg)
    if scalar is not None:
        f[4][0][3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz] = scalar
    cases.upload_initial(o, cfg, f)
    o.init()
    diags = [o.step(1) for _ in range(nsteps)]
    return o, diags


def test_projection_is_divergence_free():
    cfg = small_cfg(16, 12)
    o, diags = run_steps(cfg, 3)
    for d in diags:
        assert d.max_div < 1e-13


def test_scalar_mass_conserved():
    """Periodic x/y, zero-flux z walls, no sources: total scalar conserved."""
    cfg = small_cfg(16, 12)
    s0 = 1.0 + cases.perturbation((16, 16, 12), 0.5, 3, 9)
    o, _ = run_steps(cfg, 5, scalar=s0)
    s = o.download(capi.F_SCALAR)[3:19, 3:19, 3:15]
    assert abs(s.sum() - s0.sum()) < 1e-12 * s0.sum()


def test_uniform_pressure_gradient_forcing():
    """Inviscid uniform flow driven by a uniform pressure gradient: U(t) =
    U0 - dp/dx t for any RK3 with sum(2 alpha) = 1 (PressureGrad,
    src/dynamics.f90:83-92; RK3 coefficients src/parameters.f90:312-322)."""
    cfg = small_cfg(8, 8, nscalars=0, wall_model=0, sgs_model=0, nu=0.0, pr_gradient_x=0.2,
                    top=capi.FREESLIP)
    cfg.btype[4] = capi.FREESLIP
    o, _ = make(cfg)
    U = ghosted(cfg)
    U[3:11, 3:11, 3:11] = 1.0
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    o.upload(capi.F_PR, ghosted(cfg))
    o.init()
    o.step(10)
    u = o.download(capi.F_U)[3:11, 3:11, 3:11]
    assert np.abs(u - (1.0 - 0.2 * 10 * 0.05)).max() < 1e-13


def test_rk3_low_storage_consistency():
    """alpha_s = (beta_s + rho_s)/2 and sum(2 alpha) = 1 for the constants
    the step uses, checked through a linear forcing that sees each stage."""
    alpha = np.array([4 / 15, 1 / 15, 1 / 6])
    beta = np.array([8 / 15, 5 / 12, 3 / 4])
    rho = np.array([0, -17 / 60, -5 / 12])
    assert np.allclose(alpha, (beta + rho) / 2)
    assert abs(2 * alpha.sum() - 1) < 1e-15


def test_wall_model_smooth_wall_law_fixed_point():
    """Smooth wall (z0 = 0, the law used on obstacle walls): the u* behind
    the replaced bottom flux satisfies the log law u/u* = ln(d u*/nu)/0.4 +
    5.5 to the iteration tolerance 1e-4 (WMFlatUstar, src/wallmodels.f90:
    1190-1240)."""
    cfg = small_cfg(8, 8, nscalars=0, z0_bottom=0.0, nu=1e-5)
    o, _ = make(cfg)
    U = ghosted(cfg)
    U[3:11, 3:11, 3:11] = 2.0
    o.upload(capi.F_U, U)
    o.upload(capi.F_V, ghosted(cfg))
    o.upload(capi.F_W, ghosted(cfg))
    o.op_bound_uvw()
    o.op_sgs()
    o.upload(capi.F_VISC, ghosted(cfg) + 1e-3)
    o.op_diffusion()
    us2 = -o.download(F_USTAR)[3:11, 3:11, 3]      # u*^2 / dz with dz = 1
    ustar = np.sqrt(us2)
    zplus = 0.5 * ustar / 1e-5
    assert np.all(zplus > 30)
    law = np.log(zplus) / 0.4 + 5.5
    assert np.allclose(2.0 / ustar, law, rtol=3e-4, atol=0)


def _bound_uvw_random(cfg, seed):
    o, _ = make(cfg)
    rng = np.random.default_rng(seed)
    arrs = []
    for c in range(3):
        a = ghosted(cfg)
        a[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz] = rng.standard_normal((cfg.nx, cfg.ny, cfg.nz))
        if c == 2:
            a[:, :, 3 + cfg.nz - 1] = 0.0    # W interior is 1..nz-1 with walls
        arrs.append(a)
        o.upload(c, a)
    o.op_bound_uvw()
    return [o.download(c) for c in range(3)], arrs


def test_ghost_fill_mirror_rules():
    """BoundUVW on random interiors (src/boundaries.f90:398-1791): periodic
    x/y copies; no-slip bottom: tangential odd mirror U(1-m) = -U(m), wall
    normal W(0) = 0 and W(-m) = -W(m); free-slip top: the tangential ghosts
    repeat the top value U(nz+m) = U(nz) (src/boundaries.f90:1743-1750),
    normal W(nz) = 0, W(nz+m) = -W(nz-m)."""
    cfg = small_cfg(8, 6, nscalars=0)
    (U, V, W), _ = _bound_uvw_random(cfg, 7)
    n, nz = 8, 6
    I = lambda a, k: a[3:3 + n, 3:3 + n, 2 + k]   # lower bound -2
    for A in (U, V, W):
        assert np.array_equal(A[0:3, 3:3 + n, 3:3 + nz], A[n:n + 3, 3:3 + n, 3:3 + nz])
        assert np.array_equal(A[3 + n:6 + n, 3:3 + n, 3:3 + nz], A[3:6, 3:3 + n, 3:3 + nz])
        assert np.array_equal(A[3:3 + n, 0:3, 3:3 + nz], A[3:3 + n, n:n + 3, 3:3 + nz])
    for A in (U, V):
        for m in (1, 2, 3):
            assert np.array_equal(I(A, 1 - m), -I(A, m))            # bottom no-slip
            assert np.array_equal(I(A, nz + m), I(A, nz))            # top free-slip
    assert not I(W, 0).any() and not I(W, nz).any()
    for m in (1, 2):
        assert np.array_equal(I(W, -m), -I(W, m))
        assert np.array_equal(I(W, nz + m), -I(W, nz - m))


def test_ghost_fill_idempotent():
    cfg = small_cfg(8, 6, nscalars=0)
    o, _ = make(cfg)
    (U, V, W), _ = _bound_uvw_random(cfg, 8)
    for c, a in enumerate((U, V, W)):
        o.upload(c, a)
    o.op_bound_uvw()
    for c, a in enumerate((U, V, W)):
        assert np.array_equal(o.download(c), a)


def test_point_source_mass_budget():
    """Fluid at rest, zero-flux walls, one point source of F units/s: the
    scalar content grows by exactly F dt per step because the three RK stage
    weights add to one (point_source, src/pointsources.f90:20-66; FluxKernel,
    src/volumesources.f90:757-771; scalar RK, src/scalars.f90:279-283)."""
    cfg = small_cfg(8, 6, wall_model=0, sgs_model=0)
    cfg.btype[4] = capi.FREESLIP
    case = capi.Case(cfg, sources=[(3.3, 4.7, 2.2, 0.75, 0)])
    o = Oracle(cfg, case)
    z = ghosted(cfg)
    for c in range(3):
        o.upload(c, z)
    o.upload(capi.F_PR, z)
    o.upload(capi.F_SCALAR, z)
    o.init()
    o.step(4)
    s = o.download(capi.F_SCALAR)[3:11, 3:11, 3:9]
    assert abs(s.sum() - 4 * 0.75 * 0.05) < 1e-14


def test_pressure_reference_top_mean_zero():
    """FixPressureToReference with the top boundary as reference
    (src/pressure.f90:994-1229): after every step the mean of Pr over the
    top layer is zero to rounding."""
    cfg = small_cfg(16, 12)
    o, _ = run_steps(cfg, 2)
    pr = o.download(capi.F_PR)[3:19, 3:19, 3 + 11]
    assert abs(pr.mean()) < 1e-15 * max(1.0, np.abs(pr).max())


# ------------------------------------------------------- smooth-wall log
def test_det_log_within_one_ulp_of_libm():
    """The logarithm of WMFlatUstar's log-law branches (src/wallmodels.f90:
    1312-1391) is restated without libm so CPU and GPU agree bit for bit; it
    stays within one ulp of the C library's log, is exact at 1, and gives
    k ln 2 for powers of two."""
    import math
    rng = np.random.default_rng(7)
    xs = np.concatenate([np.exp(rng.uniform(-700, 700, 20000)), rng.uniform(0.5, 2.0, 20000),
                         rng.uniform(5.0, 1e4, 20000)])
    worst = max(abs(det_log(x) - math.log(x)) / np.spacing(abs(math.log(x))) for x in xs if x != 1.0)
    assert worst <= 1.0
    assert det_log(1.0) == 0.0
    for k in (-1000, -3, 1, 7, 1000):
        assert abs(det_log(2.0 ** k) - k * math.log(2.0)) <= np.spacing(abs(k * math.log(2.0)))
