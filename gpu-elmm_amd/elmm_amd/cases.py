"""Case definitions of the benchmark configurations and their seeded
initial fields.

Random perturbations come from numpy's Philox counter-based generator keyed
by the case seed (platform independent) and are rounded to float32, so every
initial value is exactly representable in the reference's float32 VTK
restart format (src/initial.f90:3674-3708).
"""
import numpy as np

from . import capi

NU_DEFAULT = 1.0 / 70000                      # molecular_viscosity, src/parameters.f90:128
NU_SCALAR_DEFAULT = (1.0 / 70000) / float(np.float32(0.7))  # 1._knd/70000/0.7 (default-real 0.7)


def perturbation(shape, amplitude, seed, stream):
    """Uniform(-a, a) values rounded to float32."""
    g = np.random.Generator(np.random.Philox(key=[seed, stream]))
    u = g.random(shape, dtype=np.float64)
    return (amplitude * (2.0 * u - 1.0)).astype(np.float32).astype(np.float64)


def uniform_config(nx, ny, nz, dx=1.0, dt=0.05, nscalars=1, top=capi.FREESLIP, z0_bottom=0.01, **kw):
    """Periodic x/y, log-law wall at the bottom, top as given (configs 1/3)."""
    args = dict(
        nx=nx, ny=ny, nz=nz, dx=dx, dy=dx, dz=dx,
        btype=[capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.NOSLIP, top],
        z0_wall=[0, 0, 0, 0, z0_bottom, 0],
        scal_btype=[capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.NEUMANN, capi.NEUMANN],
        side_scal=[0] * 6,
        nu=NU_DEFAULT, nu_scalar=NU_SCALAR_DEFAULT, dt=dt, nscalars=nscalars, cfl_check=1)
    args.update(kw)
    return capi.make_config(**args)


def config1(n=64, **kw):
    """Uniform 64^3 case: dx=1 m, periodic x/y, free-slip top, no-slip log-law
    bottom z0=0.01 m, CD4 advection, Smagorinsky, FFT pressure, dt=0.05 s."""
    return uniform_config(n, n, n, **kw)


def channel_zw(nz, lz=2.0, gamma=1.5):
    """Wall-clustered tanh grid zW_k = h (1 - tanh(gamma (1 - 2k/nz)) / tanh(gamma)),
    h = lz/2, k = 0..nz (walls at 0 and lz exactly)."""
    k = np.arange(nz + 1, dtype=np.float64)
    zw = (lz / 2) * (1.0 - np.tanh(gamma * (1.0 - 2.0 * k / nz)) / np.tanh(gamma))
    zw[0], zw[-1] = 0.0, lz
    return zw


def channel_config(nx=256, ny=128, nz=128, lx=2 * np.pi, ly=np.pi, lz=2.0, gamma=1.5, nu=1e-3, dt=2e-4, **kw):
    """Turbulent channel (Re_tau = u_tau h / nu = 1000 for u_tau = 1, h = 1):
    periodic x/y, smooth-wall (z0 = 0) wall model at both z walls, tanh
    stretched z, variable-z CD2 advection, driven by pr_gradient_x = -1, no
    scalars."""
    args = dict(
        nx=nx, ny=ny, nz=nz, dx=lx / nx, dy=ly / ny, dz=lz / nz,
        btype=[capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.NOSLIP, capi.NOSLIP],
        z0_wall=[0] * 6,
        scal_btype=[capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.PERIODIC, capi.NEUMANN, capi.NEUMANN],
        side_scal=[0] * 6, nu=nu, nu_scalar=nu, dt=dt, nscalars=0, cfl_check=1,
        advection_method=2, pr_gradient_x=-1.0, zw=channel_zw(nz, lz, gamma))
    args.update(kw)
    return capi.make_config(**args)


def channel_fields(cfg, seed=2, amp=0.05, utau=1.0):
    """Law-of-the-wall start: U(d) = u_tau d+ for d+ < 11, else
    u_tau (ln(d+)/0.41 + 5.5), d the distance to the nearer wall; U, V and W
    carry uniform noise of amp times the local profile value."""
    zw = np.asarray(cfg._zw_keep)
    lz = zw[-1]

    def law(d):
        dp = d * utau / cfg.nu
        return utau * np.where(dp < 11.0, dp, np.log(np.maximum(dp, 1e-300)) / 0.41 + 5.5)

    zpr = (zw[:-1] + zw[1:]) / 2
    prof = law(np.minimum(zpr, lz - zpr))
    profw = law(np.minimum(zw[1:-1], lz - zw[1:-1]))
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    shape = (nx + 6, ny + 6, nz + 6)
    U, V, W = (np.zeros(shape, order="F") for _ in range(3))
    U[3:3 + nx, 3:3 + ny, 3:3 + nz] = prof * (1.0 + perturbation((nx, ny, nz), amp, seed, 0))
    V[3:3 + nx, 3:3 + ny, 3:3 + nz] = prof * perturbation((nx, ny, nz), amp, seed, 1)
    W[3:3 + nx, 3:3 + ny, 3:3 + nz - 1] = profw * perturbation((nx, ny, nz - 1), amp, seed, 2)
    return U, V, W, np.zeros(shape, order="F"), []


def initial_fields(cfg, u0=1.0, amp=0.1, seed=1, v0=0.0):
    """U = u0 + d, V = v0 + d, W = d on each component's interior; Pr = 0;
    scalars = 0.  Arrays are ghosted (lower bound -2), Fortran order."""
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    n = [[0] * 3 for _ in range(3)]
    per = [cfg.btype[1] == capi.PERIODIC, cfg.btype[3] == capi.PERIODIC, cfg.btype[5] == capi.PERIODIC]
    for c in range(3):
        for d, N in enumerate((cfg.nx, cfg.ny, cfg.nz)):
            n[c][d] = N if (d != c or per[d]) else N - 1
    out = []
    for c, base in enumerate((u0, v0, 0.0)):
        a = np.zeros(shape, order="F")
        nc = n[c]
        d = perturbation((nc[0], nc[1], nc[2]), amp, seed, c)
        a[3:3 + nc[0], 3:3 + nc[1], 3:3 + nc[2]] = base + d
        out.append(a)
    pr = np.zeros(shape, order="F")
    scal = [np.zeros(shape, order="F") for _ in range(cfg.nscalars)]
    return out[0], out[1], out[2], pr, scal


def upload_initial(target, cfg, fields):
    """Upload (U, V, W, Pr, [scalars]) to a Solver or an oracle wrapper."""
    U, V, W, Pr, scal = fields
    target.upload(capi.F_U, U)
    target.upload(capi.F_V, V)
    target.upload(capi.F_W, W)
    target.upload(capi.F_PR, Pr)
    for s, a in enumerate(scal):
        target.upload(capi.F_SCALAR + s, a)
