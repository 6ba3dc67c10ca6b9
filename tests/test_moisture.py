"""Moisture (enable_moisture, with enable_buoyancy): its RK3 stage with its
own boundary conditions (src/scalars.f90:202-208, BoundMoisture
src/scalarboundaries.f90:311-323), Moisture = moisture_ref in solid cells
after it, and the virtual temperature Tv = T (1 + 0.61 q) in the buoyancy
force (src/dynamics.f90:620-641, 695-710, 766-770), added to the W faces in
the reference's two passes over odd and even k.

CPU pins of the oracle: hydrostatic balance at rest with a linear
temperature and uniform moisture (the pressure steps equal dz A (mean Tv -
T_ref) face by face; a dropped 0.61 fails it); the force within rounding of
its closed form W + A/2 (Tv(k) + Tv(k+1)) - A T_ref and equal bit for bit to
a literal emulation of the two-pass loop; the solid-cell values against a
brute force over Prtype; the moisture's ghost cells against the
Dirichlet/Neumann/periodic rules.
GPU: bit-identical to the oracle (FP64) with obstacles, after the initial
tail and 1 and 3 steps; the precision form within 1e-5.
"""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state, velocity_rel_l2
from oracle_py import F_WSTAR, Oracle
from test_buoyancy import OBST, temperature_bcs, temperature_field

Q_BOT, Q_TOP = 0.012, 0.008


def moist_config(nx=32, ny=32, nz=16, **kw):
    args = dict(dx=1.0, dt=0.05, nscalars=1, enable_buoyancy=1, grav_acc=9.81, temperature_ref=295.0,
                **temperature_bcs(nz), enable_moisture=1, moisture_ref=0.01,
                moist_btype=[capi.PERIODIC] * 4 + [capi.DIRICHLET, capi.NEUMANN],
                side_moist=[0, 0, 0, 0, Q_BOT, 0.0])
    args.update(kw)
    return cases.uniform_config(nx, ny, nz, **args)


def moisture_field(cfg, seed=7, amp=0.001, q0=0.01):
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    q = np.full(shape, q0, order="F")
    if amp:
        g = np.random.Generator(np.random.Philox(key=[seed, 9]))
        q[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz] += g.uniform(-amp, amp, (cfg.nx, cfg.ny, cfg.nz))
    return q


def test_oracle_moist_hydrostatic_balance():
    """At rest, linear temperature, uniform moisture q0 with zero-gradient
    walls: the velocity stays zero, Pr(k+1) - Pr(k) = dz A (Tv(k+1) + Tv(k))/2
    - dz A T_ref with Tv = T (1 + 0.61 q0), and the fields are unchanged."""
    q0 = 0.01
    cfg = moist_config(moist_btype=[capi.PERIODIC] * 4 + [capi.NEUMANN, capi.NEUMANN], side_moist=[0] * 6)
    o = Oracle(cfg, capi.Case(cfg))
    zero = np.zeros((cfg.nx + 6, cfg.ny + 6, cfg.nz + 6), order="F")
    for f in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR, capi.F_SCALAR):
        o.upload(f, zero)
    T0 = temperature_field(cfg, amp=0.0)
    o.upload(capi.F_TEMPERATURE, T0)
    o.upload(capi.F_MOISTURE, moisture_field(cfg, amp=0.0, q0=q0))
    o.init()
    o.step(3)
    for c in (capi.F_U, capi.F_V, capi.F_W):
        assert np.abs(o.download(c)).max() <= 1e-12
    s = (slice(3, 3 + cfg.nx), slice(3, 3 + cfg.ny), slice(3, 3 + cfg.nz))
    Pr = o.download(capi.F_PR)[s]
    T = o.download(capi.F_TEMPERATURE)[s]
    A = cfg.grav_acc / cfg.temperature_ref
    tv = T * (1 + 0.61 * q0)
    want = cfg.dz * A * ((tv[:, :, 1:] + tv[:, :, :-1]) / 2 - cfg.temperature_ref)
    got = np.diff(Pr, axis=2)
    assert np.abs(got - want).max() <= 1e-9 * np.abs(want).max()
    # without the moisture factor the balance is off by far more
    dry = cfg.dz * A * ((T[:, :, 1:] + T[:, :, :-1]) / 2 - cfg.temperature_ref)
    assert np.abs(got - dry).max() > 1e-3
    assert np.abs(o.download(capi.F_MOISTURE)[s] - q0).max() <= 1e-15
    o.close()


def test_oracle_moist_force_order_and_closed_form():
    cfg = moist_config()
    o = Oracle(cfg, capi.Case(cfg))
    T = temperature_field(cfg, amp=0.5)
    q = moisture_field(cfg, amp=0.003)
    o.upload(capi.F_TEMPERATURE, T)
    o.upload(capi.F_MOISTURE, q)
    ws = np.asfortranarray(np.random.default_rng(5).uniform(-1, 1, T.shape))
    o.upload(F_WSTAR, ws)
    o.op_buoyancy()
    got = o.download(F_WSTAR)
    A = cfg.grav_acc / cfg.temperature_ref
    A2 = A / 2.0
    tv = T * (1.0 + 0.61 * q)
    nx, ny, wn = cfg.nx, cfg.ny, cfg.nz - 1   # W faces 1..nz-1 between the z walls
    s = (slice(3, 3 + nx), slice(3, 3 + ny), slice(3, 3 + wn))
    up = (slice(3, 3 + nx), slice(3, 3 + ny), slice(4, 4 + wn))
    closed = ws[s] + A2 * (tv[s] + tv[up]) - A * cfg.temperature_ref
    assert np.abs(got[s] - closed).max() <= 1e-12 * np.abs(closed).max()
    # the two passes literally, in Python floats (odd k, then even k)
    want = ws.copy()
    for start in (1, 2):
        for k in range(start, wn + 2, 2):
            for j in range(1, ny + 1):
                for i in range(1, nx + 1):
                    t = float(tv[i + 2, j + 2, k + 2])
                    want[i + 2, j + 2, k + 2] = (want[i + 2, j + 2, k + 2] + A2 * t) - A * cfg.temperature_ref
                    want[i + 2, j + 2, k + 1] = want[i + 2, j + 2, k + 1] + A2 * t
    assert np.array_equal(got[s], want[s])
    # the parity order matters: the sum in the other order differs somewhere
    other = ws[s] + A2 * tv[up] + A2 * tv[s] - A * cfg.temperature_ref
    assert not np.array_equal(got[s], other)
    o.close()


def test_oracle_moisture_in_solids_brute_force():
    cfg = moist_config()
    case = capi.Case(cfg, cuboids=OBST)
    o = Oracle(cfg, case)
    q = np.asfortranarray(np.random.default_rng(6).uniform(0, 0.02, (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)))
    o.upload(capi.F_MOISTURE, q)
    o.op_moist_fill()
    got = o.download(capi.F_MOISTURE)
    pt = case.masks()[0]
    want = np.where(pt > 0, cfg.moisture_ref, q)
    assert (pt > 0).sum() > 100
    assert np.array_equal(got, want)
    o.close()


def test_oracle_moisture_ghost_rules():
    """Periodic x and y; Dirichlet bottom q(1-m) = 2 q_b - q(m), Neumann top
    q(nz+m) = q(nz) + m dz g (g = 0.003 here), m = 1, 2, 3."""
    cfg = moist_config(side_moist=[0, 0, 0, 0, Q_BOT, 0.003])
    o = Oracle(cfg, capi.Case(cfg))
    q = np.asfortranarray(np.random.default_rng(8).uniform(0, 0.02, (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)))
    o.upload(capi.F_MOISTURE, q)
    o.op_bound_moisture()
    got = o.download(capi.F_MOISTURE)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    I, J = slice(3, 3 + nx), slice(3, 3 + ny)
    for m in (1, 2, 3):
        assert np.array_equal(got[I, J, 3 - m], Q_BOT - (q[I, J, 2 + m] - Q_BOT))
        assert np.array_equal(got[I, J, 2 + nz + m], q[I, J, 2 + nz] + 0.003 * m * cfg.dz)
    K = slice(3, 3 + nz)
    assert np.array_equal(got[0:3, J, K], got[nx:nx + 3, J, K])
    assert np.array_equal(got[3 + nx:6 + nx, J, K], got[3:6, J, K])
    assert np.array_equal(got[3:3 + nx, 0:3, K], got[3:3 + nx, ny:ny + 3, K])
    o.close()


def test_moisture_needs_buoyancy():
    cfg = moist_config(enable_buoyancy=0)
    with pytest.raises(capi.ElmmError, match="enable_buoyancy"):
        capi.Case(cfg)


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
def test_gpu_moisture_bit_identical_to_oracle(require_gpu):
    cfg = moist_config(48, 32, 16)
    case = capi.Case(cfg, cuboids=OBST, sources=[(4.5, 4.5, 0.5, 1.0, 0)])
    f = cases.initial_fields(cfg, u0=1.0, amp=0.2, seed=31)
    T = temperature_field(cfg, seed=32)
    q = moisture_field(cfg, seed=33)
    o, s = Oracle(cfg, case), capi.Solver(cfg, case)
    s32 = capi.Solver(capi.with_precision(cfg, capi.FP32_REFINED), case)
    for x in (o, s, s32):
        cases.upload_initial(x, cfg, f)
        x.upload(capi.F_TEMPERATURE, T)
        x.upload(capi.F_MOISTURE, q)
        x.init()
    fields = ["U", "V", "W", "Pr", "Visc", "TDiff", "S0"]

    def same():
        err = compare_state(s, o, cfg, fields)
        assert max(err.values()) == 0.0, err
        for fid in (capi.F_TEMPERATURE, capi.F_MOISTURE):
            assert np.array_equal(s.download(fid), o.download(fid)), fid

    same()
    for n in (1, 2):
        o.step(n)
        s.step(n)
        s32.step(n)
        same()
        l2 = velocity_rel_l2(s32, o, cfg)
        q32 = np.abs(s32.download(capi.F_MOISTURE) - o.download(capi.F_MOISTURE)).max() / 0.012
        print(f"moisture after {n} more steps: precision form velocity L2 {l2:.2e}, moisture {q32:.2e}", flush=True)
        assert l2 <= 1e-5 and q32 <= 1e-5, (l2, q32)
    for x in (o, s, s32):
        x.close()
