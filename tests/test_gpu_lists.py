"""GPU parity on hand-made list dependencies: an IBMomentum list whose
targets are sources of other points (two-phase gather/scatter,
src/tsteps.f90:368-418) and a ScalFl viscosity list whose points read earlier
points' targets (serial loop, src/dynamics.f90:899-902, which the device
runs in dependency levels).  Bit-identical to the oracle after 1 and 3
steps, in the FP64 form."""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state
from oracle_py import Oracle
from test_oracle_pins_lists import SCALFL_CHAIN, ListsCase

pytestmark = pytest.mark.gpu

IB_CHAIN = (np.array([[4, 4, 2], [5, 4, 2], [6, 4, 2]]),
            np.array([0, 2, 4, 6]),
            np.array([[5, 4, 2], [4, 5, 2], [6, 4, 2], [4, 4, 3], [4, 4, 2], [7, 4, 2]]),
            np.array([0.7, 0.3, 0.6, 0.4, 1.5, -0.5]))


def test_gpu_chained_lists_bit_identical(require_gpu):
    cfg = cases.uniform_config(16, 16, 8, dx=1.0, dt=0.05, nscalars=1)
    case = capi.Case(cfg)
    lc = ListsCase(ListsCase(case, 0, *IB_CHAIN), 3, *SCALFL_CHAIN)
    f = cases.initial_fields(cfg, seed=9)
    o = Oracle(cfg, lc)
    s = capi.Solver(cfg, lc)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg, f)
    o.init()
    s.init()
    for n in (1, 2):
        o.step(n)
        s.step(n)
        err = compare_state(s, o, cfg)
        print(f"chained lists after {n} more steps: {err}", flush=True)
        assert max(err.values()) == 0.0, err


@pytest.mark.parametrize("channel", [False, True])
def test_gpu_wm_ustar_matches_oracle(require_gpu, channel):
    """WMPoints%ustar (ComputeViscsWM, src/wallmodels.f90:2050-2152): rough
    and smooth obstacle walls over a rough floor, and the smooth-wall
    stretched channel; bit-identical to the oracle after the initial tail and
    after 1 and 3 more steps (FP64), and within 1e-5 in the precision form."""
    if channel:
        cfg = cases.channel_config(32, 16, 16)
        case = capi.Case(cfg)
        f = cases.channel_fields(cfg, seed=3)
    else:
        cfg = cases.uniform_config(48, 32, 16, dx=1.0, dt=0.05, nscalars=1)
        cub = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (24.0, 32.0, 14.0, 24.0, 0.0, 9.0, 0.02)]
        case = capi.Case(cfg, cuboids=cub)
        f = cases.initial_fields(cfg, u0=2.0, amp=0.3, seed=12)
    o = Oracle(cfg, case)
    s = capi.Solver(cfg, case)
    s32 = capi.Solver(capi.with_precision(cfg, capi.FP32_REFINED), case)
    for x in (o, s, s32):
        cases.upload_initial(x, cfg, f)
        x.init()
    want = o.wm_ustar()
    assert len(want) > 0 and np.all(want != 1.0)
    assert np.array_equal(s.wm_ustar(), want)
    err = compare_state(s, o, cfg, ["U", "V", "W", "Pr", "Visc"])
    assert max(err.values()) == 0.0, err
    for n in (1, 2):
        o.step(n)
        s.step(n)
        s32.step(n)
        want = o.wm_ustar()
        got = s.wm_ustar()
        print(f"{'channel' if channel else 'obstacles'}: {len(want)} points, u* max "
              f"{want.max():.4f}, FP64 max diff {np.abs(got - want).max():.2e}", flush=True)
        assert np.array_equal(got, want)
        rel = np.abs(s32.wm_ustar() - want).max() / np.abs(want).max()
        assert rel <= 1e-5, rel
    for x in (o, s, s32):
        x.close()
