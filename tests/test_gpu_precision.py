"""Precision form (ELMM_FP32_REFINED) against the FP64 oracle.

FP32 fields and stencils, FP64 divergence/RHS and pressure potential; the
pressure solve is one FP64 spectral solve (refine_iters = 0, the default) or
FP32 spectral solves inside FP64 iterative refinement (refine_iters > 0);
both are tested.  Bars (stated per field):
relative L2 velocity error <= 1e-5 after 10 steps; the FP64 divergence of the
corrected velocity (before it is stored in FP32) <= 1e-10; the stored FP32
field's divergence is bounded by FP32 rounding of the velocity."""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state, velocity_rel_l2
from oracle_py import Oracle

pytestmark = pytest.mark.gpu

OBSTACLES = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (20.0, 26.0, 18.0, 30.0, 0.0, 10.0, 0.01)]


def pair32(cfg, cuboids=(), sources=(), u0=1.0, seed=1, refine=0):
    case = capi.Case(cfg, cuboids=cuboids, sources=sources)
    f = cases.initial_fields(cfg, u0=u0, amp=0.1 * u0, seed=seed)
    cfg32 = capi.with_precision(cfg, capi.FP32_REFINED, refine)
    o = Oracle(cfg, case)
    s = capi.Solver(cfg32, case)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg32, f)
    o.init()
    s.init()
    return o, s


@pytest.mark.parametrize("refine", [0, 1])
@pytest.mark.parametrize("obstacles", [False, True])
def test_fp32_refined_ten_steps(require_gpu, obstacles, refine):
    cfg = cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05, nscalars=1)
    kw = dict(cuboids=OBSTACLES, sources=[(4.5, 4.5, 0.5, 1.0, 0)], u0=3.0, seed=3) if obstacles else {}
    o, s = pair32(cfg, refine=refine, **kw)
    o.step(9)
    s.step(9)
    do, ds = o.step(1), s.step(1)
    l2 = velocity_rel_l2(s, o, cfg)
    err = compare_state(s, o, cfg)
    print("rel L2 velocity", l2, "max-rel per field", err, "div64", ds.max_div, "stored", ds.stored_div,
          "res", ds.poisson_res, "oracle div", do.max_div)
    assert l2 <= 1e-5
    assert ds.refinements == refine
    assert ds.max_div <= 1e-10
    # the stored FP32 velocity carries rounding-level divergence only
    assert ds.stored_div <= 64 * np.finfo(np.float32).eps * 3.3 * 3
    assert err["Pr"] <= 1e-3 and err["S0"] <= 1e-3


def test_fp32_refinement_reduces_residual(require_gpu):
    """The FP64 residual of the pressure solve falls with each FP32
    refinement sweep until it reaches the FP64 rounding level."""
    cfg = cases.config1(32)
    res = {}
    for it in (1, 2, 4):
        _, s = pair32(cfg, refine=it)
        d = s.step(2)
        assert d.refinements == it
        res[it] = d.poisson_res
    print("residual by refinement sweeps", res)
    assert res[2] <= res[1] and res[4] <= 1e-9


@pytest.mark.parametrize("refine", [0, 1])
def test_fp32_refined_divergence_every_step(require_gpu, refine):
    """Larger obstacle case (128 x 128 x 64): after every step the FP64
    divergence of the corrected velocity stays <= 1e-10, with the FP64 solve
    and with one refinement sweep of FP32 solves."""
    cfg = cases.uniform_config(128, 128, 64, dx=1.0, dt=0.05, nscalars=1)
    cub = [(16.0 + 32 * a, 32.0 + 32 * a, 16.0 + 32 * b, 32.0 + 32 * b, 0.0, 8.0 + 4 * (a + b)) for a in range(4)
           for b in range(4)]
    case = capi.Case(cfg, cuboids=cub, sources=[(5.5, 5.5, 0.5, 1.0, 0)])
    s = capi.Solver(capi.with_precision(cfg, capi.FP32_REFINED, refine), case)
    cases.upload_initial(s, cfg, cases.initial_fields(cfg, u0=3.0, amp=0.3, seed=3))
    s.init()
    worst = 0.0
    for _ in range(12):
        d = s.step(1)
        worst = max(worst, d.max_div)
    print("worst FP64 divergence over 12 steps", worst)
    assert worst <= 1e-10
