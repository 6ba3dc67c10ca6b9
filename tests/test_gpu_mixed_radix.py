"""Grids whose x/y extents are 2^a 3^b but not powers of two (the batched
ensemble's 96, 144, 192): the mixed-radix transform passes against the
oracle's restatement, bit for bit in the FP64 form, on a uniform case, an
obstacle case with a point source, non-square grids and the ensemble's
largest case, 192 x 192 x 96 (solver: src/poisson.f90:76-140 takes any
Prnx, Prny)."""
import numpy as np
import pytest

from elmm_amd import capi, cases
from helpers import compare_state, velocity_rel_l2
from oracle_py import Oracle

pytestmark = pytest.mark.gpu


def _pair(cfg, cuboids=(), sources=(), seed=3, u0=2.0):
    case = capi.Case(cfg, cuboids=cuboids, sources=sources)
    f = cases.initial_fields(cfg, u0=u0, amp=0.1 * u0, seed=seed)
    o, s = Oracle(cfg, case), capi.Solver(cfg, case)
    cases.upload_initial(o, cfg, f)
    cases.upload_initial(s, cfg, f)
    o.init()
    s.init()
    return o, s, case, f


@pytest.mark.parametrize("nx,ny,nz", [(48, 48, 24), (96, 36, 16), (64, 96, 16), (144, 48, 8)])
def test_mixed_radix_fp64_bit_identical(require_gpu, nx, ny, nz):
    cfg = cases.uniform_config(nx, ny, nz, dx=1.0, dt=0.05, nscalars=1)
    cub = [(6.0, 14.0, 6.0, 12.0, 0.0, 5.0), (float(nx - 10), float(nx - 2), float(ny - 9), float(ny + 3), 0.0, 4.0)]
    o, s, _, _ = _pair(cfg, cuboids=cub, sources=[(2.5, 2.5, 0.5, 1.0, 0)])
    err = compare_state(s, o, cfg)
    assert max(err.values()) == 0.0, ("init", err)
    for n in (1, 2):
        o.step(n)
        s.step(n)
        err = compare_state(s, o, cfg)
        print(f"{nx}x{ny}x{nz} after {n} more steps: {err}", flush=True)
        assert max(err.values()) == 0.0, err


def test_mixed_radix_192_bit_identical(require_gpu):
    cfg = cases.uniform_config(192, 192, 96, dx=1.0, dt=0.05, nscalars=1)
    cub = [(float(x), float(x + 10), float(y), float(y + 10), 0.0, 12.0) for x in range(10, 190, 48)
           for y in range(20, 190, 48)]
    o, s, _, _ = _pair(cfg, cuboids=cub, sources=[(3.5, 3.5, 0.5, 1.0, 0)])
    do, ds = o.step(1), s.step(1)
    err = compare_state(s, o, cfg)
    print(f"192x192x96 1 step: {err}, max_div {ds.max_div:.3e}", flush=True)
    assert max(err.values()) == 0.0, err
    assert ds.max_div <= 1e-12


def test_mixed_radix_precision_form(require_gpu):
    cfg = cases.uniform_config(96, 96, 32, dx=1.0, dt=0.05, nscalars=1)
    o, _, case, f = _pair(cfg, cuboids=[(30.0, 46.0, 30.0, 46.0, 0.0, 10.0)])
    c32 = capi.with_precision(cfg, capi.FP32_REFINED)
    for it in (0, 1):
        s = capi.Solver(capi.with_precision(cfg, capi.FP32_REFINED, refine_iters=it), case)
        cases.upload_initial(s, c32, f)
        s.init()
        for _ in range(5):
            d = s.step(1)
            assert d.max_div <= 1e-10, (it, d.max_div)
        if it == 0:
            o.step(5)
        l2 = velocity_rel_l2(s, o, cfg)
        print(f"96x96x32 precision form refine_iters={it}: rel L2 {l2:.3e}", flush=True)
        assert l2 <= 1e-5
        s.close()
