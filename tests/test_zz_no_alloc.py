"""elmm_divergence_noalloc (the allocation-free variant of elmm_divergence)
against elmm_divergence through the C ABI: the same value on two calls in a
row, the second on a smaller grid, and the size limit of its static scratch.
"""
import numpy as np
import pytest

from elmm_amd import capi, cases

pytestmark = pytest.mark.gpu


def _solver(cfg, cuboids, seed, steps):
    case = capi.Case(cfg, cuboids=cuboids)
    s = capi.Solver(cfg, case)
    cases.upload_initial(s, cfg, cases.initial_fields(cfg, u0=2.0, amp=0.3, seed=seed))
    s.init()
    if steps:
        s.step(steps)
    return s


def test_noalloc_divergence_equals_allocating_entry_point(require_gpu):
    big = cases.uniform_config(96, 64, 40, dx=1.0, dt=0.05, nscalars=1)
    small = capi.with_precision(cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05), capi.FP32_REFINED)
    runs = [(big, [(20.0, 36.0, 12.0, 28.0, 0.0, 17.0)], 3, 2),
            (small, [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0)], 4, 1)]
    for cfg, cub, seed, steps in runs:
        s = _solver(cfg, cub, seed, steps)
        want = s.divergence()
        got = s.divergence_noalloc()
        print(f"{cfg.nx}x{cfg.ny}x{cfg.nz}: elmm_divergence {want:.6e}, noalloc {got:.6e}", flush=True)
        assert np.isfinite(got) and got == want
        # repeated call on the same state: the scratch is rewritten, same value
        assert s.divergence_noalloc() == want
        s.close()


def test_noalloc_divergence_rejects_grids_beyond_its_scratch(require_gpu):
    # 1024 x 1024 x 72 = 75,497,472 cells > 67,108,864; no obstacles, FP32 form
    cfg = capi.with_precision(cases.uniform_config(1024, 1024, 72, dx=1.0, dt=0.05, nscalars=0), capi.FP32_REFINED)
    s = capi.Solver(cfg, capi.Case(cfg))
    with pytest.raises(capi.ElmmError) as e:
        s.divergence_noalloc()
    assert "scratch limit" in str(e.value)
    s.close()
