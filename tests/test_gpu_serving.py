"""Serving form (elmm_server_*): requests arriving while others run are
admitted as memory frees up, their work overlaps on the device, and every
result equals the same case run alone (bit for bit: each case's arithmetic
is independent of the mix)."""
import time

import numpy as np
import pytest

from elmm_amd import capi, cases

pytestmark = pytest.mark.gpu


def _cases():
    out = []
    for n, nz, nsteps, cub, prec in [(32, 16, 7, [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0)], capi.FP64),
                                      (64, 16, 5, [(20.0, 30.0, 12.0, 40.0, 0.0, 8.0, 0.01)], capi.FP64),
                                      (32, 32, 9, [], capi.FP32_REFINED),
                                      (64, 32, 4, [(4.0, 12.0, 40.0, 56.0, 0.0, 10.0)], capi.FP32_REFINED),
                                      (32, 16, 12, [], capi.FP64)]:
        cfg = capi.with_precision(cases.uniform_config(n, n, nz, dx=1.0, dt=0.05, nscalars=1), prec)
        case = capi.Case(cfg, cuboids=cub, sources=[(3.5, 3.5, 0.5, 1.0, 0)])
        fields = cases.initial_fields(cfg, u0=2.0, amp=0.2, seed=n + nz)
        out.append((cfg, case, fields, nsteps))
    return out


def _alone(cfg, case, fields, nsteps):
    s = capi.Solver(cfg, case)
    cases.upload_initial(s, cfg, fields)
    s.init()
    s.step(nsteps)
    r = [s.download(c) for c in range(3)] + [s.download(capi.F_PR), s.download(capi.F_SCALAR)]
    s.close()
    return r


def test_mix_equals_alone_and_budget_is_respected(require_gpu):
    work = _cases()
    # a budget that holds about two of the cases at once forces queueing
    fps = [capi.footprint(w[0]) for w in work]
    budget = int(sorted(fps)[-1] + sorted(fps)[-2] + 1)
    srv = capi.Server(budget, max_active=8, chunk_steps=3)
    ids = []
    for k, (cfg, case, fields, nsteps) in enumerate(work):
        ids.append(srv.submit(cfg, case, fields, nsteps))
        srv.progress()            # requests arrive while others run
    peak_used = 0
    t0 = time.time()
    while True:
        a, q = srv.progress()
        peak_used = max(peak_used, srv.stats()[1])
        if a == 0 and q == 0:
            break
        assert time.time() - t0 < 60
        time.sleep(0.001)
    assert peak_used <= budget
    cell_steps, _ = srv.stats()
    assert cell_steps == sum(w[0].nx * w[0].ny * w[0].nz * w[3] for w in work)
    for rid, (cfg, case, fields, nsteps) in zip(ids, work):
        st = srv.status(rid)
        assert st.state == 2 and st.steps_done == nsteps and st.latency_s > 0
        U, V, W, Pr, S = srv.result(rid)
        ref = _alone(cfg, case, fields, nsteps)
        for got, want in zip([U, V, W, Pr, S[0]], ref):
            assert np.array_equal(got, want)
    srv.close()


def _hip_free_bytes():
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32_REFINED])
def test_case_footprint_covers_device_allocation(require_gpu, precision):
    """The admission footprint is at least what elmm_create takes from the
    device (hipMemGetInfo before and after) on an obstacle case."""
    cfg = capi.with_precision(cases.uniform_config(128, 128, 64, dx=1.0, dt=0.05, nscalars=1), precision)
    cub = [(float(x), float(x + 12), float(y), float(y + 12), 0.0, 20.0) for x in range(8, 120, 32)
           for y in range(8, 120, 32)]
    case = capi.Case(cfg, cuboids=cub, sources=[(3.5, 3.5, 0.5, 1.0, 0)])
    fp = capi.footprint(cfg, case)
    s0 = capi.Solver(capi.with_precision(cases.uniform_config(32, 32, 16), precision), capi.Case(
        cases.uniform_config(32, 32, 16)))   # runtime and code objects loaded before measuring
    before = _hip_free_bytes()
    s = capi.Solver(cfg, case)
    after = _hip_free_bytes()
    drop = before - after
    print(f"precision {precision}: footprint {fp} B, device memory taken by elmm_create {drop} B", flush=True)
    assert fp >= drop > 0
    s.close()
    s0.close()
