"""Asynchronous field frames (elmm_frame_async / elmm_frame_wait): the VTK
files written while later steps run hold the state at the call, the same as
elmm_download of that state through elmm_vtk_write (frame writers:
src/frames.f90:369-420; formats: src/outputs.f90:1560-1854, 2260-2490)."""
import filecmp
import os

import pytest

from elmm_amd import capi, cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [capi.FP64, capi.FP32_REFINED])
def test_frame_is_the_state_at_the_call(require_gpu, tmp_path, precision):
    cfg = capi.with_precision(cases.uniform_config(48, 32, 16, dx=1.0, dt=0.05, nscalars=2), precision)
    case = capi.Case(cfg, cuboids=[(10.0, 18.0, 8.0, 16.0, 0.0, 6.0)], sources=[(3.5, 3.5, 0.5, 1.0, 1)])
    f = cases.initial_fields(cfg, u0=2.0, amp=0.2, seed=4)
    ref = capi.Solver(cfg, case)
    s = capi.Solver(cfg, case)
    for x in (ref, s):
        cases.upload_initial(x, cfg, f)
        x.init()
    ref.step(3)
    got = tmp_path / "got"
    got.mkdir()
    s.step(3)
    s.step_async(2)             # queued before the frame: part of it
    s.frame_async(str(got))
    s.step_async(4)             # queued after: runs while the files are written
    s.frame_wait()
    s.wait()
    ref.step(2)
    want2 = tmp_path / "want2"
    want2.mkdir()
    capi.vtk_write(str(want2), cfg, *[ref.download(c) for c in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR)],
                   [ref.download(capi.F_SCALAR + q) for q in range(2)])
    for name in ("out.vtk", "scalars.vtk", "U.vtk", "V.vtk", "W.vtk"):
        assert filecmp.cmp(got / name, want2 / name, shallow=False), name
    assert os.path.getsize(got / "out.vtk") > 48 * 32 * 16 * 16
