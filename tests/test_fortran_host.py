"""The Fortran host (gpu-elmm_amd/fortran): iso_c_binding types agree with
the C header, the namelist host reaches the library and reports its errors,
and on a GPU it produces the same state as the C ABI driven from Python."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from elmm_amd import build, capi, cases

HOST = build.host_path()
needs_host = pytest.mark.skipif(not os.path.exists(HOST), reason="Fortran host not built (amdflang absent)")


@needs_host
def test_interoperable_sizes_match_header():
    out = subprocess.run([HOST, "--sizes"], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out[1:5]]
    assert sizes == [C.sizeof(capi.Config), C.sizeof(capi.Cuboid), C.sizeof(capi.PointSource),
                     C.sizeof(capi.Diag)]


def _namelist(path, n, nz, steps, cuboids=(), sources=(), init="", out="", precision=0, device_lists=None):
    lines = [f"&grid nx={n}, ny={n}, nz={nz}, dx=1.0, dy=1.0, dz=1.0 /",
             f"&physics nscalars=1, precision={precision} /",
             f"&time_stepping dt=0.05, nsteps={steps}, report_every={steps} /"]
    if cuboids:
        vals = ", ".join(f"{v:.17g}" for c in cuboids for v in (list(c) + [0.0] * (7 - len(c))))
        dl = "" if device_lists is None else f", device_lists={'.true.' if device_lists else '.false.'}"
        lines.append(f"&obstacles n_cuboids={len(cuboids)}, cuboids={vals}{dl} /")
    if sources:
        vals = ", ".join(f"{v:.17g}" for s in sources for v in (s[0], s[1], s[2], s[3], s[4] + 1))
        lines.append(f"&point_sources n_sources={len(sources)}, sources={vals} /")
    lines.append(f"&io init_file='{init}', out_file='{out}' /")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


@needs_host
@pytest.mark.parametrize("device_lists", [None, False])
def test_host_reports_library_errors(tmp_path, device_lists):
    """A bad case (solid with empty extent) stops the host with the
    library's message (error_stop semantics of the reference), from the
    device case builder (the default) and the host one; the cuboids are
    checked before any device work, so no GPU is needed."""
    nml = tmp_path / "bad.nml"
    _namelist(nml, 16, 8, 1, cuboids=[(5.0, 2.0, 1.0, 4.0, 0.0, 3.0)], device_lists=device_lists)
    r = subprocess.run([HOST, str(nml)], capture_output=True, text=True)
    assert r.returncode != 0
    entry = "elmm_case_build" if device_lists is False else "elmm_case_build_device"
    assert f"{entry} failed" in r.stdout and "cuboid" in r.stdout


OBSTACLES = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (20.0, 26.0, 18.0, 30.0, 0.0, 10.0, 0.01)]
SOURCES = [(4.5, 4.5, 0.5, 1.0, 0)]


@pytest.mark.gpu
@needs_host
def test_fortran_host_matches_c_abi(require_gpu, tmp_path):
    n, nz, steps = 32, 16, 3
    cfg = cases.uniform_config(n, n, nz, dx=1.0, dt=0.05, nscalars=1)
    U, V, W, Pr, S = cases.initial_fields(cfg, u0=3.0, amp=0.3, seed=3)
    init = tmp_path / "init.bin"
    with open(init, "wb") as f:
        for a in (U, V, W):
            f.write(np.asfortranarray(a).tobytes(order="F"))
        f.write(np.asfortranarray(Pr[1:-1, 1:-1, 1:-1]).tobytes(order="F"))    # Pr has lower bound -1
        f.write(np.asfortranarray(S[0]).tobytes(order="F"))
    out = tmp_path / "out.bin"
    nml = tmp_path / "case.nml"
    _namelist(nml, n, nz, steps, OBSTACLES, SOURCES, init, out)
    r = subprocess.run([HOST, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr

    case = capi.Case(cfg, cuboids=OBSTACLES, sources=SOURCES)
    s = capi.Solver(cfg, case)
    cases.upload_initial(s, cfg, (U, V, W, Pr, S))
    s.init()
    s.step(steps)
    g = (n + 6, n + 6, nz + 6)
    gp = (n + 4, n + 4, nz + 4)
    raw = np.fromfile(out, dtype=np.float64)
    sizes = [np.prod(g)] * 3 + [np.prod(gp), np.prod(g)]
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    for c in range(3):
        assert np.array_equal(parts[c].reshape(g, order="F"), s.download(c))
    assert np.array_equal(parts[3].reshape(gp, order="F"), s.download(capi.F_PR)[1:-1, 1:-1, 1:-1])
    assert np.array_equal(parts[4].reshape(g, order="F"), s.download(capi.F_SCALAR))


@pytest.mark.gpu
@needs_host
def test_fortran_host_stretched_channel(require_gpu, tmp_path):
    """zgrid file (the reference's zgrid.txt form) -> stretched channel
    through the host, identical to the C ABI driven from Python."""
    nx, ny, nz, steps = 32, 16, 16, 3
    cfg = cases.channel_config(nx, ny, nz)
    U, V, W, Pr, _ = cases.channel_fields(cfg)
    zg = tmp_path / "zgrid.txt"
    zg.write_text("\n".join(f"{z:.17g}" for z in cfg._zw_keep) + "\n")
    init = tmp_path / "init.bin"
    with open(init, "wb") as f:
        for a in (U, V, W):
            f.write(np.asfortranarray(a).tobytes(order="F"))
        f.write(np.asfortranarray(Pr[1:-1, 1:-1, 1:-1]).tobytes(order="F"))
    out = tmp_path / "out.bin"
    nml = tmp_path / "chan.nml"
    nml.write_text("\n".join([
        f"&grid nx={nx}, ny={ny}, dx={cfg.dx:.17g}, dy={cfg.dy:.17g}, zgrid_file='{zg}' /",
        f"&physics nu={cfg.nu:.17g}, nu_scalar={cfg.nu_scalar:.17g}, nscalars=0, z0_bottom=0, z0_top=0, "
        f"top_bc={capi.NOSLIP}, advection_method=2, pr_gradient_x=-1.0 /",
        f"&time_stepping dt={cfg.dt:.17g}, nsteps={steps}, report_every={steps} /",
        f"&io init_file='{init}', out_file='{out}' /"]) + "\n")
    r = subprocess.run([HOST, str(nml)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    s = capi.Solver(cfg, capi.Case(cfg))
    cases.upload_initial(s, cfg, (U, V, W, Pr, []))
    s.init()
    s.step(steps)
    g = (nx + 6, ny + 6, nz + 6)
    gp = (nx + 4, ny + 4, nz + 4)
    raw = np.fromfile(out, dtype=np.float64)
    parts = np.split(raw, np.cumsum([np.prod(g)] * 3)[:3])
    for c in range(3):
        assert np.array_equal(parts[c].reshape(g, order="F"), s.download(c))
    assert np.array_equal(parts[3].reshape(gp, order="F"), s.download(capi.F_PR)[1:-1, 1:-1, 1:-1])
