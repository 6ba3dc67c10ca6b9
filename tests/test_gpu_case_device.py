"""Solid-body masks and lists built on the device (elmm_case_build_device):
the type arrays, the IB lists, the solid-body wall-model points, the
nullified points and the scalar face masks and lists, equal byte for byte and
in order to the host builder (elmm_case_build), which the solids tests pin
against point-in-box tests, binary dilation and erosion, interpolation of
wall-linear fields and closed-form wall distances; with them the wall-model
points' thermal classes, surface moisture and moisture fluxes.  Cases:
cuboids on a periodic grid, cuboids across the periodic x and y boundaries,
a rough no-slip top and a body touching it with a heat- and moisture-flux
floor and a Dirichlet top, the wall model off, a non-unit
spacing with cuboid faces off the grid, dispersion scenarios of 0.3-2.5 M
cells with 30-140 bodies, the 512x512x128 urban case with 64 bodies; no
bodies; and an invalid cuboid, rejected with the host builder's message."""
import numpy as np
import pytest

from elmm_amd import capi, cases, serving, workloads
from test_solids import _case_arrays

pytestmark = pytest.mark.gpu

CUB = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (20.0, 26.0, 18.0, 30.0, 0.0, 10.0, 0.01)]


def _cases():
    out = []
    cfg = cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05, nscalars=1)
    out.append(("periodic", cfg, CUB, []))
    wrap = [(-3.0, 5.0, 10.0, 20.0, 0.0, 6.0), (26.0, 35.0, -2.0, 4.0, 0.0, 4.0, 0.05),
            (12.0, 14.0, 12.0, 14.0, 0.0, 15.5)]
    out.append(("wrap", cfg, wrap, []))
    # no-slip top; buoyancy with a heat-flux floor and a Dirichlet top, and
    # moisture with a flux floor: wall points of both walls carry thermal
    # classes, surface moisture and moisture fluxes
    top = cases.uniform_config(32, 24, 16, dx=1.0, dt=0.05, nscalars=1, top=capi.NOSLIP, enable_buoyancy=1,
                               temp_btype=[capi.PERIODIC] * 4 + [capi.CONSTFLUX, capi.DIRICHLET],
                               side_temp=[0, 0, 0, 0, 0.1, 296.0], enable_moisture=1, moisture_ref=0.01,
                               moist_btype=[capi.PERIODIC] * 4 + [capi.CONSTFLUX, capi.DIRICHLET],
                               side_moist=[0, 0, 0, 0, 1e-4, 0.008])
    top.z0_wall[5] = 0.02
    out.append(("noslip top", top, [(0.0, 6.0, 4.0, 10.0, 0.0, 5.0), (20.0, 32.0, 18.0, 24.0, 0.0, 8.0, 0.03, 0.2, 3e-5),
                                    (10.0, 14.0, 10.0, 14.0, 12.0, 16.0)], [(15.5, 12.5, 0.5, 1.0, 0)]))
    out.append(("no wall model", cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05, nscalars=1, wall_model=0), CUB, []))
    dx07 = cases.uniform_config(40, 36, 20, dx=0.7, dt=0.05, nscalars=1)
    out.append(("dx 0.7", dx07, [(3.3, 9.1, 2.2, 8.8, 0.0, 4.1), (15.0, 15.9, 10.0, 20.0, 0.0, 7.7)], []))
    for seed, n in ((1, 3e5), (3, 1.5e6), (4, 2.5e6)):
        c, cub, src, _ = serving.scenario(seed, n, capi.FP32_REFINED)
        out.append((f"scenario {seed}", c, cub, src))
    return out


def _same(name, cfg, cub, src):
    host = _case_arrays(capi.Case(cfg, cuboids=cub, sources=src))
    dev_case = capi.Case(cfg, cuboids=cub, sources=src, device=True)
    dev = _case_arrays(dev_case)
    assert len(host) == len(dev)
    for n, (a, b) in enumerate(zip(host, dev)):
        assert a.dtype == b.dtype and a.shape == b.shape, (name, n, a.shape, b.shape)
        assert a.tobytes() == b.tobytes(), (name, n)
    return dev_case.counts()


@pytest.mark.parametrize("k", range(8))
def test_device_case_equals_host_case(require_gpu, k):
    name, cfg, cub, src = _cases()[k]
    counts = _same(name, cfg, cub, src)
    print(f"{name}: {cfg.nx}x{cfg.ny}x{cfg.nz}, {len(cub)} bodies, IB points {counts['ib']}, "
          f"WM Pr points {counts['wmp']}, null points {counts['null']}", flush=True)
    assert sum(counts["ib"]) > 0


def test_device_case_urban512(require_gpu):
    cfg, cub, src = workloads.get("urban512").build()
    counts = _same("urban512", cfg, cub, src)
    assert sum(counts["ib"]) > 100000


def test_device_case_without_bodies(require_gpu):
    cfg = cases.uniform_config(16, 16, 8, dx=1.0, dt=0.05, nscalars=1)
    _same("none", cfg, [], [(3.5, 3.5, 0.5, 1.0, 0)])


def test_device_case_rejects_a_bad_cuboid(require_gpu):
    cfg = cases.uniform_config(16, 16, 8, dx=1.0, dt=0.05, nscalars=1)
    with pytest.raises(RuntimeError, match="cuboid 0: empty extent"):
        capi.Case(cfg, cuboids=[(4.0, 2.0, 1.0, 5.0, 0.0, 3.0)], device=True)


def test_device_case_runs_like_the_host_case(require_gpu):
    """A solver created from the device-built case steps to the same bits as
    one created from the host-built case (FP64, 2 steps)."""
    cfg, cub, src, fields = serving.scenario(11, 4e5, capi.FP64)
    out = []
    for device in (False, True):
        case = capi.Case(cfg, cuboids=cub, sources=src, device=device)
        s = capi.Solver(cfg, case)
        cases.upload_initial(s, cfg, fields)
        s.init()
        s.step(2)
        out.append([s.download(f) for f in (capi.F_U, capi.F_V, capi.F_W, capi.F_PR)])
        s.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
