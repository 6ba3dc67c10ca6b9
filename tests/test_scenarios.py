"""Scenario generator of the ensemble and serving forms (host only)."""
import numpy as np

from elmm_amd import capi, serving


def test_grid_for_cells_shapes():
    for n in (262144, 1.4e6, 3538944, 8e6):
        nx, ny, nz = serving.grid_for_cells(n)
        assert nx == ny and nx & (nx - 1) == 0 and nz % 8 == 0 and 16 <= nz <= 256
        assert 0.5 < nx * ny * nz / n < 2.0


def test_scenario_obstacles_and_source():
    cfg, cub, src, fields = serving.scenario(11, 1.0e6)
    area = sum((c[1] - c[0]) * (c[3] - c[2]) for c in cub)
    assert 0.10 * cfg.nx * cfg.ny <= area <= 0.30 * cfg.nx * cfg.ny + 144
    for c in cub:
        assert 0 <= c[0] < c[1] <= cfg.nx and 0 <= c[2] < c[3] <= cfg.ny and c[5] < cfg.nz
        assert not (c[0] <= src[0][0] <= c[1] and c[2] <= src[0][1] <= c[3])
    case = capi.Case(cfg, cuboids=cub, sources=src)
    pt = case.masks()[0]
    assert (pt[3:-3, 3:-3, 3:-3] > 0).any()
    U, V = fields[0], fields[1]
    speed = np.hypot(U[3:-3, 3:-3, 3:-3].mean(), V[3:-3, 3:-3, 3:-3].mean())
    assert 0.9 <= speed <= 5.1
