"""Scenario generator of the ensemble and serving forms (host only)."""
import numpy as np

from elmm_amd import capi, serving


def _two_three(n):
    while n % 3 == 0:
        n //= 3
    return n & (n - 1) == 0


def test_grid_for_cells_shapes():
    """Nx = Ny a multiple of 16 of the form 2^a 3^b near (2N)^(1/3), Nz a
    multiple of 8; both ends of the ensemble's range as stated."""
    for n in (262144, 1.4e6, 3538944, 8e6, 0.25e6):
        nx, ny, nz = serving.grid_for_cells(n)
        assert nx == ny and nx % 16 == 0 and _two_three(nx) and nz % 8 == 0 and 8 <= nz <= 256
        assert 0.5 < nx * ny * nz / n < 2.0
    assert serving.grid_for_cells(3538944) == (192, 192, 96)
    assert serving.grid_for_cells(64 ** 3) == (96, 96, 32)


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
