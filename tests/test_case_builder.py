"""Case set-up (host side): outer-wall wall-model lists and scalar wall-face
masks for a flat bottom (GetOutsideBoundariesWM_UVW/_Pr,
src/wallmodels.f90:467-904; InitWMMasks, src/wallmodels.f90:911-1113)."""
import numpy as np

from elmm_amd import capi, cases


def test_flat_bottom_lists():
    cfg = cases.config1(16)
    case = capi.Case(cfg)
    cnt = case.counts()
    # one bottom point per column for U, V, W (dir zm = 4), nothing else
    for comp in range(3):
        for d in range(6):
            assert cnt["wm"][comp * 6 + d] == (16 * 16 if d == 4 else 0)
    ijk, dist, z0 = case.wm_list(0 * 6 + 4)
    assert np.all(ijk[:, 2] == 1)
    # points in (j outer, i inner) order
    assert np.array_equal(ijk[:, 0], np.tile(np.arange(1, 17), 16))
    assert np.array_equal(ijk[:, 1], np.repeat(np.arange(1, 17), 16))
    assert np.all(dist == [0, 0, -0.5])      # zW(0) - zPr(1)
    assert np.all(z0 == 0.01)
    _, dist_w, _ = case.wm_list(2 * 6 + 4)
    assert np.all(dist_w == [0, 0, -1.0])    # zW(0) - zW(1) for the W point
    pijk, pdist, _ = case.wmp_list()
    assert len(pijk) == 256 and np.all(pijk[:, 2] == 1) and np.all(pdist[:, 2] == -0.5)
    assert cnt["scfl"] == [0, 0, 0] and cnt["ib"] == [0, 0, 0, 0] and cnt["null"] == [0, 0, 0]


def test_noslip_top_adds_top_lists():
    cfg = cases.config1(8, top=capi.NOSLIP, z0_bottom=0.0)
    cfg.z0_wall[5] = 0.02
    case = capi.Case(cfg)
    cnt = case.counts()
    assert cnt["wm"][0 * 6 + 5] == 64 and cnt["wm"][2 * 6 + 5] == 64
    ijk, dist, z0 = case.wm_list(0 * 6 + 5)
    assert np.all(ijk[:, 2] == 8) and np.all(dist[:, 2] == 0.5) and np.all(z0 == 0.02)
    ijk, dist, _ = case.wm_list(2 * 6 + 5)
    # W top point is Wnz = nz-1 = 7: zW(Wnz+1) - zW(7) = 1
    assert np.all(ijk[:, 2] == 7) and np.all(dist[:, 2] == 1.0)
    assert case.counts()["wmp"] == 128


def test_point_source_cell_and_flux():
    cfg = cases.config1(8)
    case = capi.Case(cfg, sources=[(2.3, 5.9, 0.2, 1.5, 0)])
    v = case.view
    assert v.src_n == 1
    assert [v.src_ijk[0], v.src_ijk[1], v.src_ijk[2]] == [3, 6, 1]   # nint(x/dx + 0.5)
    assert v.src_flux[0] == 1.5                                        # / cell volume 1


def test_masks_default_fluid():
    cfg = cases.config1(8)
    pt, ut, vt, wt = capi.Case(cfg).masks()
    assert pt.shape == (14, 14, 14)
    assert not pt.any() and not ut.any() and not vt.any() and not wt.any()
