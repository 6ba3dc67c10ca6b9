"""Solid-body case set-up for cuboids: type masks, immersed-boundary
interpolation lists, solid-body wall-model points and nullified points
(FindInsideCells, src/solidbodies.f90:75-206; FindNeighbouringCells,
src/immersedboundary.f90:1475-1506; InitImBoundaries / TVelIBPoint_Create /
InterpolationPoints / TScalFlIBPoint_Create, src/immersedboundary.f90:
690-1410, 1650-1720; GetSolidBodiesWM(_UVW), src/immersedboundary.f90:14-308;
SetNullifiedPoints, src/initial.f90:4861-4962).

Pins: the masks against a point-in-box test on cell centres, the 27-point
neighbour marking against scipy's binary dilation, the nullified points
against binary erosion, interpolation against fields the reference's
interpolation reproduces (linear in the wall-normal direction and zero at the
wall, within the bisection tolerance of ClosestOnLineOut), partition of unity
of the scalar-flux weights, and the wall distances of grid-aligned faces."""
import numpy as np
import pytest
from scipy import ndimage

from elmm_amd import capi, cases

CUB = [(8.0, 16.0, 8.0, 16.0, 0.0, 6.0), (20.0, 26.0, 18.0, 30.0, 0.0, 10.0, 0.01)]
EPS = 0.05  # (dx dy dz)^(1/3) / 20


@pytest.fixture(scope="module")
def built():
    cfg = cases.uniform_config(32, 32, 16, dx=1.0, dt=0.05, nscalars=1)
    return cfg, capi.Case(cfg, cuboids=CUB)


def coords(n, stag):
    i = np.arange(-2, n + 4, dtype=float)
    return i if stag else i - 0.5


def expected_mask(cfg, sx, sy, sz):
    x, y, z = coords(cfg.nx, sx), coords(cfg.ny, sy), coords(cfg.nz, sz)
    # periodic x, y: coordinates beyond the first ghost face are wrapped by
    # the domain length (Body%Inside, src/geometricshapes.f90:2917-2945)
    x = np.where(x > cfg.nx + 1, x - cfg.nx, np.where(x < 0, x + cfg.nx, x))
    y = np.where(y > cfg.ny + 1, y - cfg.ny, np.where(y < 0, y + cfg.ny, y))
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    m = np.zeros(X.shape, np.int32)
    for b, c in enumerate(CUB, start=1):
        ins = ((X >= c[0] - EPS) & (X <= c[1] + EPS) & (Y >= c[2] - EPS) & (Y <= c[3] + EPS)
               & (Z >= c[4] - EPS) & (Z <= c[5] + EPS))
        m[ins] = b
    return m


def test_masks_match_point_in_box_and_dilation(built):
    cfg, case = built
    got = case.masks()
    for t, stag in zip(got, [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)]):
        solid = expected_mask(cfg, *stag)
        assert np.array_equal(np.where(t > 0, t, 0), solid)
        # -1 = free cell within the 27-point neighbourhood of a solid one
        # (interior+1 ring only; the outermost ghost layer is not visited)
        near = ndimage.binary_dilation(solid > 0, structure=np.ones((3, 3, 3), bool)) & (solid == 0)
        inner = np.zeros_like(near)
        inner[1:-1, 1:-1, 1:-1] = True
        assert np.array_equal(t == -1, near & inner)


def test_nullified_points_are_eroded_solid(built):
    cfg, case = built
    n = [(cfg.nx, cfg.ny, cfg.nz), (cfg.nx, cfg.ny, cfg.nz), (cfg.nx, cfg.ny, cfg.nz - 1)]
    for comp, t in enumerate(case.masks()[1:]):
        er = ndimage.binary_erosion(t > 0, structure=ndimage.generate_binary_structure(3, 1), border_value=0)
        ijk = np.argwhere(er[3:3 + n[comp][0], 3:3 + n[comp][1], 3:3 + n[comp][2]]) + 1
        # SetNullifiedPoints lists them in k, j, i order
        ijk = ijk[np.lexsort((ijk[:, 0], ijk[:, 1], ijk[:, 2]))]
        assert np.array_equal(case.null_list(comp), ijk)


def _wall_coordinate(cub, axis, target_coord, step):
    """Face of the cuboid between a forcing point and its outside neighbour."""
    lo, hi = cub[2 * axis], cub[2 * axis + 1]
    return lo if step < 0 else hi


def test_velocity_ib_reproduces_wall_linear_field(built):
    cfg, case = built
    masks = case.masks()
    checked = 0
    for comp in range(3):
        ijk, ptr, src, coef = case.ib_list(comp)
        stag = [0, 0, 0]
        stag[comp] = 1
        for p in range(len(ijk)):
            a, b = ptr[p], ptr[p + 1]
            if b == a:
                # null point: value forced to zero; it must lie on (or within
                # 0.1 dx of) a face of its body
                continue
            taps = src[a:b] - ijk[p]
            axes = [np.flatnonzero(t)[0] for t in taps]
            if len(set(axes)) != 1:
                continue
            ax = axes[0]
            step = int(np.sign(taps[0][ax]))
            body = masks[comp + 1][tuple(ijk[p] + 2)]
            w = _wall_coordinate(CUB[body - 1], ax, None, step)
            pos = lambda q: q[ax] - (0.0 if stag[ax] else 0.5)
            f_t = pos(ijk[p]) - w
            f_i = sum(c * (pos(s) - w) for c, s in zip(coef[a:b], src[a:b]))
            assert abs(f_i - f_t) <= 2e-3 * abs(f_t) + 1e-12, (comp, ijk[p], f_i, f_t)
            checked += 1
    assert checked > 500


def test_velocity_ib_null_points_on_faces(built):
    cfg, case = built
    for comp in range(3):
        ijk, ptr, _, _ = case.ib_list(comp)
        nulls = ijk[np.diff(ptr) == 0]
        pos = nulls.astype(float) - 0.5
        pos[:, comp] += 0.5
        onface = np.zeros(len(nulls), bool)
        for c in CUB:
            for ax in range(3):
                for f in (c[2 * ax], c[2 * ax + 1]):
                    onface |= np.abs(pos[:, ax] - f) <= 0.1
        assert onface.all()


def test_scalar_flux_ib_weights(built):
    cfg, case = built
    pt = case.masks()[0]
    ijk, ptr, src, coef = case.ib_list(3)
    assert len(ijk) > 0
    sums = np.add.reduceat(coef, ptr[:-1])
    assert np.allclose(sums, 1.0, atol=1e-14, rtol=0)
    assert np.all(pt[tuple((src + 2).T)] <= 0)
    # every solid cell with a free face neighbour is listed, in k, j, i order
    s = pt > 0
    free = ~s
    touch = np.zeros_like(s)
    for ax in range(3):
        for d in (1, -1):
            touch |= np.roll(free, d, axis=ax)
    want = np.argwhere((s & touch)[3:3 + cfg.nx, 3:3 + cfg.ny, 3:3 + cfg.nz]) + 1
    want = want[np.lexsort((want[:, 0], want[:, 1], want[:, 2]))]
    assert np.array_equal(ijk, want)


def test_solid_wall_model_points(built):
    cfg, case = built
    pijk, pdist, pz0 = case.wmp_list()
    pt = case.masks()[0]
    # Pr cells next to a grid-aligned face: distance 0.5 along the normal,
    # z0 from the body (second body rough, first smooth)
    side = np.zeros(len(pijk), bool)
    for q, (ijk, d) in enumerate(zip(pijk, pdist)):
        nb = [pt[tuple(ijk + 2 + e)] for e in np.vstack([np.eye(3, dtype=int), -np.eye(3, dtype=int)])]
        if max(nb) > 0:
            side[q] = True
            assert np.isclose(np.linalg.norm(d), 0.5) and np.count_nonzero(d) == 1
            assert pz0[q] == (0.01 if max(nb) == 2 else 0.0)
    assert side.sum() > 0
    # U points at x = 7 (one cell west of the first body's west face)
    ijk, dist, _ = case.wm_list(0 * 6 + 1)   # U, +x neighbour solid
    sel = ijk[:, 0] == 7
    assert sel.any() and np.allclose(dist[sel], [1.0, 0, 0])


def test_bad_cuboid_is_rejected():
    cfg = cases.uniform_config(16, 16, 8, dx=1.0, dt=0.05, nscalars=1)
    with pytest.raises(RuntimeError, match="cuboid"):
        capi.Case(cfg, cuboids=[(4.0, 2.0, 1.0, 5.0, 0.0, 3.0)])


def _case_arrays(case):
    out = list(case.masks())
    for l in range(18):
        out.extend(case.wm_list(l))
        out.extend(case.wm_thermal(l))
        out.extend(case.wm_moist(l))
    out.extend(case.wmp_list())
    out.extend(case.wm_thermal(-1))
    out.extend(case.wm_moist(-1))
    for c in range(4):
        out.extend(case.ib_list(c))
    for c in range(3):
        out.append(case.null_list(c))
    v = case.view
    n = int(np.prod(case.shape()))
    for d in range(3):
        out.append(np.ctypeslib.as_array(v.scfl_mask[d], shape=(n,)).copy())
        out.append(np.ctypeslib.as_array(v.scfl_ijk[d], shape=(3 * v.scfl_n[d],)).copy() if v.scfl_n[d] else
                   np.zeros(0, np.int32))
    return out


def test_builder_threads_do_not_change_the_lists():
    """The builders of the four type arrays, the four IB lists, the Pr and
    per-component wall-model lists, the nullified points and the three scalar
    face directions run on their own threads; every list must come out the
    same, in the same order, on every build of a case with many bodies."""
    from elmm_amd import serving
    cfg, cub, src, _ = serving.scenario(4007, 400000, capi.FP32_REFINED)
    assert len(cub) > 10
    ref = _case_arrays(capi.Case(cfg, cuboids=cub, sources=src))
    for _ in range(3):
        got = _case_arrays(capi.Case(cfg, cuboids=cub, sources=src))
        assert len(got) == len(ref)
        for a, b in zip(ref, got):
            assert a.shape == b.shape and a.tobytes() == b.tobytes()
