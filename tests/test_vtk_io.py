"""Field files in the reference's format (elmm_vtk_write / elmm_vtk_read,
host only): legacy VTK, RECTILINEAR_GRID, BINARY, big-endian float32, as
written by OutputOut, OutputScalars and OutputUVW (src/outputs.f90:1560-1854,
2260-2490) and read by ReadInitialConditions (src/initial.f90:3575-3811).
The files are parsed here independently of the library."""
import os

import numpy as np
import pytest

from elmm_amd import capi, cases


def parse(path):
    """(header records, {name: float32 array}) of a legacy binary VTK file."""
    data = open(path, "rb").read()
    pos, recs, blocks = 0, [], {}

    def line():
        nonlocal pos
        e = data.index(b"\n", pos)
        s = data[pos:e].decode()
        pos = e + 1
        return s

    for _ in range(5):
        recs.append(line())
    dims = [int(x) for x in recs[4].split()[1:]]
    npts = dims[0] * dims[1] * dims[2]
    for ax, n in zip("XYZ", dims):
        r = line()
        recs.append(r)
        assert r.split()[:3] == [f"{ax}_COORDINATES", str(n), "float"]
        blocks[ax] = np.frombuffer(data, ">f4", n, pos).copy()
        pos += 4 * n
        assert data[pos:pos + 1] == b"\n"
        pos += 1
    recs.append(line())
    assert recs[-1].split() == ["POINT_DATA", str(npts)]
    while pos < len(data):
        head = line().split()
        if head[0] == "SCALARS":
            assert line() == "LOOKUP_TABLE default"
            blocks[head[1]] = np.frombuffer(data, ">f4", npts, pos).reshape(dims, order="F").copy()
            pos += 4 * npts
        else:
            assert head == ["VECTORS", "u", "float"]
            blocks["u"] = np.frombuffer(data, ">f4", 3 * npts, pos).reshape([3] + dims, order="F").copy()
            pos += 12 * npts
        assert data[pos:pos + 1] == b"\n"
        pos += 1
    return recs, blocks


@pytest.fixture()
def fields():
    cfg = cases.uniform_config(12, 10, 8, dx=0.5, dt=0.05, nscalars=2)
    g = np.random.Generator(np.random.Philox(key=[17, 3]))
    shape = (cfg.nx + 6, cfg.ny + 6, cfg.nz + 6)
    U, V, W, Pr, S0, S1, nu = (np.asfortranarray(g.uniform(-2, 2, shape)) for _ in range(7))
    return cfg, U, V, W, Pr, [S0, S1], nu


def test_write_matches_the_reference_layout(tmp_path, fields):
    cfg, U, V, W, Pr, S, nu = fields
    capi.vtk_write(str(tmp_path), cfg, U, V, W, Pr, S, visc=nu)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    I = np.s_[3:3 + nx, 3:3 + ny, 3:3 + nz]
    recs, b = parse(tmp_path / "out.vtk")
    assert recs[:4] == ["# vtk DataFile Version 2.0", "CLMM output file", "BINARY", "DATASET RECTILINEAR_GRID"]
    assert all(len(r) == 70 for r in recs[4:])                 # character(70) records
    assert recs[4][:10] == "DIMENSIONS" and recs[4][10] == " "
    assert recs[5][:13] == "X_COORDINATES" and recs[5][14:26] == "   12  float"
    # cell centres xPr(i) = (i - 1/2) dx (x0 = 0)
    np.testing.assert_array_equal(b["X"], ((np.arange(1, nx + 1) - 0.5) * 0.5).astype(np.float32))
    np.testing.assert_array_equal(b["Z"], ((np.arange(1, nz + 1) - 0.5) * 0.5).astype(np.float32))
    np.testing.assert_array_equal(b["p"], Pr[I].astype(np.float32))
    np.testing.assert_array_equal(b["visc"], nu[I].astype(np.float32))
    uc = (U[I] + U[2:2 + nx, 3:3 + ny, 3:3 + nz]) / 2.0
    vc = (V[I] + V[3:3 + nx, 2:2 + ny, 3:3 + nz]) / 2.0
    wc = (W[I] + W[3:3 + nx, 3:3 + ny, 2:2 + nz]) / 2.0
    np.testing.assert_array_equal(b["u"], np.stack([uc, vc, wc]).astype(np.float32))
    _, bs = parse(tmp_path / "scalars.vtk")
    np.testing.assert_array_equal(bs["scalar01"], S[0][I].astype(np.float32))
    np.testing.assert_array_equal(bs["scalar02"], S[1][I].astype(np.float32))
    # staggered files: U on xU(1..nx) (periodic x), W on zW(1..nz-1) (walls)
    _, bu = parse(tmp_path / "U.vtk")
    np.testing.assert_array_equal(bu["X"], (np.arange(1, nx + 1) * 0.5).astype(np.float32))
    np.testing.assert_array_equal(bu["U"], U[I].astype(np.float32))
    _, bw = parse(tmp_path / "W.vtk")
    assert bw["W"].shape == (nx, ny, nz - 1)
    np.testing.assert_array_equal(bw["Z"], (np.arange(1, nz) * 0.5).astype(np.float32))


def test_restart_round_trip(tmp_path, fields):
    cfg, U, V, W, Pr, S, _ = fields
    capi.vtk_write(str(tmp_path), cfg, U, V, W, Pr, S)
    U2, V2, W2, P2, S2 = capi.vtk_read(str(tmp_path), cfg, nscalars=2)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    f32 = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    I = np.s_[3:3 + nx, 3:3 + ny, 3:3 + nz]
    assert np.array_equal(P2[I], f32(Pr[I]))
    assert np.array_equal(S2[1][I], f32(S[1][I]))
    assert np.array_equal(U2[I], f32(U[I]))
    assert np.array_equal(W2[3:3 + nx, 3:3 + ny, 3:2 + nz], f32(W[3:3 + nx, 3:3 + ny, 3:2 + nz]))


def test_restart_from_cell_centred_vectors(tmp_path, fields):
    """Without U/V/W.vtk the faces are rebuilt from out.vtk's vectors as
    (u_c(i) + u_c(i+1)) / 2, the periodic last face from u_c(n) and u_c(1)."""
    cfg, U, V, W, Pr, S, _ = fields
    capi.vtk_write(str(tmp_path), cfg, U, V, W, Pr, S)
    for f in ("U.vtk", "V.vtk", "W.vtk"):
        os.remove(tmp_path / f)
    _, b = parse(tmp_path / "out.vtk")
    uc = b["u"].astype(np.float64)
    U2, V2, W2, _, _ = capi.vtk_read(str(tmp_path), cfg, nscalars=2)
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    want_u = (uc[0] + np.roll(uc[0], -1, axis=0)) / 2
    np.testing.assert_array_equal(U2[3:3 + nx, 3:3 + ny, 3:3 + nz], want_u)
    want_v = (uc[1] + np.roll(uc[1], -1, axis=1)) / 2
    np.testing.assert_array_equal(V2[3:3 + nx, 3:3 + ny, 3:3 + nz], want_v)
    want_w = (uc[2][:, :, :-1] + uc[2][:, :, 1:]) / 2          # top is a wall: no wrap
    np.testing.assert_array_equal(W2[3:3 + nx, 3:3 + ny, 3:2 + nz], want_w)
