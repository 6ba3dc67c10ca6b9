// case_setup.cpp - host-side construction of the masks and integer point
// lists that the time step consumes (no GPU involved).
//
// Restates the list builders of the reference's initialisation:
//   grid coordinates          src/initial.f90:4395-4502
//   outer-wall WM points      GetOutsideBoundariesWM_Pr/_UVW,
//                             src/wallmodels.f90:467-904
//   duplicate removal         RemoveDuplicateWMPoints, src/wallmodels.f90:373-438
//   scalar wall-face masks    InitWMMasks, src/wallmodels.f90:911-1113
//   point sources             point_source + GridCoords,
//                             src/pointsources.f90:20-41, src/boundaries.f90:54-92
//   solid bodies              cuboid masks + IB/WM/null lists, see solids.cpp
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "case_setup.hpp"
#include "elmm_gpu.h"
#include "errors.hpp"

namespace elmm {

Grid1D make_grid(const elmm_config &c) {
  Grid1D g;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  auto alloc = [](std::vector<double> &v, int n) { v.assign(n + 8, 0.0); };
  alloc(g.xU, nx); alloc(g.yV, ny); alloc(g.zW, nz);
  alloc(g.xPr, nx); alloc(g.yPr, ny); alloc(g.zPr, nz);
  alloc(g.dxPr, nx); alloc(g.dyPr, ny); alloc(g.dzPr, nz);
  alloc(g.dxU, nx); alloc(g.dyV, ny); alloc(g.dzW, nz);
  g.dxmin = c.dx;
  g.dymin = c.dy;
  for (int i = -3; i <= nx + 3; i++) g.at(g.xU, i) = i * g.dxmin + c.x0;
  for (int j = -3; j <= ny + 3; j++) g.at(g.yV, j) = j * g.dymin + c.y0;
  if (c.stretched_z) {
    for (int k = 0; k <= nz; k++) g.at(g.zW, k) = c.zw[k];
    if (c.btype[ELMM_BO] == ELMM_BC_PERIODIC) {
      for (int j = -1; j >= -3; j--) g.at(g.zW, j) = g.at(g.zW, 0) - (g.at(g.zW, nz - 1) - g.at(g.zW, nz - 1 + j));
    } else {
      for (int j = -1; j >= -3; j--) g.at(g.zW, j) = g.at(g.zW, 0) - (g.at(g.zW, -j) - g.at(g.zW, 0));
    }
    if (c.btype[ELMM_TO] == ELMM_BC_PERIODIC) {
      for (int j = nz + 1; j <= nz + 3; j++)
        g.at(g.zW, j) = g.at(g.zW, nz - 1) + (g.at(g.zW, j - nz + 1) - g.at(g.zW, 0));
    } else {
      for (int j = nz + 1; j <= nz + 3; j++)
        g.at(g.zW, j) = g.at(g.zW, nz - 1) + (g.at(g.zW, nz - 1) - g.at(g.zW, nz - 1 - (j - nz + 1)));
    }
  } else {
    g.dzmin = c.dz;
    for (int k = -3; k <= nz + 3; k++) g.at(g.zW, k) = k * g.dzmin + c.z0;
  }
  for (int i = -2; i <= nx + 3; i++) {
    g.at(g.xPr, i) = (g.at(g.xU, i - 1) + g.at(g.xU, i)) / 2.0;
    g.at(g.dxPr, i) = g.at(g.xU, i) - g.at(g.xU, i - 1);
  }
  for (int j = -2; j <= ny + 3; j++) {
    g.at(g.yPr, j) = (g.at(g.yV, j - 1) + g.at(g.yV, j)) / 2.0;
    g.at(g.dyPr, j) = g.at(g.yV, j) - g.at(g.yV, j - 1);
  }
  for (int k = -2; k <= nz + 3; k++) {
    g.at(g.zPr, k) = (g.at(g.zW, k - 1) + g.at(g.zW, k)) / 2.0;
    g.at(g.dzPr, k) = g.at(g.zW, k) - g.at(g.zW, k - 1);
  }
  for (int i = -2; i <= nx + 2; i++) g.at(g.dxU, i) = g.at(g.xPr, i + 1) - g.at(g.xPr, i);
  for (int j = -2; j <= ny + 2; j++) g.at(g.dyV, j) = g.at(g.yPr, j + 1) - g.at(g.yPr, j);
  for (int k = -2; k <= nz + 2; k++) g.at(g.dzW, k) = g.at(g.zPr, k + 1) - g.at(g.zPr, k);
  if (c.stretched_z) {
    double m = g.at(g.dzPr, -2);
    for (int k = -2; k <= nz + 3; k++) m = std::min(m, g.at(g.dzPr, k));
    g.dzmin = m;
  }
  return g;
}

void component_extents(const elmm_config &c, int n[3][3]) {
  // src/initial.f90:871-896
  n[0][0] = c.btype[ELMM_EA] == ELMM_BC_PERIODIC ? c.nx : c.nx - 1;
  n[0][1] = c.ny;
  n[0][2] = c.nz;
  n[1][0] = c.nx;
  n[1][1] = c.btype[ELMM_NO] == ELMM_BC_PERIODIC ? c.ny : c.ny - 1;
  n[1][2] = c.nz;
  n[2][0] = c.nx;
  n[2][1] = c.ny;
  n[2][2] = c.btype[ELMM_TO] == ELMM_BC_PERIODIC ? c.nz : c.nz - 1;
}

bool mo_moist_flux_wall(const elmm_config &c, int s) {
  return c.enable_buoyancy && c.enable_moisture && c.wall_model && c.btype[s] == ELMM_BC_NOSLIP &&
         c.moist_btype[s] == ELMM_BC_CONSTFLUX;
}

bool mo_wall(const elmm_config &c, int s) {
  return c.enable_buoyancy && c.wall_model && c.btype[s] == ELMM_BC_NOSLIP &&
         (c.temp_btype[s] == ELMM_BC_DIRICHLET || c.temp_btype[s] == ELMM_BC_CONSTFLUX);
}

std::string validate_config(const elmm_config &c) {
  if (c.nx < 4 || c.ny < 4 || c.nz < 4) return "grid must be at least 4 cells in every direction";
  if (!(c.dx > 0) || !(c.dy > 0)) return "dx and dy must be positive";
  if (!c.stretched_z && !(c.dz > 0)) return "dz must be positive for a uniform z grid";
  if (c.stretched_z && !c.zw) return "stretched_z requires zw";
  for (int s = 0; s < 6; s++) {
    int b = c.btype[s];
    if (b != ELMM_BC_NOSLIP && b != ELMM_BC_FREESLIP && b != ELMM_BC_PERIODIC)
      return "velocity boundary type " + std::to_string(b) + " on side " + std::to_string(s) +
             " is not supported (noslip, freeslip, periodic)";
  }
  for (int a = 0; a < 3; a++)
    if ((c.btype[2 * a] == ELMM_BC_PERIODIC) != (c.btype[2 * a + 1] == ELMM_BC_PERIODIC))
      return "periodic boundaries must be paired";
  if (c.btype[ELMM_WE] != ELMM_BC_PERIODIC || c.btype[ELMM_SO] != ELMM_BC_PERIODIC ||
      c.btype[ELMM_BO] == ELMM_BC_PERIODIC)
    return "the pressure solver needs periodic x and y and walls in z";
  for (int s = 0; s < 6; s++) {
    int b = c.scal_btype[s];
    if (c.nscalars > 0 && b != ELMM_BC_PERIODIC && b != ELMM_BC_NEUMANN && b != ELMM_BC_DIRICHLET)
      return "scalar boundary type must be periodic, Neumann or Dirichlet";
  }
  if (c.stretched_z) {
    for (int k = 0; k < c.nz; k++)
      if (!(c.zw[k + 1] > c.zw[k])) return "zw must increase strictly";
    // CD4div uses the uniform dzmin (src/advection.f90:473); the stretched
    // grid has the variable-z CD2 of src/advection_variable_z.f90 only
    if (c.advection_method != 2) return "stretched z grids need advection_method 2 (variable-z CD2)";
    if (c.nscalars > 0) return "passive scalars on stretched z grids are not built in this version";
  } else if (c.advection_method != 4) {
    return "advection_method must be 4 (CD4div) on uniform grids";
  }
  if (c.sgs_model != 0 && c.sgs_model != 1) return "sgs_model must be 0 or 1";
  if (c.nscalars < 0 || c.nscalars > 8) return "nscalars must be in 0..8";
  if (c.precision != ELMM_FP64 && c.precision != ELMM_FP32_REFINED) return "unknown precision";
  if (c.enable_moisture && !c.enable_buoyancy) return "moisture needs enable_buoyancy (the temperature)";
  // A wall with wall-model points (no-slip, wall model on) takes a Dirichlet
  // or flux temperature into its points (prescribed surface temperature or
  // heat flux: the Monin-Obukhov laws, src/wallmodels.f90:592-622,
  // 2086-2118); a flux temperature or moisture exists only there.
  if (c.enable_buoyancy)
    for (int s = ELMM_BO; s <= ELMM_TO; s++) {
      const bool wm_wall = c.wall_model && c.btype[s] == ELMM_BC_NOSLIP;
      if (c.temp_btype[s] == ELMM_BC_CONSTFLUX && !wm_wall)
        return "a temperature flux (ELMM_BC_CONSTFLUX) needs a no-slip wall with the wall model (side " +
               std::to_string(s) + ")";
      if (c.enable_moisture && c.moist_btype[s] == ELMM_BC_CONSTFLUX && !wm_wall)
        return "a moisture flux (ELMM_BC_CONSTFLUX) needs a no-slip wall with the wall model (side " +
               std::to_string(s) + ")";
    }
  if (c.variable_dt) {
    if (!(c.cfl_target > 0) || !(c.dt_max > 0)) return "variable time steps need cfl_target > 0 and dt_max > 0";
  } else if (!(c.dt > 0)) {
    return "dt must be positive";
  }
  return "";
}

namespace {

struct UVWPoint {
  int i, j, k;
  double d[3];
  double z0;
};

// Thermal class of the wall-model points of domain side s (the velocity
// lists' direction dir equals the side): a prescribed surface temperature
// (Dirichlet) or heat flux (CONSTFLUX, positive upward: the top wall's is
// -sideTemp(To)) on a Monin-Obukhov wall, else neutral
// (src/wallmodels.f90:592-602, 652-662, 821-831, 872-882).
void wall_thermal(const elmm_config &c, int s, int8_t &tc, double &tv, double &qv, double &qf) {
  tc = ELMM_WM_NEUTRAL;
  tv = 0;
  qv = 0;
  qf = 0;
  if (s < ELMM_BO) return;
  // surface moisture (Dirichlet) or moisture flux (flux; the top's is
  // -sideMoist(To)) of a no-slip wall-model wall (wallmodels.f90:604-612)
  if (c.enable_buoyancy && c.enable_moisture && c.wall_model && c.btype[s] == ELMM_BC_NOSLIP) {
    if (c.moist_btype[s] == ELMM_BC_DIRICHLET) qv = c.side_moist[s];
    if (c.moist_btype[s] == ELMM_BC_CONSTFLUX) qf = s == ELMM_BO ? c.side_moist[s] : -c.side_moist[s];
  }
  if (!mo_wall(c, s)) return;
  if (c.temp_btype[s] == ELMM_BC_DIRICHLET) {
    tc = ELMM_WM_SURFACE_TEMP;
    tv = c.side_temp[s];
  } else {
    tc = ELMM_WM_SURFACE_FLUX;
    tv = s == ELMM_BO ? c.side_temp[s] : -c.side_temp[s];
  }
}

// GetOutsideBoundariesWM_UVW helper, src/wallmodels.f90:640-904
void outer_uvw(const elmm_config &c, const Grid1D &g, CaseData &cd, int comp) {
  const int *n = cd.n[comp];
  const int nx = n[0], ny = n[1], nz = n[2];
  const std::vector<int32_t> &T = comp == 0 ? cd.utype : comp == 1 ? cd.vtype : cd.wtype;
  auto X = [&](int i) { return comp == 0 ? g.at(g.xU, i) : g.at(g.xPr, i); };
  auto Y = [&](int j) { return comp == 1 ? g.at(g.yV, j) : g.at(g.yPr, j); };
  auto Z = [&](int k) { return comp == 2 ? g.at(g.zW, k) : g.at(g.zPr, k); };
  auto add = [&](int dir, int i, int j, int k, double dx, double dy, double dz, double z0) {
    int8_t tc;
    double tv, qv, qf;
    wall_thermal(c, dir, tc, tv, qv, qf);
    cd.wm[comp * 6 + dir].push(i, j, k, dx, dy, dz, z0, tc, tv, qv, qf);
  };
  auto t = [&](int i, int j, int k) { return T[cd.lin(i, j, k)]; };
  if (c.btype[ELMM_WE] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        if (t(1, j, k) == 0 && t(0, j, k) <= 0) add(0, 1, j, k, g.at(g.xU, 0) - X(1), 0, 0, c.z0_wall[ELMM_WE]);
  if (c.btype[ELMM_EA] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        if (t(nx, j, k) == 0 && t(nx + 1, j, k) <= 0)
          add(1, nx, j, k, g.at(g.xU, cd.n[0][0] + 1) - X(nx), 0, 0, c.z0_wall[ELMM_EA]);
  if (c.btype[ELMM_SO] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++)
        if (t(i, 1, k) == 0 && t(i, 0, k) <= 0) add(2, i, 1, k, 0, g.at(g.yV, 0) - Y(1), 0, c.z0_wall[ELMM_SO]);
  if (c.btype[ELMM_NO] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++)
        if (t(i, ny, k) == 0 && t(i, ny + 1, k) <= 0)
          add(3, i, ny, k, 0, g.at(g.yV, cd.n[1][1] + 1) - Y(ny), 0, c.z0_wall[ELMM_NO]);
  if (c.btype[ELMM_BO] == ELMM_BC_NOSLIP)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        if (t(i, j, 1) == 0 && t(i, j, 0) <= 0) add(4, i, j, 1, 0, 0, g.at(g.zW, 0) - Z(1), c.z0_wall[ELMM_BO]);
  if (c.btype[ELMM_TO] == ELMM_BC_NOSLIP)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        if (t(i, j, nz) == 0 && t(i, j, nz + 1) <= 0)
          add(5, i, j, nz, 0, 0, g.at(g.zW, cd.n[2][2] + 1) - Z(nz), c.z0_wall[ELMM_TO]);
}

// GetOutsideBoundariesWM_Pr, src/wallmodels.f90:467-637
void outer_pr(const elmm_config &c, const Grid1D &g, CaseData &cd) {
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  auto add = [&](int side, int i, int j, int k, double dx, double dy, double dz, double z0) {
    int8_t tc;
    double tv, qv, qf;
    wall_thermal(c, side, tc, tv, qv, qf);
    cd.wmp.push(i, j, k, dx, dy, dz, z0, tc, tv, qv, qf);
  };
  auto t = [&](int i, int j, int k) { return cd.prtype[cd.lin(i, j, k)]; };
  if (c.btype[ELMM_WE] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        if (t(1, j, k) == 0) add(ELMM_WE, 1, j, k, g.at(g.xU, 0) - g.at(g.xPr, 1), 0, 0, c.z0_wall[ELMM_WE]);
  if (c.btype[ELMM_EA] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        if (t(nx, j, k) == 0)
          add(ELMM_EA, nx, j, k, g.at(g.xU, cd.n[0][0] + 1) - g.at(g.xPr, nx), 0, 0, c.z0_wall[ELMM_EA]);
  if (c.btype[ELMM_SO] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++)
        if (t(i, 1, k) == 0) add(ELMM_SO, i, 1, k, 0, g.at(g.yV, 0) - g.at(g.yPr, 1), 0, c.z0_wall[ELMM_SO]);
  if (c.btype[ELMM_NO] == ELMM_BC_NOSLIP)
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++)
        if (t(i, ny, k) == 0)
          add(ELMM_NO, i, ny, k, 0, g.at(g.yV, cd.n[1][1] + 1) - g.at(g.yPr, ny), 0, c.z0_wall[ELMM_NO]);
  if (c.btype[ELMM_BO] == ELMM_BC_NOSLIP)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        if (t(i, j, 1) == 0) add(ELMM_BO, i, j, 1, 0, 0, g.at(g.zW, 0) - g.at(g.zPr, 1), c.z0_wall[ELMM_BO]);
  if (c.btype[ELMM_TO] == ELMM_BC_NOSLIP)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++)
        if (t(i, j, nz) == 0)
          add(ELMM_TO, i, j, nz, 0, 0, g.at(g.zW, cd.n[2][2] + 1) - g.at(g.zPr, nz), c.z0_wall[ELMM_TO]);
}

}  // namespace

// RemoveDuplicateWMPoints, src/wallmodels.f90:373-438.  Points are ordered by
// the linear cell index xi+(yj-1)*Prnx+(zk-1)*Prnx*Prny, then by distance;
// the first of each cell is kept.  The reference sorts with C qsort, whose
// order among equal keys is unspecified; a stable sort (earlier-added first)
// is used here.
void remove_duplicate_wm_points(const elmm_config &c, WMList &L) {
  size_t n = L.z0.size();
  if (n < 2) return;
  std::vector<size_t> idx(n);
  for (size_t p = 0; p < n; p++) idx[p] = p;
  auto key = [&](size_t p) {
    return (long)L.ijk[3 * p] + (long)(L.ijk[3 * p + 1] - 1) * c.nx + (long)(L.ijk[3 * p + 2] - 1) * c.nx * c.ny;
  };
  auto d2 = [&](size_t p) {
    const double *d = &L.dist[3 * p];
    return d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
  };
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
    long ka = key(a), kb = key(b);
    if (ka != kb) return ka < kb;
    return d2(a) < d2(b);
  });
  WMList out;
  for (size_t q = 0; q < n; q++) {
    size_t p = idx[q];
    if (q > 0) {
      size_t r = idx[q - 1];
      if (L.ijk[3 * r] == L.ijk[3 * p] && L.ijk[3 * r + 1] == L.ijk[3 * p + 1] && L.ijk[3 * r + 2] == L.ijk[3 * p + 2])
        continue;
    }
    out.push(L.ijk[3 * p], L.ijk[3 * p + 1], L.ijk[3 * p + 2], L.dist[3 * p], L.dist[3 * p + 1], L.dist[3 * p + 2],
             L.z0[p], L.tclass[p], L.tval[p], L.qval[p], L.qflux[p]);
  }
  L = std::move(out);
}

// InitWMMasks scalar part, src/wallmodels.f90:911-1021
static void init_scalar_masks(const elmm_config &c, CaseData &cd) {
  for (int d = 0; d < 3; d++) cd.scfl_mask[d].assign(cd.ntot, 1);
  for (int d = 0; d < 3; d++) cd.scfl[d].clear();
  if (c.wall_model == 0) return;
  const int nx = c.nx, ny = c.ny, nz = c.nz;
  // face d of cell (i, j, k) lies between it and its +d neighbour: masked
  // when either side is solid, listed when exactly one is (one task per
  // direction, each in k, j, i order)
  run_tasks(3, [&](int d) {
    const long off = d == 0 ? 1 : d == 1 ? cd.sy : cd.sz;
    const int lo[3] = {d == 0 ? 0 : 1, d == 1 ? 0 : 1, d == 2 ? 0 : 1};
    const int32_t *pt = cd.prtype.data();
    uint8_t *mask = cd.scfl_mask[d].data();
    std::vector<int32_t> &list = cd.scfl[d];
    for (int k = lo[2]; k <= nz; k++)
      for (int j = lo[1]; j <= ny; j++)
        for (int i = lo[0]; i <= nx; i++) {
          const long p = cd.lin(i, j, k);
          const bool a = pt[p] > 0, b = pt[p + off] > 0;
          if (a || b) mask[p] = 0;
          if (a != b) list.insert(list.end(), {i, j, k});
        }
  });
}

// point_source + GridCoords, src/pointsources.f90:20-41, src/boundaries.f90:54-92
static void point_sources(const elmm_config &c, const Grid1D &g, CaseData &cd, int n,
                          const elmm_point_source *src) {
  for (int s = 0; s < n; s++) {
    const elmm_point_source &p = src[s];
    double xmin = g.at(g.xU, 0), xmax = g.at(g.xU, c.nx);
    double ymin = g.at(g.yV, 0), ymax = g.at(g.yV, c.ny);
    double zmin = g.at(g.zW, 0), zmax = g.at(g.zW, c.nz);
    if (!(p.x >= xmin && p.x <= xmax && p.y >= ymin && p.y <= ymax && p.z >= zmin && p.z <= zmax)) continue;
    int xi, yj, zk;
    auto nint = [](double v) { return (int)std::lround(v); };
    if (!c.stretched_z) {
      xi = std::min(std::max(nint((p.x - g.at(g.xU, 0)) / g.dxmin + 0.5), 1), c.nx);
      yj = std::min(std::max(nint((p.y - g.at(g.yV, 0)) / g.dymin + 0.5), 1), c.ny);
      zk = std::min(std::max(nint((p.z - g.at(g.zW, 0)) / g.dzmin + 0.5), 1), c.nz);
    } else {
      xi = c.nx;
      for (int i = 1; i <= c.nx; i++)
        if (g.at(g.xU, i) >= p.x) { xi = i; break; }
      yj = c.ny;
      for (int i = 1; i <= c.ny; i++)
        if (g.at(g.yV, i) >= p.y) { yj = i; break; }
      zk = c.nz;
      for (int i = 1; i <= c.nz; i++)
        if (g.at(g.zW, i) >= p.z) { zk = i; break; }
    }
    double vol = g.at(g.dxPr, xi) * g.at(g.dyPr, yj) * g.at(g.dzPr, zk);
    cd.src_ijk.insert(cd.src_ijk.end(), {xi, yj, zk});
    cd.src_sc.push_back(p.scalar);
    cd.src_flux.push_back(p.flux / vol);
  }
}

int build_case(const elmm_config &c, int ncub, const elmm_cuboid *cub, int nsrc, const elmm_point_source *src,
               CaseData &cd, bool device) {
  std::string err = validate_config(c);
  if (!err.empty()) return fail(ELMM_E_ARG, err);
  if (c.stretched_z && ncub > 0) return fail(ELMM_E_ARG, "solid bodies on stretched z grids are not built in this version");
  cd.cfg = c;
  cd.nx = c.nx;
  cd.ny = c.ny;
  cd.nz = c.nz;
  cd.sy = c.nx + 6;
  cd.sz = (long)(c.nx + 6) * (c.ny + 6);
  cd.ntot = cd.sz * (c.nz + 6);
  component_extents(c, cd.n);
  Grid1D g = make_grid(c);
  const bool on_device = device && ncub > 0;
  if (!on_device) {  // the device builder writes every element
    cd.prtype.assign(cd.ntot, 0);
    cd.utype.assign(cd.ntot, 0);
    cd.vtype.assign(cd.ntot, 0);
    cd.wtype.assign(cd.ntot, 0);
  }
  if (ncub > 0) {
    int rc = on_device ? build_solids_device(c, g, ncub, cub, cd) : build_solids(c, g, ncub, cub, cd);
    if (rc) return rc;
  }
  if (c.wall_model) {
    // GetOutsideBoundariesWM: UVW lists first (wallmodels.f90:455-459), solid
    // body WM points were appended by build_solids before (initial.f90:4658-4662)
    for (int comp = 0; comp < 3; comp++) outer_uvw(c, g, cd, comp);
    outer_pr(c, g, cd);
    remove_duplicate_wm_points(c, cd.wmp);
  }
  if (!(on_device && c.wall_model)) init_scalar_masks(c, cd);  // else built with the solids
  point_sources(c, g, cd, nsrc, src);
  return 0;
}

void view_case(const CaseData &cd, elmm_lists &v) {
  std::memset(&v, 0, sizeof v);
  v.prtype = cd.prtype.data();
  v.utype = cd.utype.data();
  v.vtype = cd.vtype.data();
  v.wtype = cd.wtype.data();
  for (int d = 0; d < 3; d++) v.scfl_mask[d] = cd.scfl_mask[d].data();
  for (int l = 0; l < 18; l++) {
    v.wm_n[l] = (int32_t)cd.wm[l].z0.size();
    v.wm_ijk[l] = cd.wm[l].ijk.data();
    v.wm_dist[l] = cd.wm[l].dist.data();
    v.wm_z0[l] = cd.wm[l].z0.data();
  }
  for (int l = 0; l < 18; l++) {
    v.wm_tclass[l] = cd.wm[l].tclass.data();
    v.wm_tval[l] = cd.wm[l].tval.data();
  }
  v.wmp_tclass = cd.wmp.tclass.data();
  v.wmp_tval = cd.wmp.tval.data();
  for (int l = 0; l < 18; l++) {
    v.wm_qval[l] = cd.wm[l].qval.data();
    v.wm_qflux[l] = cd.wm[l].qflux.data();
  }
  v.wmp_qval = cd.wmp.qval.data();
  v.wmp_qflux = cd.wmp.qflux.data();
  v.wmp_n = (int32_t)cd.wmp.z0.size();
  v.wmp_ijk = cd.wmp.ijk.data();
  v.wmp_dist = cd.wmp.dist.data();
  v.wmp_z0 = cd.wmp.z0.data();
  for (int d = 0; d < 3; d++) {
    v.scfl_n[d] = (int32_t)(cd.scfl[d].size() / 3);
    v.scfl_ijk[d] = cd.scfl[d].data();
  }
  for (int c = 0; c < 4; c++) {
    v.ib_n[c] = (int32_t)(cd.ib[c].ijk.size() / 3);
    v.ib_ijk[c] = cd.ib[c].ijk.data();
    v.ib_ptr[c] = cd.ib[c].ptr.data();
    v.ib_src_ijk[c] = cd.ib[c].src.data();
    v.ib_coef[c] = cd.ib[c].coef.data();
  }
  for (int c = 0; c < 3; c++) {
    v.null_n[c] = (int32_t)(cd.null_ijk[c].size() / 3);
    v.null_ijk[c] = cd.null_ijk[c].data();
  }
  v.src_n = (int32_t)cd.src_flux.size();
  v.src_ijk = cd.src_ijk.data();
  v.src_scalar = cd.src_sc.data();
  v.src_flux = cd.src_flux.data();
}

}  // namespace elmm

struct elmm_case {
  elmm::CaseData data;
};

namespace elmm {
// Device bytes of a case's surface lists once elmm_create has copied them:
// wall-model points (ijk, list id, dist, z0, ustar, fluxp, fluxm, log term:
// 64 B), their ordered correction terms (about two 16-byte terms and a
// 16-byte cell entry per point and pass), scalar WM and face points, the IB
// lists (target 12 B, pointer 4 B, taps 20 B each) with their mass-source
// terms and the ghost-delta list, and the nullified points.
int64_t case_list_bytes(const elmm_case *c) {
  if (!c) return 0;
  const CaseData &d = c->data;
  int64_t b = 0;
  for (int l = 0; l < 18; l++) b += (int64_t)(d.wm[l].ijk.size() / 3) * (64 + 2 * (32 + 16));
  b += (int64_t)(d.wmp.ijk.size() / 3) * (12 + 8 + 2 * 32);
  for (int a = 0; a < 3; a++) b += (int64_t)(d.scfl[a].size() / 3) * 2 * 32;
  for (int q = 0; q < 4; q++) {
    const int64_t n = (int64_t)(d.ib[q].ijk.size() / 3), taps = (int64_t)d.ib[q].coef.size();
    b += n * (12 + 4 + 8) + taps * 20 + (q < 3 ? n * (2 * 32 + 16 * 6) : 0);
  }
  for (int a = 0; a < 3; a++) b += (int64_t)d.null_ijk[a].size() * 4;
  b += (int64_t)d.src_flux.size() * 48;
  return b;
}
}  // namespace elmm

static int case_build(const elmm_config *cfg, int32_t n_cuboids, const elmm_cuboid *cuboids, int32_t n_sources,
                      const elmm_point_source *sources, elmm_case **out, bool device) {
  if (!cfg || !out || (n_cuboids > 0 && !cuboids) || (n_sources > 0 && !sources))
    return elmm::fail(ELMM_E_ARG, "null argument to elmm_case_build");
  *out = nullptr;
  try {
    auto *c = new elmm_case;
    int rc = elmm::build_case(*cfg, n_cuboids, cuboids, n_sources, sources, c->data, device);
    if (rc) {
      delete c;
      return rc;
    }
    *out = c;
    return 0;
  } catch (const std::bad_alloc &) {
    return elmm::fail(ELMM_E_NOMEM, "out of host memory in elmm_case_build");
  }
}

extern "C" int elmm_case_build(const elmm_config *cfg, int32_t n_cuboids, const elmm_cuboid *cuboids,
                               int32_t n_sources, const elmm_point_source *sources, elmm_case **out) {
  return case_build(cfg, n_cuboids, cuboids, n_sources, sources, out, false);
}

extern "C" int elmm_case_build_device(const elmm_config *cfg, int32_t n_cuboids, const elmm_cuboid *cuboids,
                                      int32_t n_sources, const elmm_point_source *sources, elmm_case **out) {
  return case_build(cfg, n_cuboids, cuboids, n_sources, sources, out, true);
}

extern "C" int elmm_case_view(const elmm_case *c, elmm_lists *view) {
  if (!c || !view) return elmm::fail(ELMM_E_ARG, "null argument to elmm_case_view");
  elmm::view_case(c->data, *view);
  return 0;
}

extern "C" void elmm_case_free(elmm_case *c) { delete c; }
