// solids.hip - solid bodies on the host: type masks, immersed-boundary
// interpolation lists, solid-body wall-model points and nullified points, for
// axis-aligned cuboids.  The per-point geometry is in solid_geom.hpp, shared
// with the device builder (k_solids.hip); this file holds the host loops,
// in the reference's loop and list order:
//   FindInsideCells            src/solidbodies.f90:75-206
//   SolidBody_Init             src/solidbodies.f90:439-454 (rough = z0 > 0)
//   FindNeighbouringCells      src/immersedboundary.f90:1475-1506
//   InitImBoundaries           src/immersedboundary.f90:1650-1720
//   GetSolidBodiesWM           src/immersedboundary.f90:14-139
//   GetSolidBodiesWM_UVW       src/immersedboundary.f90:144-308
//   SetNullifiedPoints         src/initial.f90:4861-4962
// The inside-cell search visits only the index ranges whose coordinates can
// satisfy a cuboid's per-axis half-space pair (each plane of an axis-aligned
// cuboid depends on one coordinate, so the result is the same as testing
// every point).  Builders that fill different arrays or lists (the four type
// arrays, the four IB lists, the WM lists of Pr and of each velocity
// component, the nullified points of each component) run on their own
// threads; each list is still filled in the reference's order, so the
// result does not depend on the threads.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <exception>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "case_setup.hpp"
#include "errors.hpp"
#include "solid_geom.hpp"

namespace elmm {

sg::Geom make_geom(const elmm_config &c, const Grid1D &g, const CaseData &cd) {
  sg::Geom G{};
  G.nx = c.nx;
  G.ny = c.ny;
  G.nz = c.nz;
  component_extents(c, G.n);
  for (int s = 0; s < 6; s++) G.per[s] = c.btype[s] == ELMM_BC_PERIODIC;
  G.dxmin = g.dxmin;
  G.dymin = g.dymin;
  G.dzmin = g.dzmin;
  G.ib_cut = 0.1 * std::pow(g.dxmin * g.dymin * g.dzmin, 1.0 / 3.0);
  G.sy = cd.sy;
  G.sz = cd.sz;
  G.xU = g.xU.data();
  G.yV = g.yV.data();
  G.zW = g.zW.data();
  G.xPr = g.xPr.data();
  G.yPr = g.yPr.data();
  G.zPr = g.zPr.data();
  return G;
}

int make_bodies(int ncub, const elmm_cuboid *cub, std::vector<sg::Body> &bodies) {
  bodies.clear();
  for (int b = 0; b < ncub; b++) {
    const elmm_cuboid &q = cub[b];
    if (!(q.xmin < q.xmax && q.ymin < q.ymax && q.zmin < q.zmax) || !(q.z0 >= 0))
      return fail(ELMM_E_ARG, "cuboid " + std::to_string(b) + ": empty extent or negative z0");
    bodies.push_back(sg::make_body(b + 1, q.xmin, q.xmax, q.ymin, q.ymax, q.zmin, q.zmax, q.z0));
    // a rough body's surface fluxes (SolidBody%rough: z0 > 0)
    if (q.z0 > 0) {
      bodies.back().tflux = q.temperature_flux;
      bodies.back().qflux = q.moisture_flux;
    }
  }
  return 0;
}

// Thermal class, value and moisture flux of a solid body's wall-model point:
// its heat flux with buoyancy (the flux class when non-zero), its moisture
// flux with moisture (src/immersedboundary.f90:111-125, 260-285)
void body_thermal(const elmm_config &c, const sg::Body &B, int8_t &tc, double &tv, double &qf) {
  tc = ELMM_WM_NEUTRAL;
  tv = 0;
  qf = 0;
  if (c.enable_buoyancy && B.tflux != 0) {
    tc = ELMM_WM_SURFACE_FLUX;
    tv = B.tflux;
  }
  if (c.enable_buoyancy && c.enable_moisture) qf = B.qflux;
}

// Inside-cell tolerance and minimum wall distance: (dxmin dymin dzmin)^(1/3) / 20
double solid_delta(const Grid1D &g) { return std::pow(g.dxmin * g.dymin * g.dzmin, 1.0 / 3.0) / 20; }

namespace {

// FindInsideCells for one type array on the (x, y, z) coordinate arrays
void set_type(const sg::Geom &G, const sg::Body &B, CaseData &cd, std::vector<int32_t> &T, const std::vector<double> &X,
              const std::vector<double> &Y, const std::vector<double> &Z, const int ext[3], double delta) {
  auto hits = [&](int axis, const std::vector<double> &coord, int hi) {
    std::vector<char> h(hi + 3, 0);
    for (int i = -2; i <= hi; i++) h[i + 2] = sg::axis_hit(G, B, axis, Grid1D::at(coord, i), delta);
    return h;
  };
  auto hx = hits(0, X, ext[0] + 3);
  auto hy = hits(1, Y, ext[1] + 3);
  auto hz = hits(2, Z, ext[2] + 3);
  for (int k = -2; k <= ext[2] + 3; k++) {
    if (!hz[k + 2]) continue;
    for (int j = -2; j <= ext[1] + 3; j++) {
      if (!hy[j + 2]) continue;
      for (int i = -2; i <= ext[0] + 3; i++)
        if (hx[i + 2]) T[cd.lin(i, j, k)] = B.id;
    }
  }
}

// AuxNeighbours: free cells (i, j, k in -1..ext+2) with a solid cell among
// their 27 neighbours become -1.  Solid cells are never changed, so the test
// reads the solid flags of the input only, and "a neighbour is solid" is an
// OR over the 3x3x3 box, computed as three one-axis ORs (x, then y, then z).
void aux_neighbours(CaseData &cd, std::vector<int32_t> &T, const int ext[3]) {
  const long n = cd.ntot;
  std::vector<uint8_t> a(n, 0), b(n, 0);
  for (long p = 0; p < n; p++) a[p] = T[p] > 0;
  for (int k = -2; k <= ext[2] + 3; k++)
    for (int j = -2; j <= ext[1] + 3; j++) {
      const long r = cd.lin(0, j, k);
      for (int i = -1; i <= ext[0] + 2; i++) b[r + i] = a[r + i - 1] | a[r + i] | a[r + i + 1];
    }
  const long sy = cd.sy, sz = cd.sz;
  for (int k = -2; k <= ext[2] + 3; k++)
    for (int j = -1; j <= ext[1] + 2; j++) {
      const long r = cd.lin(0, j, k);
      for (int i = -1; i <= ext[0] + 2; i++) a[r + i] = b[r + i - sy] | b[r + i] | b[r + i + sy];
    }
  for (int k = -1; k <= ext[2] + 2; k++)
    for (int j = -1; j <= ext[1] + 2; j++) {
      const long r = cd.lin(0, j, k);
      for (int i = -1; i <= ext[0] + 2; i++)
        if (T[r + i] == 0 && (a[r + i - sz] | a[r + i] | a[r + i + sz])) T[r + i] = -1;
    }
}

void push_ib(IBList &L, int i, int j, int k, const sg::Tap *taps, int n) {
  L.ijk.insert(L.ijk.end(), {i, j, k});
  for (int s = 0; s < n; s++) {
    L.src.insert(L.src.end(), {taps[s].i, taps[s].j, taps[s].k});
    L.coef.push_back(taps[s].coef);
  }
  L.ptr.push_back((int32_t)L.coef.size());
}

// GetSolidBodiesWM: free Pr points with a solid face neighbour (only cells
// marked -1 by FindNeighbouringCells can have one).  Returns an error
// message, empty on success.
std::string solid_wm_pr(const sg::Geom &G, const std::vector<sg::Body> &bodies, CaseData &cd,
                        double min_wall_distance) {
  for (int k = 1; k <= G.nz; k++)
    for (int j = 1; j <= G.ny; j++)
      for (int i = 1; i <= G.nx; i++) {
        if (cd.prtype[cd.lin(i, j, k)] >= 0) continue;
        double dv[3], z0 = 0;
        int bid = 0;
        const int r =
            sg::wm_pr_point(G, bodies.data(), cd.prtype.data(), i, j, k, min_wall_distance, dv, &z0, &bid);
        if (r < 0) return sg::solid_error(-r);
        if (r == 0) continue;
        int8_t tc;
        double tv, qf;
        body_thermal(cd.cfg, bodies[bid - 1], tc, tv, qf);
        cd.wmp.push(i, j, k, dv[0], dv[1], dv[2], z0, tc, tv, 0.0, qf);
      }
  return std::string();
}

// GetSolidBodiesWM_UVW for one velocity component: -1 points, per direction.
std::string solid_wm_uvw(const sg::Geom &G, const std::vector<sg::Body> &bodies, CaseData &cd, int comp,
                         const std::vector<int32_t> &T, const std::vector<double> &cx, const std::vector<double> &cy,
                         const std::vector<double> &cz, double min_wall_distance) {
  for (int k = 1; k <= G.n[comp][2]; k++)
    for (int j = 1; j <= G.n[comp][1]; j++)
      for (int i = 1; i <= G.n[comp][0]; i++) {
        if (T[cd.lin(i, j, k)] >= 0) continue;
        for (int dir = 0; dir < 6; dir++) {
          double dv[3], z0 = 0;
          int bid = 0;
          const int r = sg::wm_uvw_dir(G, bodies.data(), T.data(), i, j, k, dir, cx.data(), cy.data(), cz.data(),
                                       min_wall_distance, dv, &z0, &bid);
          if (r < 0) return sg::solid_error(-r);
          if (r == 0) continue;
          int8_t tc;
          double tv, qf;
          body_thermal(cd.cfg, bodies[bid - 1], tc, tv, qf);
          cd.wm[comp * 6 + dir].push(i, j, k, dv[0], dv[1], dv[2], z0, tc, tv, 0.0, qf);
        }
      }
  return std::string();
}

}  // namespace

void run_tasks(int n, const std::function<void(int)> &f) {
  std::vector<std::exception_ptr> ex(n);
  auto one = [&](int a) {
    try {
      f(a);
    } catch (...) {
      ex[a] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  for (int a = 1; a < n; a++) th.emplace_back(one, a);
  one(0);
  for (auto &t : th) t.join();
  for (auto &e : ex)
    if (e) std::rethrow_exception(e);
}

int build_solids(const elmm_config &c, const Grid1D &g, int ncub, const elmm_cuboid *cub, CaseData &cd) {
  std::vector<sg::Body> bodies;
  if (int rc = make_bodies(ncub, cub, bodies)) return rc;
  const sg::Geom G = make_geom(c, g, cd);
  const double delta = solid_delta(g);
  int ext[4][3];
  for (int d = 0; d < 3; d++) {
    ext[0][d] = d == 0 ? c.nx : d == 1 ? c.ny : c.nz;
    for (int comp = 0; comp < 3; comp++) ext[comp + 1][d] = G.n[comp][d];
  }
  // per-array and per-component tasks: T[0] Pr, T[1..3] U, V, W
  std::vector<int32_t> *T[4] = {&cd.prtype, &cd.utype, &cd.vtype, &cd.wtype};
  const std::vector<double> *cx[4] = {&g.xPr, &g.xU, &g.xPr, &g.xPr}, *cy[4] = {&g.yPr, &g.yPr, &g.yV, &g.yPr},
                            *cz[4] = {&g.zPr, &g.zPr, &g.zPr, &g.zW};
  // FindInsideCells (body-major per array) and FindNeighbouringCells
  run_tasks(4, [&](int a) {
    for (const sg::Body &B : bodies) set_type(G, B, cd, *T[a], *cx[a], *cy[a], *cz[a], ext[a], delta);
    aux_neighbours(cd, *T[a], ext[a]);
  });
  // InitImBoundaries (serial k, j, i order per list); lists 0-2 velocity,
  // 3 scalar flux
  run_tasks(4, [&](int a) {
    IBList &L = cd.ib[a];
    L = IBList();
    sg::Tap taps[6];
    if (a < 3) {
      const std::vector<int32_t> &Tc = *T[a + 1];
      for (int k = 1; k <= G.n[a][2]; k++)
        for (int j = 1; j <= G.n[a][1]; j++)
          for (int i = 1; i <= G.n[a][0]; i++)
            if (Tc[cd.lin(i, j, k)] > 0 && sg::has_free_face_neighbour(G, Tc.data(), i, j, k)) {
              const int n = sg::vel_ib_point(G, bodies.data(), Tc.data(), i, j, k, cx[a + 1]->data(),
                                             cy[a + 1]->data(), cz[a + 1]->data(), taps);
              push_ib(L, i, j, k, taps, n);
            }
    } else {
      for (int k = 1; k <= c.nz; k++)
        for (int j = 1; j <= c.ny; j++)
          for (int i = 1; i <= c.nx; i++)
            if (cd.prtype[cd.lin(i, j, k)] > 0 && sg::has_free_face_neighbour(G, cd.prtype.data(), i, j, k))
              push_ib(L, i, j, k, taps, sg::scalfl_ib_point(G, bodies.data(), cd.prtype.data(), i, j, k, taps));
    }
  });

  // wall-model points (task 0: Pr, tasks 1-3: U, V, W) and nullified points
  // (tasks 1-3); a task that finds an error stops and keeps its message, and
  // the first one in task order is reported (the one the serial order meets
  // first)
  std::string errs[4];
  const double min_wall_distance = delta;
  run_tasks(4, [&](int a) {
    if (c.wall_model && a == 0) errs[0] = solid_wm_pr(G, bodies, cd, min_wall_distance);
    if (a == 0) return;
    const int comp = a - 1;
    if (c.wall_model) errs[a] = solid_wm_uvw(G, bodies, cd, comp, *T[a], *cx[a], *cy[a], *cz[a], min_wall_distance);
    if (!errs[a].empty()) return;
    // SetNullifiedPoints: velocity points whose 7-point stencil is solid
    const std::vector<int32_t> &Tc = *T[a];
    cd.null_ijk[comp].clear();
    for (int k = 1; k <= G.n[comp][2]; k++)
      for (int j = 1; j <= G.n[comp][1]; j++)
        for (int i = 1; i <= G.n[comp][0]; i++)
          if (sg::null_point(G, Tc.data(), i, j, k)) cd.null_ijk[comp].insert(cd.null_ijk[comp].end(), {i, j, k});
  });
  for (const std::string &e : errs)
    if (!e.empty()) return fail(ELMM_E_ARG, e);
  return 0;
}

}  // namespace elmm
