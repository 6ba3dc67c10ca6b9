// case_setup.hpp - host-side case description: grid coordinates, masks and
// point lists in the reference's conventions (1-based i,j,k; ghosted arrays
// with lower bound -2 and extents n+6, Fortran order).
#pragma once
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "elmm_gpu.h"

namespace elmm {

// 1-D grid coordinate arrays of InitBoundaryConditions (src/initial.f90:
// 4395-4502), stored with offset 3 so that index -3 is element 0.
struct Grid1D {
  int nx = 0, ny = 0, nz = 0;
  double dxmin = 0, dymin = 0, dzmin = 0;
  std::vector<double> xU, yV, zW, xPr, yPr, zPr, dxPr, dyPr, dzPr, dxU, dyV, dzW;
  static double &at(std::vector<double> &v, int i) { return v[i + 3]; }
  static double at(const std::vector<double> &v, int i) { return v[i + 3]; }
};

struct WMList {
  std::vector<int32_t> ijk;  // 3 per point
  std::vector<double> dist;  // 3 per point
  std::vector<double> z0;
  std::vector<int8_t> tclass;  // ELMM_WM_* per point
  std::vector<double> tval;    // surface temperature or heat flux per point
  std::vector<double> qval;    // surface moisture per point
  std::vector<double> qflux;   // initial moisture flux per point
  // a point (by default neutral and dry)
  void push(int i, int j, int k, double dx, double dy, double dz, double z, int8_t tc = 0, double tv = 0,
            double qv = 0, double qf = 0) {
    ijk.insert(ijk.end(), {i, j, k});
    dist.insert(dist.end(), {dx, dy, dz});
    z0.push_back(z);
    tclass.push_back(tc);
    tval.push_back(tv);
    qval.push_back(qv);
    qflux.push_back(qf);
  }
};

// The bottom (s = ELMM_BO) or top (ELMM_TO) wall takes the temperature into
// its wall-model points (a prescribed surface temperature or heat flux, the
// Monin-Obukhov laws) and the temperature's ghost rule there becomes the
// mirror (TempBtype = 0): buoyancy, the wall model, a no-slip wall and a
// Dirichlet or flux temperature (src/wallmodels.f90:574-625, 636-680).
bool mo_wall(const elmm_config &c, int s);
// The same wall takes a flux moisture into its points (MoistBtype = 0 there,
// src/wallmodels.f90:627-629): moisture, the wall model, a no-slip wall and
// a flux moisture.
bool mo_moist_flux_wall(const elmm_config &c, int s);

struct IBList {
  std::vector<int32_t> ijk;      // target, 3 per point
  std::vector<int32_t> ptr{0};   // n+1 offsets
  std::vector<int32_t> src;      // interpolation sources, 3 per tap
  std::vector<double> coef;      // coefficient per tap
};

struct CaseData {
  elmm_config cfg{};
  int nx = 0, ny = 0, nz = 0;
  int n[3][3] = {};  // component extents n[c][d]
  long sy = 0, sz = 0, ntot = 0;
  std::vector<int32_t> prtype, utype, vtype, wtype;
  std::vector<uint8_t> scfl_mask[3];
  WMList wm[18];
  WMList wmp;
  std::vector<int32_t> scfl[3];
  IBList ib[4];
  std::vector<int32_t> null_ijk[3];
  std::vector<int32_t> src_ijk, src_sc;
  std::vector<double> src_flux;
  long lin(int i, int j, int k) const { return (long)(i + 2) + (long)(j + 2) * sy + (long)(k + 2) * sz; }
};

Grid1D make_grid(const elmm_config &c);
void component_extents(const elmm_config &c, int n[3][3]);
std::string validate_config(const elmm_config &c);
void remove_duplicate_wm_points(const elmm_config &c, WMList &L);
// device: the solid-body masks and lists on the current HIP device
// (build_solids_device) instead of the host (build_solids); the same result.
int build_case(const elmm_config &c, int ncub, const elmm_cuboid *cub, int nsrc, const elmm_point_source *src,
               CaseData &cd, bool device = false);
void view_case(const CaseData &cd, elmm_lists &v);
// Estimated device bytes of a case's surface lists (elmm_create's copies).
int64_t case_list_bytes(const elmm_case *c);
// Solid bodies (axis-aligned cuboids): masks, IB lists, solid-body WM points,
// nullified points.  Defined in solids.hip.
int build_solids(const elmm_config &c, const Grid1D &g, int ncub, const elmm_cuboid *cub, CaseData &cd);
// The same on the current HIP device, and the scalar face masks and lists
// when the wall model is on (defined in k_solids.hip).
int build_solids_device(const elmm_config &c, const Grid1D &g, int ncub, const elmm_cuboid *cub, CaseData &cd);
// Runs f(0) .. f(n-1) on n threads (f(0) on the caller's) and joins them; an
// exception in a task is rethrown after every task has finished.
void run_tasks(int n, const std::function<void(int)> &f);
// The calling thread's current HIP device, and making d current on it (the
// server's worker threads run on the device the server was created on).
// Defined in engine.hip.
int current_device();
void use_device(int d);

}  // namespace elmm
