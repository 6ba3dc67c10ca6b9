// solids.cpp - solid bodies: type masks, immersed-boundary interpolation
// lists, solid-body wall-model points and nullified points, for axis-aligned
// cuboids.
//
// A cuboid is the reference's convex polyhedron of six half-spaces (the
// POLYHEDRON form of the obstacle file, src/geometricshapes.f90:2655-2760),
// with outward normals in the side order W, E, S, N, B, T and gl = .false.
// (inside where a*x+b*y+c*z+d <= eps*|n|).  The builders restate, in the
// reference's loop and list order:
//   FindInsideCells            src/solidbodies.f90:75-206
//   SolidBody_Init             src/solidbodies.f90:439-454 (rough = z0 > 0)
//   Body%Inside (periodic wrap), ClosestOnLineOut
//                              src/geometricshapes.f90:2917-3008
//   Plane Inside/Closest       src/geometricshapes.f90:543-660
//   ConvexPolyhedron Inside/Closest/ClosestOut
//                              src/geometricshapes.f90:702-966
//   Line Closest               src/geometricshapes.f90:441-455
//   FindNeighbouringCells      src/immersedboundary.f90:1475-1506
//   InitImBoundaries           src/immersedboundary.f90:1650-1720
//   TVelIBPoint_Create         src/immersedboundary.f90:690-904
//   InterpolationPoints        src/immersedboundary.f90:907-1020
//   IB_interpolation_coefs_1st_order  src/immersedboundary.f90:1047-1062
//   TScalFlIBPoint_Create      src/immersedboundary.f90:1065-1410
//   GetSolidBodiesWM           src/immersedboundary.f90:14-139
//   GetSolidBodiesWM_UVW       src/immersedboundary.f90:144-308
//   SetNullifiedPoints         src/initial.f90:4861-4962
// Differences: the inside-cell search visits only the index ranges whose
// coordinates can satisfy a cuboid's per-axis half-space pair (each plane
// of an axis-aligned cuboid depends on one coordinate, so the result is the
// same as testing every point); a singular 3x3 system in Closest, where the
// reference would continue with unset values, is reported as an error;
// norm2 is sqrt(x*x+y*y+z*z) (used only in comparisons and thresholds).
// Builders that fill different arrays or lists (the four type arrays, the
// four IB lists, the WM lists of Pr and of each velocity component, the
// nullified points of each component) run on their own threads; each list
// is still filled in the reference's order, so the result does not depend
// on the threads.
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

namespace elmm {

namespace {

struct Plane {
  double a, b, c, d;
};

struct Body {
  int id;  // numofbody, 1-based
  double z0;
  Plane pl[6];
};

// Plane_Init_3r: point and outward normal (src/geometricshapes.f90:543-554)
Plane make_plane(double xc, double yc, double zc, double a, double b, double c) {
  Plane p;
  p.a = a;
  p.b = b;
  p.c = c;
  p.d = -(a * xc + b * yc + c * zc);
  return p;
}

Body make_body(int id, const elmm_cuboid &q) {
  Body B;
  B.id = id;
  B.z0 = q.z0;
  B.pl[0] = make_plane(q.xmin, 0, 0, -1, 0, 0);
  B.pl[1] = make_plane(q.xmax, 0, 0, 1, 0, 0);
  B.pl[2] = make_plane(0, q.ymin, 0, 0, -1, 0);
  B.pl[3] = make_plane(0, q.ymax, 0, 0, 1, 0);
  B.pl[4] = make_plane(0, 0, q.zmin, 0, 0, -1);
  B.pl[5] = make_plane(0, 0, q.zmax, 0, 0, 1);
  return B;
}

inline double norm3(double x, double y, double z) { return std::sqrt(x * x + y * y + z * z); }

// Plane_Inside with gl = .false.
inline bool plane_inside(const Plane &p, double x, double y, double z, double eps) {
  return p.a * x + p.b * y + p.c * z + p.d <= eps * std::sqrt(p.a * p.a + p.b * p.b + p.c * p.c);
}

// Plane_Closest
inline void plane_closest(const Plane &p, double &xn, double &yn, double &zn, double x, double y, double z) {
  double t;
  if (std::fabs(p.a) > DBL_MIN || std::fabs(p.b) > DBL_MIN || std::fabs(p.c) > DBL_MIN)
    t = -(p.a * x + p.b * y + p.c * z + p.d) / std::sqrt(p.a * p.a + p.b * p.b + p.c * p.c);
  else
    t = 0;
  xn = x + p.a * t;
  yn = y + p.b * t;
  zn = z + p.c * t;
}

// ConvexPolyhedron_Inside (no wrap)
inline bool poly_inside(const Body &B, double x, double y, double z, double eps) {
  for (int i = 0; i < 6; i++)
    if (!plane_inside(B.pl[i], x, y, z, eps)) return false;
  return true;
}

struct Geometry {
  const elmm_config &c;
  const Grid1D &g;
  int n[3][3];
  double dxmin, dymin, dzmin;
  std::string err;

  Geometry(const elmm_config &cc, const Grid1D &gg) : c(cc), g(gg) {
    component_extents(c, n);
    dxmin = g.dxmin;
    dymin = g.dymin;
    dzmin = g.dzmin;
  }
  double xU(int i) const { return g.at(g.xU, i); }
  double yV(int j) const { return g.at(g.yV, j); }
  double zW(int k) const { return g.at(g.zW, k); }
  double xPr(int i) const { return g.at(g.xPr, i); }
  double yPr(int j) const { return g.at(g.yPr, j); }
  double zPr(int k) const { return g.at(g.zPr, k); }

  // Body%Inside = CInside: periodic wrap, then the shape's InsideEps
  bool inside(const Body &B, double x, double y, double z, double eps) const {
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    double x2 = x, y2 = y, z2 = z;
    const double lx = xU(nx) - xU(0), ly = yV(ny) - yV(0), lz = zW(nz) - zW(0);
    if (c.btype[ELMM_EA] == ELMM_BC_PERIODIC && x2 > xU(nx + 1)) x2 = x2 - lx;
    if (c.btype[ELMM_NO] == ELMM_BC_PERIODIC && y2 > yV(ny + 1)) y2 = y2 - ly;
    if (c.btype[ELMM_TO] == ELMM_BC_PERIODIC && z2 > zW(nz + 1)) z2 = z2 - lz;
    if (c.btype[ELMM_WE] == ELMM_BC_PERIODIC && x2 < xU(0)) x2 = x2 + lx;
    if (c.btype[ELMM_SO] == ELMM_BC_PERIODIC && y2 < yV(0)) y2 = y2 + ly;
    if (c.btype[ELMM_BO] == ELMM_BC_PERIODIC && z2 < zW(0)) z2 = z2 + lz;
    return poly_inside(B, x2, y2, z2, eps);
  }

  // ConvexPolyhedron_ClosestOut: nearest plane (first of equals)
  void closest_out(const Body &B, double &xn, double &yn, double &zn, double x, double y, double z) const {
    double xP[6], yP[6], zP[6], dists[6];
    for (int i = 0; i < 6; i++) {
      plane_closest(B.pl[i], xP[i], yP[i], zP[i], x, y, z);
      dists[i] = norm3(x - xP[i], y - yP[i], z - zP[i]);
    }
    int in = 0;
    double minv = DBL_MAX;
    for (int i = 0; i < 6; i++)
      if (dists[i] >= 0 && dists[i] < minv) {
        in = i;
        minv = dists[i];
      }
    xn = xP[in];
    yn = yP[in];
    zn = zP[in];
  }

  // solve3x3 by Cramer's rule (inside ConvexPolyhedron_Closest); a[r][c]
  bool solve3(const double a[3][3], const double b[3], double x[3]) const {
    double ai[3][3];
    ai[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    ai[0][1] = a[2][1] * a[0][2] - a[2][2] * a[0][1];
    ai[0][2] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    ai[1][0] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    ai[1][1] = a[2][2] * a[0][0] - a[2][0] * a[0][2];
    ai[1][2] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    ai[2][0] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    ai[2][1] = a[2][0] * a[0][1] - a[2][1] * a[0][0];
    ai[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    double det = a[0][0] * ai[0][0] + a[0][1] * ai[1][0] + a[0][2] * ai[2][0];
    if (std::fabs(det) < 1e-8) return false;
    det = 1 / det;
    x[0] = det * (ai[0][0] * b[0] + ai[0][1] * b[1] + ai[0][2] * b[2]);
    x[1] = det * (ai[1][0] * b[0] + ai[1][1] * b[1] + ai[1][2] * b[2]);
    x[2] = det * (ai[2][0] * b[0] + ai[2][1] * b[1] + ai[2][2] * b[2]);
    return true;
  }

  // ConvexPolyhedron_Closest: nearest point of the body to an outside point
  bool closest(const Body &B, double &xn, double &yn, double &zn, double x, double y, double z) {
    double xP[6], yP[6], zP[6], dists[6];
    const double eps_def = 10 * DBL_EPSILON;
    for (int i = 0; i < 6; i++) {
      plane_closest(B.pl[i], xP[i], yP[i], zP[i], x, y, z);
      dists[i] = norm3(x - xP[i], y - yP[i], z - zP[i]);
      if (plane_inside(B.pl[i], x, y, z, eps_def)) dists[i] = -std::fabs(dists[i]);
    }
    int in1 = -1;
    double minv = DBL_MAX;
    for (int i = 0; i < 6; i++)
      if (dists[i] >= DBL_MIN && dists[i] < minv) {
        in1 = i;
        minv = dists[i];
      }
    if (in1 < 0) {
      err = "solid body: no nearest face for a point outside the body";
      return false;
    }
    if (poly_inside(B, xP[in1], yP[in1], zP[in1], std::min(dxmin / 10000., std::min(dymin / 10000., dzmin / 10000.)))) {
      xn = xP[in1];
      yn = yP[in1];
      zn = zP[in1];
      return true;
    }
    for (int i = 0; i < 6; i++) dists[i] = std::fabs(dists[i]);
    int in2 = -1, in3 = -1;
    minv = DBL_MAX;
    for (int i = 0; i < 6; i++)
      if (i != in1 && dists[i] < minv) {
        in2 = i;
        minv = dists[i];
      }
    minv = DBL_MAX;
    for (int i = 0; i < 6; i++)
      if (i != in1 && i != in2 && dists[i] < minv) {
        in3 = i;
        minv = dists[i];
      }
    const Plane &p1 = B.pl[in1], &p2 = B.pl[in2], &p3 = B.pl[in3];
    double al = p1.b * p2.c - p1.c * p2.b;
    double bl = p1.c * p2.a - p1.a * p2.c;
    double cl = p1.a * p2.b - p1.b * p2.a;
    if (std::fabs(al) <= DBL_EPSILON && std::fabs(bl) <= DBL_EPSILON && std::fabs(cl) <= DBL_EPSILON) {
      // parallel planes: the reference keeps the projection on the second
      // (its comparison of the two candidates is always false)
      plane_closest(p2, xn, yn, zn, x, y, z);
      return true;
    }
    double p = std::sqrt(al * al + bl * bl + cl * cl);
    al = al / p;
    bl = bl / p;
    cl = cl / p;
    double ag[3][3], bg[3], xg[3];
    if (std::fabs(cl) >= 0.1) {
      double A[3][3] = {{p1.a, p1.b, p1.c}, {p2.a, p2.b, p2.c}, {0, 0, 1}};
      double Bv[3] = {-p1.d, -p2.d, (zW(n[2][2] + 1) + zW(0)) / 2.0};
      std::copy(&A[0][0], &A[0][0] + 9, &ag[0][0]);
      std::copy(Bv, Bv + 3, bg);
    } else if (std::fabs(bl) >= 0.1) {
      double A[3][3] = {{p1.a, p1.b, p1.c}, {p2.a, p2.b, p2.c}, {0, 1, 0}};
      double Bv[3] = {-p1.d, -p2.d, (yV(n[1][1] + 1) + yV(0)) / 2.0};
      std::copy(&A[0][0], &A[0][0] + 9, &ag[0][0]);
      std::copy(Bv, Bv + 3, bg);
    } else {
      double A[3][3] = {{p1.a, p1.b, p1.c}, {p2.a, p2.b, p2.c}, {1, 0, 0}};
      double Bv[3] = {-p1.d, -p2.d, (xU(n[0][0] + 1) + xU(0)) / 2.0};
      std::copy(&A[0][0], &A[0][0] + 9, &ag[0][0]);
      std::copy(Bv, Bv + 3, bg);
    }
    if (!solve3(ag, bg, xg)) {
      err = "solid body: singular edge system in the nearest-point search";
      return false;
    }
    // Line_Closest on the intersection line of the two nearest planes
    double t = 0;
    if (al != 0 || bl != 0 || cl != 0)
      t = (al * (x - xg[0]) + bl * (y - xg[1]) + cl * (z - xg[2])) / (al * al + bl * bl + cl * cl);
    const double xl = xg[0] + al * t, yl = xg[1] + bl * t, zl = xg[2] + cl * t;
    if (poly_inside(B, xl, yl, zl, std::min(dxmin / 1000., std::min(dymin / 10000., dzmin / 10000.)))) {
      xn = xl;
      yn = yl;
      zn = zl;
      return true;
    }
    double A3[3][3] = {{p1.a, p1.b, p1.c}, {p2.a, p2.b, p2.c}, {p3.a, p3.b, p3.c}};
    double B3[3] = {-p1.d, -p2.d, -p3.d};
    if (!solve3(A3, B3, xg)) {
      err = "solid body: singular vertex system in the nearest-point search";
      return false;
    }
    xn = xg[0];
    yn = xg[1];
    zn = xg[2];
    return true;
  }

  // Body%ClosestOnLineOut: t with x+(x2-x)*t on the boundary, x inside
  double closest_on_line_out(const Body &B, double x, double y, double z, double x2, double y2, double z2) const {
    double t1 = 0, t2 = 1, t;
    if (inside(B, x2, y2, z2, 0.0)) {
      t1 = 1;
      for (;;) {
        t2 = t2 * 1.1;
        if (!inside(B, x + (x2 - x) * t2, y + (y2 - y) * t2, z + (z2 - z) * t2, 0.0)) break;
        if (t2 >= 2) return 10;
      }
    }
    t = (t1 + t2) / 2.0;
    const double tol = std::min(dxmin / 1000., std::min(dymin / 1000., dzmin / 1000.));
    for (int i = 1; i <= 20; i++) {
      if (inside(B, x + (x2 - x) * t, y + (y2 - y) * t, z + (z2 - z) * t, 0.0))
        t1 = t;
      else
        t2 = t;
      t = (t1 + t2) / 2.0;
      if (std::fabs(t1 - t2) < tol) break;
    }
    return t;
  }
};

// index ranges where one axis' half-space pair can hold (exact per-axis test)
std::vector<char> axis_hits(const Geometry &G, const Body &B, int axis, const std::vector<double> &coord, int lo,
                            int hi, double eps) {
  std::vector<char> hit(hi - lo + 1, 0);
  const Plane &pm = B.pl[2 * axis], &pp = B.pl[2 * axis + 1];
  const int nn = axis == 0 ? G.c.nx : axis == 1 ? G.c.ny : G.c.nz;
  const std::vector<double> &face = axis == 0 ? G.g.xU : axis == 1 ? G.g.yV : G.g.zW;
  const int sm = axis == 0 ? ELMM_WE : axis == 1 ? ELMM_SO : ELMM_BO;
  const int sp = sm + 1;
  const double L = Grid1D::at(face, nn) - Grid1D::at(face, 0);
  for (int i = lo; i <= hi; i++) {
    double v = Grid1D::at(coord, i);
    if (G.c.btype[sp] == ELMM_BC_PERIODIC && v > Grid1D::at(face, nn + 1)) v = v - L;
    if (G.c.btype[sm] == ELMM_BC_PERIODIC && v < Grid1D::at(face, 0)) v = v + L;
    double q[3] = {0, 0, 0};
    q[axis] = v;
    hit[i - lo] = plane_inside(pm, q[0], q[1], q[2], eps) && plane_inside(pp, q[0], q[1], q[2], eps);
  }
  return hit;
}

// FindInsideCells for one type array on the (x, y, z) coordinate arrays
void set_type(const Geometry &G, const Body &B, CaseData &cd, std::vector<int32_t> &T, const std::vector<double> &X,
              const std::vector<double> &Y, const std::vector<double> &Z, const int ext[3], double delta) {
  auto hx = axis_hits(G, B, 0, X, -2, ext[0] + 3, delta);
  auto hy = axis_hits(G, B, 1, Y, -2, ext[1] + 3, delta);
  auto hz = axis_hits(G, B, 2, Z, -2, ext[2] + 3, delta);
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

struct Tap {
  int i, j, k;
  double coef;
};

void push_ib(IBList &L, int i, int j, int k, const std::vector<Tap> &taps) {
  L.ijk.insert(L.ijk.end(), {i, j, k});
  for (const Tap &t : taps) {
    L.src.insert(L.src.end(), {t.i, t.j, t.k});
    L.coef.push_back(t.coef);
  }
  L.ptr.push_back((int32_t)L.coef.size());
}

// IB_interpolation_coefs_1st_order with n_ls = 2
void ib_coefs(double xr, const double x[3], double out[2]) {
  double A = (x[1] - xr) * (x[1] - xr) + (x[2] - xr) * (x[2] - xr);
  out[0] = (x[1] - xr) * (x[0] - xr) / A;
  out[1] = (x[2] - xr) * (x[0] - xr) / A;
}

// TVelIBPoint_Create + InterpolationPoints for one velocity point
std::vector<Tap> vel_ib_point(Geometry &G, const std::vector<Body> &bodies, const CaseData &cd,
                              const std::vector<int32_t> &T, int xi, int yj, int zk, const std::vector<double> &XU,
                              const std::vector<double> &YU, const std::vector<double> &ZU) {
  std::vector<Tap> none;
  auto X = [&](int i) { return Grid1D::at(XU, i); };
  auto Y = [&](int j) { return Grid1D::at(YU, j); };
  auto Z = [&](int k) { return Grid1D::at(ZU, k); };
  auto t = [&](int i, int j, int k) { return T[cd.lin(i, j, k)]; };
  const double x = X(xi), y = Y(yj), z = Z(zk);
  const Body &SB = bodies[t(xi, yj, zk) - 1];
  // (the nested-domain boundary checks of the reference apply to domain
  // boundary codes 2000-2010 only, which this build does not accept)
  if (!G.inside(SB, x, y, z, 0.0)) return none;
  double xn, yn, zn;
  G.closest_out(SB, xn, yn, zn, x, y, z);
  if (norm3(x - xn, y - yn, z - zn) <= 0.1 * std::pow(G.dxmin * G.dymin * G.dzmin, 1.0 / 3.0)) return none;
  const int nls = 2;
  auto free_run = [&](int a, int b, int c) {
    for (int s = 1; s <= nls; s++)
      if (t(xi + s * a, yj + s * b, zk + s * c) > 0) return false;
    return true;
  };
  bool freexm = free_run(-1, 0, 0), freeym = free_run(0, -1, 0), freezm = free_run(0, 0, -1);
  bool freexp = free_run(1, 0, 0), freeyp = free_run(0, 1, 0), freezp = free_run(0, 0, 1);
  bool f1xm = t(xi - 1, yj, zk) <= 0, f1ym = t(xi, yj - 1, zk) <= 0, f1zm = t(xi, yj, zk - 1) <= 0;
  bool f1xp = t(xi + 1, yj, zk) <= 0, f1yp = t(xi, yj + 1, zk) <= 0, f1zp = t(xi, yj, zk + 1) <= 0;
  if ((f1xm && !freexm) || (f1ym && !freeym) || (f1zm && !freezm) || (f1xp && !freexp) || (f1yp && !freeyp) ||
      (f1zp && !freezp))
    return none;
  int nsolid = 0;
  for (int c = -1; c <= 1; c++)
    for (int b = -1; b <= 1; b++)
      for (int a = -1; a <= 1; a++)
        if (t(xi + a, yj + b, zk + c) > 0) nsolid++;
  if (nsolid > 21) return none;
  if ((f1xm && f1xp) || (f1ym && f1yp) || (f1zm && f1zp)) return none;
  int dirx = (!freexm && !freexp) ? 0 : freexm ? -1 : 1;
  int diry = (!freeym && !freeyp) ? 0 : freeym ? -1 : 1;
  int dirz = (!freezm && !freezp) ? 0 : freezm ? -1 : 1;
  double distx = 0, disty = 0, distz = 0;
  if (dirx != 0) {
    double tx = G.closest_on_line_out(SB, x, y, z, X(xi + dirx), y, z);
    if (tx > 1 || tx < 0.05) {
      dirx = 0;
      tx = 0;
    }
    distx = tx * (X(xi + dirx) - X(xi));
  }
  if (diry != 0) {
    double ty = G.closest_on_line_out(SB, x, y, z, x, Y(yj + diry), z);
    if (ty > 1 || ty < 0.05) {
      diry = 0;
      ty = 0;
    }
    disty = ty * (Y(yj + diry) - Y(yj));
  }
  if (dirz != 0) {
    double tz = G.closest_on_line_out(SB, x, y, z, x, y, Z(zk + dirz));
    if (tz > 1 || tz < 0.05) {
      dirz = 0;
      tz = 0;
    }
    distz = tz * (Z(zk + dirz) - Z(zk));
  }
  if (dirx == 0 && diry == 0 && dirz == 0) return none;
  // InterpolationPoints
  Tap tmp[6];
  double d[3] = {1, 1, 1};
  if (dirx != 0) {
    double xr = X(xi) + distx, xs[3] = {X(xi), X(xi + dirx), X(xi + 2 * dirx)}, cf[2];
    ib_coefs(xr, xs, cf);
    for (int s = 0; s < 2; s++) tmp[s] = {xi + (s + 1) * dirx, yj, zk, cf[s]};
    d[0] = std::fabs(xs[0] - xr);
  }
  if (diry != 0) {
    double yr = Y(yj) + disty, ys[3] = {Y(yj), Y(yj + diry), Y(yj + 2 * diry)}, cf[2];
    ib_coefs(yr, ys, cf);
    for (int s = 0; s < 2; s++) tmp[2 + s] = {xi, yj + (s + 1) * diry, zk, cf[s]};
    d[1] = std::fabs(ys[0] - yr);
  }
  if (dirz != 0) {
    double zr = Z(zk) + distz, zs[3] = {Z(zk), Z(zk + dirz), Z(zk + 2 * dirz)}, cf[2];
    ib_coefs(zr, zs, cf);
    for (int s = 0; s < 2; s++) tmp[4 + s] = {xi, yj, zk + (s + 1) * dirz, cf[s]};
    d[2] = std::fabs(zs[0] - zr);
  }
  // Peller et al. weights b_d = prod(d)/d_d^2
  const double prod = d[0] * d[1] * d[2];
  double b[3] = {0, 0, 0};
  if (dirx != 0) b[0] = prod / (d[0] * d[0]);
  if (diry != 0) b[1] = prod / (d[1] * d[1]);
  if (dirz != 0) b[2] = prod / (d[2] * d[2]);
  const double csum = b[0] + b[1] + b[2];
  std::vector<Tap> res;
  const int dirs[3] = {dirx, diry, dirz};
  for (int a = 0; a < 3; a++)
    if (dirs[a] != 0)
      for (int s = 0; s < 2; s++) {
        Tap tp = tmp[2 * a + s];
        tp.coef = tp.coef * b[a] / csum;
        res.push_back(tp);
      }
  return res;
}

inline int fsign_nint(double v) { return std::signbit(v) ? -1 : 1; }

// TScalFlIBPoint_Create for one Pr cell
std::vector<Tap> scalfl_ib_point(Geometry &G, const std::vector<Body> &bodies, const CaseData &cd, int xi, int yj,
                                 int zk) {
  auto pt = [&](int i, int j, int k) { return cd.prtype[cd.lin(i, j, k)]; };
  const double x = G.xPr(xi), y = G.yPr(yj), z = G.zPr(zk);
  const Body &SB = bodies[pt(xi, yj, zk) - 1];
  double xn, yn, zn;
  G.closest_out(SB, xn, yn, zn, x, y, z);
  double distx = xn - x, disty = yn - y, distz = zn - z;
  int dirx = fsign_nint(distx), diry = fsign_nint(disty), dirz = fsign_nint(distz);
  if (std::fabs(distx) < (G.xPr(xi + 1) - G.xPr(xi - 1)) / 100.) {
    distx = 0;
    dirx = 0;
  }
  if (std::fabs(disty) < (G.yPr(yj + 1) - G.yPr(yj - 1)) / 100.) {
    disty = 0;
    diry = 0;
  }
  if (std::fabs(distz) < (G.zW(zk + 1) - G.zW(zk - 1)) / 100.) {
    distz = 0;
    dirz = 0;
  }
  const bool fp00 = pt(xi + 1, yj, zk) <= 0, fm00 = pt(xi - 1, yj, zk) <= 0;
  const bool f0p0 = pt(xi, yj + 1, zk) <= 0, f0m0 = pt(xi, yj - 1, zk) <= 0;
  const bool f00p = pt(xi, yj, zk + 1) <= 0, f00m = pt(xi, yj, zk - 1) <= 0;
  if (!(fp00 || fm00)) {
    dirx = 0;
    distx = 0;
  }
  if (!(f0p0 || f0m0)) {
    diry = 0;
    disty = 0;
  }
  if (!(f00p || f00m)) {
    dirz = 0;
    distz = 0;
  }
  const int nfree = fp00 + f0p0 + f00p + fm00 + f0m0 + f00m;
  const int ndirs = (dirx != 0) + (diry != 0) + (dirz != 0);
  if (nfree > ndirs || (fp00 && fm00) || (f0p0 && f0m0) || (f00p && f00m)) {
    distx = disty = distz = 0;
    dirx = diry = dirz = 0;
  }
  std::vector<Tap> r;
  if (dirx == 0 && diry == 0 && dirz == 0) {
    int dx2 = 0, dy2 = 0, dz2 = 0;
    if (fp00) dx2 += 1;
    if (fm00) dx2 -= 1;
    if (f0p0) dy2 += 1;
    if (f0m0) dy2 -= 1;
    if (f00p) dz2 += 1;
    if (f00m) dz2 -= 1;
    if (dx2 == 0 && dy2 == 0 && dz2 == 0) {
      const double w = 1.0 / nfree;
      if (fp00) r.push_back({xi + 1, yj, zk, w});
      if (fm00) r.push_back({xi - 1, yj, zk, w});
      if (f0p0) r.push_back({xi, yj + 1, zk, w});
      if (f0m0) r.push_back({xi, yj - 1, zk, w});
      if (f00p) r.push_back({xi, yj, zk + 1, w});
      if (f00m) r.push_back({xi, yj, zk - 1, w});
    } else {
      r.push_back({xi + dx2, yj + dy2, zk + dz2, 1.0});
    }
  } else if (nfree == 1 || ndirs == 1) {
    r.push_back({xi + dirx, yj + diry, zk + dirz, 1.0});
  } else if (nfree == 2 || ndirs == 2) {
    double c1;
    if (dirx == 0) {  // plane yz
      if (std::fabs(disty / distz) < std::fabs(G.yPr(yj + diry) - y) / std::fabs(G.zPr(zk + dirz) - z)) {
        double t = (G.zPr(zk + dirz) - z) / distz;
        c1 = std::fabs(disty * t) / std::fabs(G.yPr(yj + diry) - G.yPr(yj));
        r.push_back({xi, yj + diry, zk + dirz, c1});
        r.push_back({xi, yj, zk + dirz, 1 - c1});
      } else {
        double t = (G.yPr(yj + diry) - y) / disty;
        c1 = std::fabs(distz * t) / std::fabs(G.zPr(zk + dirz) - G.zPr(zk));
        r.push_back({xi, yj + diry, zk + dirz, c1});
        r.push_back({xi, yj + diry, zk, 1 - c1});
      }
    } else if (diry == 0) {  // plane xz
      if (std::fabs(distx / distz) < std::fabs(G.xPr(xi + dirx) - x) / std::fabs(G.zPr(zk + dirz) - z)) {
        double t = (G.zPr(zk + dirz) - z) / distz;
        c1 = std::fabs(distx * t) / std::fabs(G.xPr(xi + dirx) - G.xPr(xi));
        r.push_back({xi + dirx, yj, zk + dirz, c1});
        r.push_back({xi, yj, zk + dirz, 1 - c1});
      } else {
        double t = (G.xPr(xi + dirx) - x) / distx;
        c1 = std::fabs(distz * t) / std::fabs(G.zPr(zk + dirz) - G.zPr(zk));
        r.push_back({xi + dirx, yj, zk + dirz, c1});
        r.push_back({xi + dirx, yj, zk, 1 - c1});
      }
    } else {  // plane xy
      if (std::fabs(distx / disty) < std::fabs(G.xPr(xi + dirx) - x) / std::fabs(G.yPr(yj + diry) - y)) {
        double t = (G.yPr(yj + diry) - y) / disty;
        c1 = std::fabs(distx * t) / std::fabs(G.xPr(xi + dirx) - G.xPr(xi));
        r.push_back({xi + dirx, yj + diry, zk, c1});
        r.push_back({xi, yj + diry, zk, 1 - c1});
      } else {
        double t = (G.xPr(xi + dirx) - x) / distx;
        c1 = std::fabs(disty * t) / std::fabs(G.yPr(yj + diry) - G.yPr(yj));
        r.push_back({xi + dirx, yj + diry, zk, c1});
        r.push_back({xi + dirx, yj, zk, 1 - c1});
      }
    }
  } else {
    const double tx = (G.xPr(xi + dirx) - x) / distx;
    const double ty = (G.yPr(yj + diry) - y) / disty;
    const double tz = (G.zPr(zk + dirz) - z) / distz;
    if (tx <= ty && tx <= tz) {
      const double fy = std::fabs(disty * tx) / std::fabs(G.yPr(yj + diry) - y);
      const double fz = std::fabs(distz * tx) / std::fabs(G.zPr(zk + dirz) - z);
      r.push_back({xi + dirx, yj + diry, zk + dirz, fy * fz});
      r.push_back({xi + dirx, yj, zk + dirz, (1 - fy) * fz});
      r.push_back({xi + dirx, yj + diry, zk, fy * (1 - fz)});
      r.push_back({xi + dirx, yj, zk, (1 - fy) * (1 - fz)});
    } else if (ty <= tx && ty <= tz) {
      const double fx = std::fabs(distx * ty) / std::fabs(G.xPr(xi + dirx) - x);
      const double fz = std::fabs(distz * ty) / std::fabs(G.zPr(zk + dirz) - z);
      r.push_back({xi + dirx, yj + diry, zk + dirz, fx * fz});
      r.push_back({xi, yj + diry, zk + dirz, (1 - fx) * fz});
      r.push_back({xi + dirx, yj + diry, zk, fx * (1 - fz)});
      r.push_back({xi, yj + diry, zk, (1 - fx) * (1 - fz)});
    } else {
      const double fx = std::fabs(distx * tz) / std::fabs(G.xPr(xi + dirx) - x);
      const double fy = std::fabs(disty * tz) / std::fabs(G.yPr(yj + diry) - y);
      r.push_back({xi + dirx, yj + diry, zk + dirz, fx * fy});
      r.push_back({xi, yj + diry, zk + dirz, (1 - fx) * fy});
      r.push_back({xi + dirx, yj, zk + dirz, fx * (1 - fy)});
      r.push_back({xi, yj, zk + dirz, (1 - fx) * (1 - fy)});
    }
  }
  return r;
}

bool has_free_face_neighbour(const CaseData &cd, const std::vector<int32_t> &T, int i, int j, int k) {
  return T[cd.lin(i + 1, j, k)] <= 0 || T[cd.lin(i - 1, j, k)] <= 0 || T[cd.lin(i, j + 1, k)] <= 0 ||
         T[cd.lin(i, j - 1, k)] <= 0 || T[cd.lin(i, j, k + 1)] <= 0 || T[cd.lin(i, j, k - 1)] <= 0;
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

// GetSolidBodiesWM: free Pr points with a solid face neighbour (only cells
// marked -1 by FindNeighbouringCells can have one).  Returns an error
// message, empty on success.
std::string solid_wm_pr(Geometry &G, const std::vector<Body> &bodies, CaseData &cd, double min_wall_distance) {
  const elmm_config &c = G.c;
  static const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  for (int k = 1; k <= c.nz; k++)
    for (int j = 1; j <= c.ny; j++)
      for (int i = 1; i <= c.nx; i++) {
        if (cd.prtype[cd.lin(i, j, k)] >= 0) continue;
        double dist = DBL_MAX, nx_ = 0, ny_ = 0, nz_ = 0;
        int body = 0;
        for (int r = 0; r < 6; r++) {
          const int t = cd.prtype[cd.lin(i + nb[r][0], j + nb[r][1], k + nb[r][2])];
          if (t > 0 && t != body) {
            if (!G.closest(bodies[t - 1], nx_, ny_, nz_, G.xPr(i), G.yPr(j), G.zPr(k))) return G.err;
            const double dd = norm3(nx_ - G.xPr(i), ny_ - G.yPr(j), nz_ - G.zPr(k));
            if (dd < dist) {
              dist = dd;
              body = t;
            }
          }
        }
        if (body == 0) continue;
        const double dx = nx_ - G.xPr(i), dy = ny_ - G.yPr(j), dz = nz_ - G.zPr(k);
        const double d = norm3(dx, dy, dz);
        if (d == 0) return "zero distance wall-model point next to a solid body";
        if (d < min_wall_distance) return "wall-model point too close to a solid body";
        const double z0 = bodies[body - 1].z0 > 0 ? bodies[body - 1].z0 : 0.0;
        cd.wmp.ijk.insert(cd.wmp.ijk.end(), {i, j, k});
        cd.wmp.dist.insert(cd.wmp.dist.end(), {dx, dy, dz});
        cd.wmp.z0.push_back(z0);
      }
  return std::string();
}

// GetSolidBodiesWM_UVW for one velocity component: -1 points, per direction.
std::string solid_wm_uvw(Geometry &G, const std::vector<Body> &bodies, CaseData &cd, int comp,
                         const std::vector<int32_t> &T, const std::vector<double> &cx, const std::vector<double> &cy,
                         const std::vector<double> &cz, double min_wall_distance) {
  static const int dvec[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  auto X = [&](int i) { return Grid1D::at(cx, i); };
  auto Y = [&](int j) { return Grid1D::at(cy, j); };
  auto Z = [&](int k) { return Grid1D::at(cz, k); };
  for (int k = 1; k <= G.n[comp][2]; k++)
    for (int j = 1; j <= G.n[comp][1]; j++)
      for (int i = 1; i <= G.n[comp][0]; i++) {
        if (T[cd.lin(i, j, k)] >= 0) continue;
        for (int dir = 0; dir < 6; dir++) {
          const int m = dvec[dir][0], n = dvec[dir][1], o = dvec[dir][2];
          const int t = T[cd.lin(i + m, j + n, k + o)];
          if (t <= 0) continue;
          const Body &SB = bodies[t - 1];
          double nx_, ny_, nz_;
          if (!G.closest(SB, nx_, ny_, nz_, X(i), Y(j), Z(k))) return G.err;
          double dv[3] = {nx_ - X(i), ny_ - Y(j), nz_ - Z(k)};
          const double dist = norm3(dv[0], dv[1], dv[2]);
          const double cosang =
              (dv[0] * m + dv[1] * n + dv[2] * o) / (dist * norm3((double)m, (double)n, (double)o));
          if (!(cosang > 0.4)) {
            const double tl = G.closest_on_line_out(SB, X(i + m), Y(j + n), Z(k + o), X(i), Y(j), Z(k));
            nx_ = X(i + m) + tl * (X(i) - X(i + m));
            ny_ = Y(j + n) + tl * (Y(j) - Y(j + n));
            nz_ = Z(k + o) + tl * (Z(k) - Z(k + o));
            dv[0] = nx_ - X(i);
            dv[1] = ny_ - Y(j);
            dv[2] = nz_ - Z(k);
            const double dl = norm3(dv[0], dv[1], dv[2]);
            if (dl < dist && dl < min_wall_distance) {
              if (!G.closest(SB, nx_, ny_, nz_, X(i), Y(j), Z(k))) return G.err;
              dv[0] = nx_ - X(i);
              dv[1] = ny_ - Y(j);
              dv[2] = nz_ - Z(k);
            }
          }
          if (norm3(dv[0], dv[1], dv[2]) < min_wall_distance)
            return "velocity wall-model point too close to a solid body";
          WMList &L = cd.wm[comp * 6 + dir];
          L.ijk.insert(L.ijk.end(), {i, j, k});
          L.dist.insert(L.dist.end(), {dv[0], dv[1], dv[2]});
          L.z0.push_back(SB.z0 > 0 ? SB.z0 : 0.0);
        }
      }
  return std::string();
}

int build_solids(const elmm_config &c, const Grid1D &g, int ncub, const elmm_cuboid *cub, CaseData &cd) {
  Geometry G(c, g);
  std::vector<Body> bodies;
  for (int b = 0; b < ncub; b++) {
    const elmm_cuboid &q = cub[b];
    if (!(q.xmin < q.xmax && q.ymin < q.ymax && q.zmin < q.zmax) || !(q.z0 >= 0))
      return fail(ELMM_E_ARG, "cuboid " + std::to_string(b) + ": empty extent or negative z0");
    bodies.push_back(make_body(b + 1, q));
  }
  const double delta = std::pow(G.dxmin * G.dymin * G.dzmin, 1.0 / 3.0) / 20;
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
    for (const Body &B : bodies) set_type(G, B, cd, *T[a], *cx[a], *cy[a], *cz[a], ext[a], delta);
    aux_neighbours(cd, *T[a], ext[a]);
  });
  // InitImBoundaries (serial k, j, i order per list); lists 0-2 velocity,
  // 3 scalar flux
  run_tasks(4, [&](int a) {
    Geometry Gt(c, g);
    IBList &L = cd.ib[a];
    L = IBList();
    if (a < 3) {
      const std::vector<int32_t> &Tc = *T[a + 1];
      for (int k = 1; k <= G.n[a][2]; k++)
        for (int j = 1; j <= G.n[a][1]; j++)
          for (int i = 1; i <= G.n[a][0]; i++)
            if (Tc[cd.lin(i, j, k)] > 0 && has_free_face_neighbour(cd, Tc, i, j, k))
              push_ib(L, i, j, k, vel_ib_point(Gt, bodies, cd, Tc, i, j, k, *cx[a + 1], *cy[a + 1], *cz[a + 1]));
    } else {
      for (int k = 1; k <= c.nz; k++)
        for (int j = 1; j <= c.ny; j++)
          for (int i = 1; i <= c.nx; i++)
            if (cd.prtype[cd.lin(i, j, k)] > 0 && has_free_face_neighbour(cd, cd.prtype, i, j, k))
              push_ib(L, i, j, k, scalfl_ib_point(Gt, bodies, cd, i, j, k));
    }
  });

  // wall-model points (task 0: Pr, tasks 1-3: U, V, W) and nullified points
  // (tasks 1-3); a task that finds an error stops and keeps its message, and
  // the first one in task order is reported (the one the serial order meets
  // first)
  std::string errs[4];
  const double min_wall_distance = delta;
  run_tasks(4, [&](int a) {
    Geometry Gt(c, g);
    if (c.wall_model && a == 0) errs[0] = solid_wm_pr(Gt, bodies, cd, min_wall_distance);
    if (a == 0) return;
    const int comp = a - 1;
    if (c.wall_model) errs[a] = solid_wm_uvw(Gt, bodies, cd, comp, *T[a], *cx[a], *cy[a], *cz[a], min_wall_distance);
    if (!errs[a].empty()) return;
    // SetNullifiedPoints: velocity points whose 7-point stencil is solid
    const std::vector<int32_t> &Tc = *T[a];
    cd.null_ijk[comp].clear();
    for (int k = 1; k <= G.n[comp][2]; k++)
      for (int j = 1; j <= G.n[comp][1]; j++)
        for (int i = 1; i <= G.n[comp][0]; i++) {
          auto sol = [&](int p, int q, int r) { return Tc[cd.lin(i + p, j + q, k + r)] > 0; };
          if (sol(0, 0, 0) && sol(0, 0, 1) && sol(0, 0, -1) && sol(0, -1, 0) && sol(0, 1, 0) && sol(-1, 0, 0) &&
              sol(1, 0, 0))
            cd.null_ijk[comp].insert(cd.null_ijk[comp].end(), {i, j, k});
        }
  });
  for (const std::string &e : errs)
    if (!e.empty()) return fail(ELMM_E_ARG, e);
  return 0;
}

}  // namespace elmm
