// solid_geom.hpp - geometry of cuboid bodies and the per-point builders of
// the solid-body lists, as host+device functions: the host case builder
// (solids.hip) and the device case builder (k_solids.hip) call the same code,
// so both produce the same bits (IEEE double operations, no contraction, no
// libm calls on either side: the one cube root is evaluated by the host and
// passed in).
//
// A cuboid is the reference's convex polyhedron of six half-spaces (the
// POLYHEDRON form of the obstacle file, src/geometricshapes.f90:2655-2760),
// outward normals in the side order W, E, S, N, B, T and gl = .false.
// (inside where a*x+b*y+c*z+d <= eps*|n|).  Restated, in the reference's
// order of operations:
//   Plane Inside/Closest       src/geometricshapes.f90:543-660
//   ConvexPolyhedron Inside/Closest/ClosestOut
//                              src/geometricshapes.f90:702-966
//   Body%Inside (periodic wrap), ClosestOnLineOut
//                              src/geometricshapes.f90:2917-3008
//   Line Closest               src/geometricshapes.f90:441-455
//   TVelIBPoint_Create         src/immersedboundary.f90:690-904
//   InterpolationPoints        src/immersedboundary.f90:907-1020
//   IB_interpolation_coefs_1st_order  src/immersedboundary.f90:1047-1062
//   TScalFlIBPoint_Create      src/immersedboundary.f90:1065-1410
//   GetSolidBodiesWM           src/immersedboundary.f90:14-139
//   GetSolidBodiesWM_UVW       src/immersedboundary.f90:144-308
//   SetNullifiedPoints         src/initial.f90:4861-4962
// Differences: a singular 3x3 system in Closest, where the reference would
// continue with unset values, is an error; norm2 is sqrt(x*x+y*y+z*z) (used
// only in comparisons and thresholds).
#pragma once
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

namespace elmm {
namespace sg {

#define SG_HD __host__ __device__ inline

struct Plane {
  double a, b, c, d;
};

struct Body {
  int id;  // numofbody, 1-based
  double z0;
  Plane pl[6];
  double tflux, qflux;  // surface heat and moisture flux (rough bodies)
};

struct Tap {
  int i, j, k;
  double coef;
};

// Error codes of the point builders (negative returns), messages in
// solid_error().
enum { E_NO_FACE = 1, E_EDGE, E_VERTEX, E_WM_ZERO, E_WM_CLOSE, E_UVW_CLOSE };

inline const char *solid_error(int code) {
  switch (code) {
    case E_NO_FACE: return "solid body: no nearest face for a point outside the body";
    case E_EDGE: return "solid body: singular edge system in the nearest-point search";
    case E_VERTEX: return "solid body: singular vertex system in the nearest-point search";
    case E_WM_ZERO: return "zero distance wall-model point next to a solid body";
    case E_WM_CLOSE: return "wall-model point too close to a solid body";
    case E_UVW_CLOSE: return "velocity wall-model point too close to a solid body";
  }
  return "solid body: unknown error";
}

// Grid and boundary description.  Coordinate arrays hold index i at element
// i + 3 (lower bound -3); type arrays use the ghosted layout with lower bound
// -2 (lin).
struct Geom {
  int nx, ny, nz;
  int n[3][3];       // component extents n[c][d]
  int per[6];        // side is periodic
  double dxmin, dymin, dzmin;
  double ib_cut;     // 0.1 (dxmin dymin dzmin)^(1/3)
  long sy, sz;
  const double *xU, *yV, *zW, *xPr, *yPr, *zPr;
  SG_HD long lin(int i, int j, int k) const { return (long)(i + 2) + (long)(j + 2) * sy + (long)(k + 2) * sz; }
  SG_HD double XU(int i) const { return xU[i + 3]; }
  SG_HD double YV(int j) const { return yV[j + 3]; }
  SG_HD double ZW(int k) const { return zW[k + 3]; }
  SG_HD double XP(int i) const { return xPr[i + 3]; }
  SG_HD double YP(int j) const { return yPr[j + 3]; }
  SG_HD double ZP(int k) const { return zPr[k + 3]; }
};

// std::min's selection: b when b < a, else a
SG_HD double dmin(double a, double b) { return (b < a) ? b : a; }
SG_HD double at3(const double *v, int i) { return v[i + 3]; }
SG_HD double norm3(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// Plane_Init_3r: point and outward normal (src/geometricshapes.f90:543-554)
SG_HD Plane make_plane(double xc, double yc, double zc, double a, double b, double c) {
  Plane p;
  p.a = a;
  p.b = b;
  p.c = c;
  p.d = -(a * xc + b * yc + c * zc);
  return p;
}

SG_HD Body make_body(int id, double xmin, double xmax, double ymin, double ymax, double zmin, double zmax,
                     double z0) {
  Body B;
  B.id = id;
  B.z0 = z0;
  B.tflux = 0;
  B.qflux = 0;
  B.pl[0] = make_plane(xmin, 0, 0, -1, 0, 0);
  B.pl[1] = make_plane(xmax, 0, 0, 1, 0, 0);
  B.pl[2] = make_plane(0, ymin, 0, 0, -1, 0);
  B.pl[3] = make_plane(0, ymax, 0, 0, 1, 0);
  B.pl[4] = make_plane(0, 0, zmin, 0, 0, -1);
  B.pl[5] = make_plane(0, 0, zmax, 0, 0, 1);
  return B;
}

// Plane_Inside with gl = .false.
SG_HD bool plane_inside(const Plane &p, double x, double y, double z, double eps) {
  return p.a * x + p.b * y + p.c * z + p.d <= eps * sqrt(p.a * p.a + p.b * p.b + p.c * p.c);
}

// Plane_Closest
SG_HD void plane_closest(const Plane &p, double &xn, double &yn, double &zn, double x, double y, double z) {
  double t;
  if (fabs(p.a) > DBL_MIN || fabs(p.b) > DBL_MIN || fabs(p.c) > DBL_MIN)
    t = -(p.a * x + p.b * y + p.c * z + p.d) / sqrt(p.a * p.a + p.b * p.b + p.c * p.c);
  else
    t = 0;
  xn = x + p.a * t;
  yn = y + p.b * t;
  zn = z + p.c * t;
}

// ConvexPolyhedron_Inside (no wrap)
SG_HD bool poly_inside(const Body &B, double x, double y, double z, double eps) {
  for (int i = 0; i < 6; i++)
    if (!plane_inside(B.pl[i], x, y, z, eps)) return false;
  return true;
}

// One axis' half-space pair of a cuboid at coordinate v of that axis, after
// the periodic wrap of Body%Inside.  A cuboid's planes each depend on one
// coordinate, so a point is inside exactly when the pair holds on all three
// axes (FindInsideCells evaluates it per axis).
SG_HD bool axis_hit(const Geom &G, const Body &B, int axis, double v, double eps) {
  const Plane &pm = B.pl[2 * axis], &pp = B.pl[2 * axis + 1];
  const int nn = axis == 0 ? G.nx : axis == 1 ? G.ny : G.nz;
  const double *face = axis == 0 ? G.xU : axis == 1 ? G.yV : G.zW;
  const double L = at3(face, nn) - at3(face, 0);
  if (G.per[2 * axis + 1] && v > at3(face, nn + 1)) v = v - L;
  if (G.per[2 * axis] && v < at3(face, 0)) v = v + L;
  double q[3] = {0, 0, 0};
  q[axis] = v;
  return plane_inside(pm, q[0], q[1], q[2], eps) && plane_inside(pp, q[0], q[1], q[2], eps);
}

// Body%Inside = CInside: periodic wrap, then the shape's InsideEps
SG_HD bool inside(const Geom &G, const Body &B, double x, double y, double z, double eps) {
  double x2 = x, y2 = y, z2 = z;
  const double lx = G.XU(G.nx) - G.XU(0), ly = G.YV(G.ny) - G.YV(0), lz = G.ZW(G.nz) - G.ZW(0);
  if (G.per[1] && x2 > G.XU(G.nx + 1)) x2 = x2 - lx;
  if (G.per[3] && y2 > G.YV(G.ny + 1)) y2 = y2 - ly;
  if (G.per[5] && z2 > G.ZW(G.nz + 1)) z2 = z2 - lz;
  if (G.per[0] && x2 < G.XU(0)) x2 = x2 + lx;
  if (G.per[2] && y2 < G.YV(0)) y2 = y2 + ly;
  if (G.per[4] && z2 < G.ZW(0)) z2 = z2 + lz;
  return poly_inside(B, x2, y2, z2, eps);
}

// ConvexPolyhedron_ClosestOut: nearest plane (first of equals)
SG_HD void closest_out(const Body &B, double &xn, double &yn, double &zn, double x, double y, double z) {
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
SG_HD bool solve3(const double a[3][3], const double b[3], double x[3]) {
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
  if (fabs(det) < 1e-8) return false;
  det = 1 / det;
  x[0] = det * (ai[0][0] * b[0] + ai[0][1] * b[1] + ai[0][2] * b[2]);
  x[1] = det * (ai[1][0] * b[0] + ai[1][1] * b[1] + ai[1][2] * b[2]);
  x[2] = det * (ai[2][0] * b[0] + ai[2][1] * b[1] + ai[2][2] * b[2]);
  return true;
}

// ConvexPolyhedron_Closest: nearest point of the body to an outside point;
// returns 0 or an error code
SG_HD int closest(const Geom &G, const Body &B, double &xn, double &yn, double &zn, double x, double y, double z) {
  double xP[6], yP[6], zP[6], dists[6];
  const double eps_def = 10 * DBL_EPSILON;
  for (int i = 0; i < 6; i++) {
    plane_closest(B.pl[i], xP[i], yP[i], zP[i], x, y, z);
    dists[i] = norm3(x - xP[i], y - yP[i], z - zP[i]);
    if (plane_inside(B.pl[i], x, y, z, eps_def)) dists[i] = -fabs(dists[i]);
  }
  int in1 = -1;
  double minv = DBL_MAX;
  for (int i = 0; i < 6; i++)
    if (dists[i] >= DBL_MIN && dists[i] < minv) {
      in1 = i;
      minv = dists[i];
    }
  if (in1 < 0) return E_NO_FACE;
  if (poly_inside(B, xP[in1], yP[in1], zP[in1], dmin(G.dxmin / 10000., dmin(G.dymin / 10000., G.dzmin / 10000.)))) {
    xn = xP[in1];
    yn = yP[in1];
    zn = zP[in1];
    return 0;
  }
  for (int i = 0; i < 6; i++) dists[i] = fabs(dists[i]);
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
  if (fabs(al) <= DBL_EPSILON && fabs(bl) <= DBL_EPSILON && fabs(cl) <= DBL_EPSILON) {
    // parallel planes: the reference keeps the projection on the second
    // (its comparison of the two candidates is always false)
    plane_closest(p2, xn, yn, zn, x, y, z);
    return 0;
  }
  const double p = sqrt(al * al + bl * bl + cl * cl);
  al = al / p;
  bl = bl / p;
  cl = cl / p;
  double ag[3][3] = {{p1.a, p1.b, p1.c}, {p2.a, p2.b, p2.c}, {0, 0, 0}};
  double bg[3] = {-p1.d, -p2.d, 0}, xg[3];
  if (fabs(cl) >= 0.1) {
    ag[2][2] = 1;
    bg[2] = (G.ZW(G.n[2][2] + 1) + G.ZW(0)) / 2.0;
  } else if (fabs(bl) >= 0.1) {
    ag[2][1] = 1;
    bg[2] = (G.YV(G.n[1][1] + 1) + G.YV(0)) / 2.0;
  } else {
    ag[2][0] = 1;
    bg[2] = (G.XU(G.n[0][0] + 1) + G.XU(0)) / 2.0;
  }
  if (!solve3(ag, bg, xg)) return E_EDGE;
  // Line_Closest on the intersection line of the two nearest planes
  double t = 0;
  if (al != 0 || bl != 0 || cl != 0)
    t = (al * (x - xg[0]) + bl * (y - xg[1]) + cl * (z - xg[2])) / (al * al + bl * bl + cl * cl);
  const double xl = xg[0] + al * t, yl = xg[1] + bl * t, zl = xg[2] + cl * t;
  if (poly_inside(B, xl, yl, zl, dmin(G.dxmin / 1000., dmin(G.dymin / 10000., G.dzmin / 10000.)))) {
    xn = xl;
    yn = yl;
    zn = zl;
    return 0;
  }
  const double A3[3][3] = {{p1.a, p1.b, p1.c}, {p2.a, p2.b, p2.c}, {p3.a, p3.b, p3.c}};
  const double B3[3] = {-p1.d, -p2.d, -p3.d};
  if (!solve3(A3, B3, xg)) return E_VERTEX;
  xn = xg[0];
  yn = xg[1];
  zn = xg[2];
  return 0;
}

// Body%ClosestOnLineOut: t with x+(x2-x)*t on the boundary, x inside
SG_HD double closest_on_line_out(const Geom &G, const Body &B, double x, double y, double z, double x2, double y2,
                                 double z2) {
  double t1 = 0, t2 = 1, t;
  if (inside(G, B, x2, y2, z2, 0.0)) {
    t1 = 1;
    for (;;) {
      t2 = t2 * 1.1;
      if (!inside(G, B, x + (x2 - x) * t2, y + (y2 - y) * t2, z + (z2 - z) * t2, 0.0)) break;
      if (t2 >= 2) return 10;
    }
  }
  t = (t1 + t2) / 2.0;
  const double tol = dmin(G.dxmin / 1000., dmin(G.dymin / 1000., G.dzmin / 1000.));
  for (int i = 1; i <= 20; i++) {
    if (inside(G, B, x + (x2 - x) * t, y + (y2 - y) * t, z + (z2 - z) * t, 0.0))
      t1 = t;
    else
      t2 = t;
    t = (t1 + t2) / 2.0;
    if (fabs(t1 - t2) < tol) break;
  }
  return t;
}

// IB_interpolation_coefs_1st_order with n_ls = 2
SG_HD void ib_coefs(double xr, const double x[3], double out[2]) {
  const double A = (x[1] - xr) * (x[1] - xr) + (x[2] - xr) * (x[2] - xr);
  out[0] = (x[1] - xr) * (x[0] - xr) / A;
  out[1] = (x[2] - xr) * (x[0] - xr) / A;
}

SG_HD bool has_free_face_neighbour(const Geom &G, const int32_t *T, int i, int j, int k) {
  return T[G.lin(i + 1, j, k)] <= 0 || T[G.lin(i - 1, j, k)] <= 0 || T[G.lin(i, j + 1, k)] <= 0 ||
         T[G.lin(i, j - 1, k)] <= 0 || T[G.lin(i, j, k + 1)] <= 0 || T[G.lin(i, j, k - 1)] <= 0;
}

// TVelIBPoint_Create + InterpolationPoints for the velocity point (xi, yj,
// zk) of a component whose point coordinates are cx, cy, cz and type array
// T: the taps (at most 6) in out, their count returned (0: a point without
// taps).
SG_HD int vel_ib_point(const Geom &G, const Body *bodies, const int32_t *T, int xi, int yj, int zk, const double *cx,
                       const double *cy, const double *cz, Tap out[6]) {
  auto t = [&](int i, int j, int k) { return T[G.lin(i, j, k)]; };
  const double x = at3(cx, xi), y = at3(cy, yj), z = at3(cz, zk);
  const Body &SB = bodies[t(xi, yj, zk) - 1];
  // (the nested-domain boundary checks of the reference apply to domain
  // boundary codes 2000-2010 only, which this build does not accept)
  if (!inside(G, SB, x, y, z, 0.0)) return 0;
  double xn, yn, zn;
  closest_out(SB, xn, yn, zn, x, y, z);
  if (norm3(x - xn, y - yn, z - zn) <= G.ib_cut) return 0;
  const int nls = 2;
  auto free_run = [&](int a, int b, int c) {
    for (int s = 1; s <= nls; s++)
      if (t(xi + s * a, yj + s * b, zk + s * c) > 0) return false;
    return true;
  };
  const bool freexm = free_run(-1, 0, 0), freeym = free_run(0, -1, 0), freezm = free_run(0, 0, -1);
  const bool freexp = free_run(1, 0, 0), freeyp = free_run(0, 1, 0), freezp = free_run(0, 0, 1);
  const bool f1xm = t(xi - 1, yj, zk) <= 0, f1ym = t(xi, yj - 1, zk) <= 0, f1zm = t(xi, yj, zk - 1) <= 0;
  const bool f1xp = t(xi + 1, yj, zk) <= 0, f1yp = t(xi, yj + 1, zk) <= 0, f1zp = t(xi, yj, zk + 1) <= 0;
  if ((f1xm && !freexm) || (f1ym && !freeym) || (f1zm && !freezm) || (f1xp && !freexp) || (f1yp && !freeyp) ||
      (f1zp && !freezp))
    return 0;
  int nsolid = 0;
  for (int c = -1; c <= 1; c++)
    for (int b = -1; b <= 1; b++)
      for (int a = -1; a <= 1; a++)
        if (t(xi + a, yj + b, zk + c) > 0) nsolid++;
  if (nsolid > 21) return 0;
  if ((f1xm && f1xp) || (f1ym && f1yp) || (f1zm && f1zp)) return 0;
  int dirx = (!freexm && !freexp) ? 0 : freexm ? -1 : 1;
  int diry = (!freeym && !freeyp) ? 0 : freeym ? -1 : 1;
  int dirz = (!freezm && !freezp) ? 0 : freezm ? -1 : 1;
  double distx = 0, disty = 0, distz = 0;
  if (dirx != 0) {
    double tx = closest_on_line_out(G, SB, x, y, z, at3(cx, xi + dirx), y, z);
    if (tx > 1 || tx < 0.05) {
      dirx = 0;
      tx = 0;
    }
    distx = tx * (at3(cx, xi + dirx) - at3(cx, xi));
  }
  if (diry != 0) {
    double ty = closest_on_line_out(G, SB, x, y, z, x, at3(cy, yj + diry), z);
    if (ty > 1 || ty < 0.05) {
      diry = 0;
      ty = 0;
    }
    disty = ty * (at3(cy, yj + diry) - at3(cy, yj));
  }
  if (dirz != 0) {
    double tz = closest_on_line_out(G, SB, x, y, z, x, y, at3(cz, zk + dirz));
    if (tz > 1 || tz < 0.05) {
      dirz = 0;
      tz = 0;
    }
    distz = tz * (at3(cz, zk + dirz) - at3(cz, zk));
  }
  if (dirx == 0 && diry == 0 && dirz == 0) return 0;
  // InterpolationPoints
  Tap tmp[6];
  double d[3] = {1, 1, 1};
  if (dirx != 0) {
    const double xr = at3(cx, xi) + distx, xs[3] = {at3(cx, xi), at3(cx, xi + dirx), at3(cx, xi + 2 * dirx)};
    double cf[2];
    ib_coefs(xr, xs, cf);
    for (int s = 0; s < 2; s++) tmp[s] = Tap{xi + (s + 1) * dirx, yj, zk, cf[s]};
    d[0] = fabs(xs[0] - xr);
  }
  if (diry != 0) {
    const double yr = at3(cy, yj) + disty, ys[3] = {at3(cy, yj), at3(cy, yj + diry), at3(cy, yj + 2 * diry)};
    double cf[2];
    ib_coefs(yr, ys, cf);
    for (int s = 0; s < 2; s++) tmp[2 + s] = Tap{xi, yj + (s + 1) * diry, zk, cf[s]};
    d[1] = fabs(ys[0] - yr);
  }
  if (dirz != 0) {
    const double zr = at3(cz, zk) + distz, zs[3] = {at3(cz, zk), at3(cz, zk + dirz), at3(cz, zk + 2 * dirz)};
    double cf[2];
    ib_coefs(zr, zs, cf);
    for (int s = 0; s < 2; s++) tmp[4 + s] = Tap{xi, yj, zk + (s + 1) * dirz, cf[s]};
    d[2] = fabs(zs[0] - zr);
  }
  // Peller et al. weights b_d = prod(d)/d_d^2
  const double prod = d[0] * d[1] * d[2];
  double b[3] = {0, 0, 0};
  if (dirx != 0) b[0] = prod / (d[0] * d[0]);
  if (diry != 0) b[1] = prod / (d[1] * d[1]);
  if (dirz != 0) b[2] = prod / (d[2] * d[2]);
  const double csum = b[0] + b[1] + b[2];
  const int dirs[3] = {dirx, diry, dirz};
  int n = 0;
  for (int a = 0; a < 3; a++)
    if (dirs[a] != 0)
      for (int s = 0; s < 2; s++) {
        Tap tp = tmp[2 * a + s];
        tp.coef = tp.coef * b[a] / csum;
        out[n++] = tp;
      }
  return n;
}

SG_HD int fsign_nint(double v) { return __builtin_signbit(v) ? -1 : 1; }

// TScalFlIBPoint_Create for the Pr cell (xi, yj, zk): taps in out (at most
// 6), their count returned.
SG_HD int scalfl_ib_point(const Geom &G, const Body *bodies, const int32_t *prtype, int xi, int yj, int zk,
                          Tap out[6]) {
  auto pt = [&](int i, int j, int k) { return prtype[G.lin(i, j, k)]; };
  const double x = G.XP(xi), y = G.YP(yj), z = G.ZP(zk);
  const Body &SB = bodies[pt(xi, yj, zk) - 1];
  double xn, yn, zn;
  closest_out(SB, xn, yn, zn, x, y, z);
  double distx = xn - x, disty = yn - y, distz = zn - z;
  int dirx = fsign_nint(distx), diry = fsign_nint(disty), dirz = fsign_nint(distz);
  if (fabs(distx) < (G.XP(xi + 1) - G.XP(xi - 1)) / 100.) {
    distx = 0;
    dirx = 0;
  }
  if (fabs(disty) < (G.YP(yj + 1) - G.YP(yj - 1)) / 100.) {
    disty = 0;
    diry = 0;
  }
  if (fabs(distz) < (G.ZW(zk + 1) - G.ZW(zk - 1)) / 100.) {
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
  int n = 0;
  auto push = [&](int i, int j, int k, double c) { out[n++] = Tap{i, j, k, c}; };
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
      if (fp00) push(xi + 1, yj, zk, w);
      if (fm00) push(xi - 1, yj, zk, w);
      if (f0p0) push(xi, yj + 1, zk, w);
      if (f0m0) push(xi, yj - 1, zk, w);
      if (f00p) push(xi, yj, zk + 1, w);
      if (f00m) push(xi, yj, zk - 1, w);
    } else {
      push(xi + dx2, yj + dy2, zk + dz2, 1.0);
    }
  } else if (nfree == 1 || ndirs == 1) {
    push(xi + dirx, yj + diry, zk + dirz, 1.0);
  } else if (nfree == 2 || ndirs == 2) {
    double c1;
    if (dirx == 0) {  // plane yz
      if (fabs(disty / distz) < fabs(G.YP(yj + diry) - y) / fabs(G.ZP(zk + dirz) - z)) {
        const double tt = (G.ZP(zk + dirz) - z) / distz;
        c1 = fabs(disty * tt) / fabs(G.YP(yj + diry) - G.YP(yj));
        push(xi, yj + diry, zk + dirz, c1);
        push(xi, yj, zk + dirz, 1 - c1);
      } else {
        const double tt = (G.YP(yj + diry) - y) / disty;
        c1 = fabs(distz * tt) / fabs(G.ZP(zk + dirz) - G.ZP(zk));
        push(xi, yj + diry, zk + dirz, c1);
        push(xi, yj + diry, zk, 1 - c1);
      }
    } else if (diry == 0) {  // plane xz
      if (fabs(distx / distz) < fabs(G.XP(xi + dirx) - x) / fabs(G.ZP(zk + dirz) - z)) {
        const double tt = (G.ZP(zk + dirz) - z) / distz;
        c1 = fabs(distx * tt) / fabs(G.XP(xi + dirx) - G.XP(xi));
        push(xi + dirx, yj, zk + dirz, c1);
        push(xi, yj, zk + dirz, 1 - c1);
      } else {
        const double tt = (G.XP(xi + dirx) - x) / distx;
        c1 = fabs(distz * tt) / fabs(G.ZP(zk + dirz) - G.ZP(zk));
        push(xi + dirx, yj, zk + dirz, c1);
        push(xi + dirx, yj, zk, 1 - c1);
      }
    } else {  // plane xy
      if (fabs(distx / disty) < fabs(G.XP(xi + dirx) - x) / fabs(G.YP(yj + diry) - y)) {
        const double tt = (G.YP(yj + diry) - y) / disty;
        c1 = fabs(distx * tt) / fabs(G.XP(xi + dirx) - G.XP(xi));
        push(xi + dirx, yj + diry, zk, c1);
        push(xi, yj + diry, zk, 1 - c1);
      } else {
        const double tt = (G.XP(xi + dirx) - x) / distx;
        c1 = fabs(disty * tt) / fabs(G.YP(yj + diry) - G.YP(yj));
        push(xi + dirx, yj + diry, zk, c1);
        push(xi + dirx, yj, zk, 1 - c1);
      }
    }
  } else {
    const double tx = (G.XP(xi + dirx) - x) / distx;
    const double ty = (G.YP(yj + diry) - y) / disty;
    const double tz = (G.ZP(zk + dirz) - z) / distz;
    if (tx <= ty && tx <= tz) {
      const double fy = fabs(disty * tx) / fabs(G.YP(yj + diry) - y);
      const double fz = fabs(distz * tx) / fabs(G.ZP(zk + dirz) - z);
      push(xi + dirx, yj + diry, zk + dirz, fy * fz);
      push(xi + dirx, yj, zk + dirz, (1 - fy) * fz);
      push(xi + dirx, yj + diry, zk, fy * (1 - fz));
      push(xi + dirx, yj, zk, (1 - fy) * (1 - fz));
    } else if (ty <= tx && ty <= tz) {
      const double fx = fabs(distx * ty) / fabs(G.XP(xi + dirx) - x);
      const double fz = fabs(distz * ty) / fabs(G.ZP(zk + dirz) - z);
      push(xi + dirx, yj + diry, zk + dirz, fx * fz);
      push(xi, yj + diry, zk + dirz, (1 - fx) * fz);
      push(xi + dirx, yj + diry, zk, fx * (1 - fz));
      push(xi, yj + diry, zk, (1 - fx) * (1 - fz));
    } else {
      const double fx = fabs(distx * tz) / fabs(G.XP(xi + dirx) - x);
      const double fy = fabs(disty * tz) / fabs(G.YP(yj + diry) - y);
      push(xi + dirx, yj + diry, zk + dirz, fx * fy);
      push(xi, yj + diry, zk + dirz, (1 - fx) * fy);
      push(xi + dirx, yj, zk + dirz, fx * (1 - fy));
      push(xi, yj, zk + dirz, (1 - fx) * (1 - fy));
    }
  }
  return n;
}

// GetSolidBodiesWM for the free Pr point (i, j, k) (prtype -1): the nearest
// solid face neighbour's body.  Returns 1 with the distance vector and z0, 0
// when no face neighbour is solid, or a negative error code.
SG_HD int wm_pr_point(const Geom &G, const Body *bodies, const int32_t *prtype, int i, int j, int k,
                      double min_wall_distance, double dv[3], double *z0, int *bid) {
  const int nb[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
  const double x = G.XP(i), y = G.YP(j), z = G.ZP(k);
  double dist = DBL_MAX, nx_ = 0, ny_ = 0, nz_ = 0;
  int body = 0;
  for (int r = 0; r < 6; r++) {
    const int t = prtype[G.lin(i + nb[r][0], j + nb[r][1], k + nb[r][2])];
    if (t > 0 && t != body) {
      const int e = closest(G, bodies[t - 1], nx_, ny_, nz_, x, y, z);
      if (e) return -e;
      const double dd = norm3(nx_ - x, ny_ - y, nz_ - z);
      if (dd < dist) {
        dist = dd;
        body = t;
      }
    }
  }
  if (body == 0) return 0;
  dv[0] = nx_ - x;
  dv[1] = ny_ - y;
  dv[2] = nz_ - z;
  const double d = norm3(dv[0], dv[1], dv[2]);
  if (d == 0) return -E_WM_ZERO;
  if (d < min_wall_distance) return -E_WM_CLOSE;
  *z0 = bodies[body - 1].z0 > 0 ? bodies[body - 1].z0 : 0.0;
  *bid = body;
  return 1;
}

// GetSolidBodiesWM_UVW for the velocity point (i, j, k) (type -1) and one
// direction dir (-x, +x, -y, +y, -z, +z): 1 with the distance vector and z0
// when the neighbour in that direction is solid, else 0, or a negative error
// code.
SG_HD int wm_uvw_dir(const Geom &G, const Body *bodies, const int32_t *T, int i, int j, int k, int dir,
                     const double *cx, const double *cy, const double *cz, double min_wall_distance, double dv[3],
                     double *z0, int *bid) {
  const int dvec[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
  const int m = dvec[dir][0], n = dvec[dir][1], o = dvec[dir][2];
  const int t = T[G.lin(i + m, j + n, k + o)];
  if (t <= 0) return 0;
  const Body &SB = bodies[t - 1];
  const double X = at3(cx, i), Y = at3(cy, j), Z = at3(cz, k);
  const double Xn = at3(cx, i + m), Yn = at3(cy, j + n), Zn = at3(cz, k + o);
  double nx_, ny_, nz_;
  int e = closest(G, SB, nx_, ny_, nz_, X, Y, Z);
  if (e) return -e;
  dv[0] = nx_ - X;
  dv[1] = ny_ - Y;
  dv[2] = nz_ - Z;
  const double dist = norm3(dv[0], dv[1], dv[2]);
  const double cosang = (dv[0] * m + dv[1] * n + dv[2] * o) / (dist * norm3((double)m, (double)n, (double)o));
  if (!(cosang > 0.4)) {
    const double tl = closest_on_line_out(G, SB, Xn, Yn, Zn, X, Y, Z);
    nx_ = Xn + tl * (X - Xn);
    ny_ = Yn + tl * (Y - Yn);
    nz_ = Zn + tl * (Z - Zn);
    dv[0] = nx_ - X;
    dv[1] = ny_ - Y;
    dv[2] = nz_ - Z;
    const double dl = norm3(dv[0], dv[1], dv[2]);
    if (dl < dist && dl < min_wall_distance) {
      e = closest(G, SB, nx_, ny_, nz_, X, Y, Z);
      if (e) return -e;
      dv[0] = nx_ - X;
      dv[1] = ny_ - Y;
      dv[2] = nz_ - Z;
    }
  }
  if (norm3(dv[0], dv[1], dv[2]) < min_wall_distance) return -E_UVW_CLOSE;
  *z0 = SB.z0 > 0 ? SB.z0 : 0.0;
  *bid = t;
  return 1;
}

// SetNullifiedPoints: a velocity point whose 7-point stencil is solid
SG_HD bool null_point(const Geom &G, const int32_t *T, int i, int j, int k) {
  auto sol = [&](int p, int q, int r) { return T[G.lin(i + p, j + q, k + r)] > 0; };
  return sol(0, 0, 0) && sol(0, 0, 1) && sol(0, 0, -1) && sol(0, -1, 0) && sol(0, 1, 0) && sol(-1, 0, 0) &&
         sol(1, 0, 0);
}

#undef SG_HD

}  // namespace sg
}  // namespace elmm
