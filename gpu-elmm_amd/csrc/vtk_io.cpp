// vtk_io.cpp - the reference's field files: legacy VTK, RECTILINEAR_GRID,
// BINARY, big-endian float32 (no GPU involved).
//
//   out.vtk      SCALARS p, [SCALARS visc], VECTORS u (cell-centred
//                velocity: (U(i)+U(i-1))/2 etc.)       OutputOut, src/outputs.f90:1560-1745
//   scalars.vtk  SCALARS scalar01 .. scalarNN            OutputScalars, src/outputs.f90:1754-1854
//   U.vtk, V.vtk, W.vtk  the staggered components on their own interiors
//                                                         OutputUVW, src/outputs.f90:2260-2490
// and the restart reader ReadInitialConditions (src/initial.f90:3575-3811):
// p (0 when absent), the scalars, and U, V, W from U/V/W.vtk when all three
// exist, else rebuilt from out.vtk's cell-centred vectors.  Header lines are
// written as the reference writes them: 70-character records ("DIMENSIONS"
// with three list-directed integers from column 12, the coordinate lines
// with (i5,2x,a) from column 15), each followed by a line feed; the
// coordinate arrays and data blocks are followed by a line feed too
// (big-endian conversion: src/endianness.f90:12-136).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "case_setup.hpp"
#include "elmm_gpu.h"
#include "errors.hpp"

namespace {

using elmm::Grid1D;

uint32_t be32(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return __builtin_bswap32(u);
}
float from_be32(uint32_t u) {
  u = __builtin_bswap32(u);
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Writer {
  FILE *f = nullptr;
  bool ok = true;
  explicit Writer(const std::string &path) { f = std::fopen(path.c_str(), "wb"); }
  ~Writer() {
    if (f) std::fclose(f);
  }
  void raw(const void *p, size_t n) {
    if (ok && std::fwrite(p, 1, n, f) != n) ok = false;
  }
  void line(const std::string &s) {
    raw(s.data(), s.size());
    raw("\n", 1);
  }
  // a character(70) record: text padded with blanks to 70 characters
  void record(const std::string &s) {
    std::string r = s;
    r.resize(70, ' ');
    line(r);
  }
  void floats(const std::vector<float> &v) {
    std::vector<uint32_t> b(v.size());
    for (size_t q = 0; q < v.size(); q++) b[q] = be32(v[q]);
    raw(b.data(), b.size() * 4);
  }
};

// Fortran list-directed integers (each in 12 columns) from column 12.
std::string dims_record(int a, int b, int c) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "DIMENSIONS %12d%12d%12d", a, b, c);
  return buf;
}
std::string coord_record(const char *axis, int n) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%s_COORDINATES %5d  float", axis, n);
  return buf;
}
std::string points_record(long n) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "POINT_DATA %12ld", n);
  return buf;
}

// Header of one file: nx x ny x nz points at coordinates xs(1..nx),
// ys(1..ny), zs(1..nz) (Grid1D arrays).
void header(Writer &w, int nx, int ny, int nz, const std::vector<double> &xs, const std::vector<double> &ys,
            const std::vector<double> &zs) {
  w.line("# vtk DataFile Version 2.0");
  w.line("CLMM output file");
  w.line("BINARY");
  w.line("DATASET RECTILINEAR_GRID");
  w.record(dims_record(nx, ny, nz));
  auto coords = [&](const char *axis, const std::vector<double> &v, int n) {
    w.record(coord_record(axis, n));
    std::vector<float> c(n);
    for (int q = 0; q < n; q++) c[q] = (float)Grid1D::at(v, q + 1);
    w.floats(c);
    w.raw("\n", 1);
  };
  coords("X", xs, nx);
  coords("Y", ys, ny);
  coords("Z", zs, nz);
  w.record(points_record((long)nx * ny * nz));
}

// Element (i, j, k) of a ghosted host array (lower bound -2, extents n+6).
struct Ghosted {
  const double *a;
  long sy, sz;
  double at(int i, int j, int k) const { return a[(i + 2) + (j + 2) * sy + (long)(k + 2) * sz]; }
};

std::string join(const char *dir, const char *name) {
  std::string d = dir ? dir : "";
  if (!d.empty() && d.back() != '/') d += '/';
  return d + name;
}

}  // namespace

extern "C" int elmm_vtk_write(const char *dir, const elmm_config *cfg, const double *U, const double *V,
                              const double *W, const double *Pr, const double *visc, int32_t nscal,
                              const double *const *scalars) {
  if (!cfg || !U || !V || !W || !Pr || nscal < 0 || (nscal > 0 && !scalars))
    return elmm::fail(ELMM_E_ARG, "null argument to elmm_vtk_write");
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  const Grid1D g = elmm::make_grid(*cfg);
  int n[3][3];
  elmm::component_extents(*cfg, n);
  const long sy = nx + 6, sz = (long)(nx + 6) * (ny + 6);
  const Ghosted u{U, sy, sz}, v{V, sy, sz}, wv{W, sy, sz}, p{Pr, sy, sz};
  const long cells = (long)nx * ny * nz;
  // out.vtk
  {
    Writer w(join(dir, "out.vtk"));
    if (!w.f) return elmm::fail(ELMM_E_ARG, "cannot open " + join(dir, "out.vtk"));
    header(w, nx, ny, nz, g.xPr, g.yPr, g.zPr);
    std::vector<float> buf(cells);
    auto scalar = [&](const char *name, const Ghosted &a) {
      w.line(std::string("SCALARS ") + name + " float");
      w.line("LOOKUP_TABLE default");
      long q = 0;
      for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
          for (int i = 1; i <= nx; i++) buf[q++] = (float)a.at(i, j, k);
      w.floats(buf);
      w.raw("\n", 1);
    };
    scalar("p", p);
    if (visc) scalar("visc", Ghosted{visc, sy, sz});
    w.line("VECTORS u float");
    std::vector<float> vec(3 * cells);
    long q = 0;
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) {
          vec[q++] = (float)((u.at(i, j, k) + u.at(i - 1, j, k)) / 2.0);
          vec[q++] = (float)((v.at(i, j, k) + v.at(i, j - 1, k)) / 2.0);
          vec[q++] = (float)((wv.at(i, j, k) + wv.at(i, j, k - 1)) / 2.0);
        }
    w.floats(vec);
    w.raw("\n", 1);
    if (!w.ok) return elmm::fail(ELMM_E_ARG, "write error in " + join(dir, "out.vtk"));
  }
  // scalars.vtk
  if (nscal > 0) {
    Writer w(join(dir, "scalars.vtk"));
    if (!w.f) return elmm::fail(ELMM_E_ARG, "cannot open " + join(dir, "scalars.vtk"));
    header(w, nx, ny, nz, g.xPr, g.yPr, g.zPr);
    std::vector<float> buf(cells);
    for (int s = 0; s < nscal; s++) {
      char name[32];
      std::snprintf(name, sizeof name, "SCALARS scalar%02d float", s + 1);
      w.line(name);
      w.line("LOOKUP_TABLE default");
      const Ghosted a{scalars[s], sy, sz};
      long q = 0;
      for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
          for (int i = 1; i <= nx; i++) buf[q++] = (float)a.at(i, j, k);
      w.floats(buf);
      w.raw("\n", 1);
    }
    if (!w.ok) return elmm::fail(ELMM_E_ARG, "write error in " + join(dir, "scalars.vtk"));
  }
  // U.vtk, V.vtk, W.vtk on the staggered positions (xU for U, yV for V, zW for W)
  const char *files[3] = {"U.vtk", "V.vtk", "W.vtk"};
  const char *names[3] = {"U", "V", "W"};
  const Ghosted *comps[3] = {&u, &v, &wv};
  for (int c = 0; c < 3; c++) {
    Writer w(join(dir, files[c]));
    if (!w.f) return elmm::fail(ELMM_E_ARG, std::string("cannot open ") + join(dir, files[c]));
    header(w, n[c][0], n[c][1], n[c][2], c == 0 ? g.xU : g.xPr, c == 1 ? g.yV : g.yPr, c == 2 ? g.zW : g.zPr);
    w.line(std::string("SCALARS ") + names[c] + " float");
    w.line("LOOKUP_TABLE default");
    std::vector<float> buf((long)n[c][0] * n[c][1] * n[c][2]);
    long q = 0;
    for (int k = 1; k <= n[c][2]; k++)
      for (int j = 1; j <= n[c][1]; j++)
        for (int i = 1; i <= n[c][0]; i++) buf[q++] = (float)comps[c]->at(i, j, k);
    w.floats(buf);
    w.raw("\n", 1);
    if (!w.ok) return elmm::fail(ELMM_E_ARG, std::string("write error in ") + join(dir, files[c]));
  }
  return 0;
}

namespace {

struct Reader {
  std::vector<char> data;
  size_t pos = 0;
  bool open(const std::string &path) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    std::fseek(f, 0, SEEK_END);
    const long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    data.resize(n > 0 ? n : 0);
    const size_t got = n > 0 ? std::fread(data.data(), 1, n, f) : 0;
    std::fclose(f);
    return got == data.size();
  }
  // skip_to: the position just past the first occurrence of s
  bool skip_to(const std::string &s) {
    for (size_t q = pos; q + s.size() <= data.size(); q++)
      if (std::memcmp(&data[q], s.data(), s.size()) == 0) {
        pos = q + s.size();
        return true;
      }
    return false;
  }
  void skip_line() {
    while (pos < data.size() && data[pos] != '\n') pos++;
    if (pos < data.size()) pos++;
  }
  bool floats(std::vector<float> &out, size_t n) {
    if (pos + 4 * n > data.size()) return false;
    out.resize(n);
    for (size_t q = 0; q < n; q++) {
      uint32_t u;
      std::memcpy(&u, &data[pos + 4 * q], 4);
      out[q] = from_be32(u);
    }
    pos += 4 * n;
    return true;
  }
};

bool exists(const std::string &p) {
  FILE *f = std::fopen(p.c_str(), "rb");
  if (!f) return false;
  std::fclose(f);
  return true;
}

}  // namespace

extern "C" int elmm_vtk_read(const char *dir, const elmm_config *cfg, double *U, double *V, double *W, double *Pr,
                             int32_t nscal, double *const *scalars) {
  if (!cfg || !U || !V || !W || !Pr || nscal < 0 || (nscal > 0 && !scalars))
    return elmm::fail(ELMM_E_ARG, "null argument to elmm_vtk_read");
  const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
  int n[3][3];
  elmm::component_extents(*cfg, n);
  const long sy = nx + 6, sz = (long)(nx + 6) * (ny + 6);
  auto at = [&](double *a, int i, int j, int k) -> double & { return a[(i + 2) + (j + 2) * sy + (long)(k + 2) * sz]; };
  const long cells = (long)nx * ny * nz;
  std::vector<float> buf;
  Reader r;
  if (!r.open(join(dir, "out.vtk"))) return elmm::fail(ELMM_E_ARG, "Error opening " + join(dir, "out.vtk"));
  if (r.skip_to("SCALARS p float")) {
    r.skip_line();
    r.skip_line();
    if (!r.floats(buf, cells)) return elmm::fail(ELMM_E_ARG, "out.vtk: p field truncated");
    long q = 0;
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) at(Pr, i, j, k) = (double)buf[q++];
  } else {
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++)
        for (int i = 1; i <= nx; i++) at(Pr, i, j, k) = 0;
  }
  if (nscal > 0) {
    Reader rs;
    if (!rs.open(join(dir, "scalars.vtk"))) return elmm::fail(ELMM_E_ARG, "Error opening " + join(dir, "scalars.vtk"));
    for (int s = 0; s < nscal; s++) {
      char name[32];
      std::snprintf(name, sizeof name, "SCALARS scalar%02d float", s + 1);
      if (!rs.skip_to(name)) return elmm::fail(ELMM_E_ARG, std::string(name + 8) + " not found in scalars.vtk");
      rs.skip_line();
      rs.skip_line();
      if (!rs.floats(buf, cells)) return elmm::fail(ELMM_E_ARG, "scalars.vtk: field truncated");
      long q = 0;
      for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
          for (int i = 1; i <= nx; i++) at(scalars[s], i, j, k) = (double)buf[q++];
    }
  }
  const char *files[3] = {"U.vtk", "V.vtk", "W.vtk"};
  const char *tags[3] = {"SCALARS U float", "SCALARS V float", "SCALARS W float"};
  double *F[3] = {U, V, W};
  if (exists(join(dir, files[0])) && exists(join(dir, files[1])) && exists(join(dir, files[2]))) {
    for (int c = 0; c < 3; c++) {
      Reader rc;
      if (!rc.open(join(dir, files[c])) || !rc.skip_to(tags[c]))
        return elmm::fail(ELMM_E_ARG, std::string("no field in ") + join(dir, files[c]));
      rc.skip_line();
      rc.skip_line();
      const long m = (long)n[c][0] * n[c][1] * n[c][2];
      if (!rc.floats(buf, m)) return elmm::fail(ELMM_E_ARG, std::string(files[c]) + ": field truncated");
      long q = 0;
      for (int k = 1; k <= n[c][2]; k++)
        for (int j = 1; j <= n[c][1]; j++)
          for (int i = 1; i <= n[c][0]; i++) at(F[c], i, j, k) = (double)buf[q++];
    }
    return 0;
  }
  // cell-centred vectors of out.vtk back to the faces, as the reference
  // does (src/initial.f90:3715-3745): U(0:n-1) = U(0:n-1) + U(1:n) (array
  // assignment: every right-hand value is the old one, so i ascends), the
  // periodic face n adds face 0, then everything is halved
  r.pos = 0;
  if (!r.skip_to("VECTORS u float")) return elmm::fail(ELMM_E_ARG, "out.vtk: no VECTORS u");
  r.skip_line();
  if (!r.floats(buf, 3 * cells)) return elmm::fail(ELMM_E_ARG, "out.vtk: vectors truncated");
  const long tot = sz * (nz + 6);
  for (int c = 0; c < 3; c++) std::fill(F[c], F[c] + tot, 0.0);
  long q = 0;
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) {
        at(U, i, j, k) = (double)buf[q++];
        at(V, i, j, k) = (double)buf[q++];
        at(W, i, j, k) = (double)buf[q++];
      }
  for (int k = 1; k <= nz; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 0; i <= nx - 1; i++) at(U, i, j, k) = at(U, i, j, k) + at(U, i + 1, j, k);
  for (int k = 1; k <= nz; k++)
    for (int j = 0; j <= ny - 1; j++)
      for (int i = 1; i <= nx; i++) at(V, i, j, k) = at(V, i, j, k) + at(V, i, j + 1, k);
  for (int k = 0; k <= nz - 1; k++)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) at(W, i, j, k) = at(W, i, j, k) + at(W, i, j, k + 1);
  if (cfg->btype[ELMM_EA] == ELMM_BC_PERIODIC)
    for (int k = 1; k <= nz; k++)
      for (int j = 1; j <= ny; j++) at(U, nx, j, k) = at(U, nx, j, k) + at(U, 0, j, k);
  if (cfg->btype[ELMM_NO] == ELMM_BC_PERIODIC)
    for (int k = 1; k <= nz; k++)
      for (int i = 1; i <= nx; i++) at(V, i, ny, k) = at(V, i, ny, k) + at(V, i, 0, k);
  if (cfg->btype[ELMM_TO] == ELMM_BC_PERIODIC)
    for (int j = 1; j <= ny; j++)
      for (int i = 1; i <= nx; i++) at(W, i, j, nz) = at(W, i, j, nz) + at(W, i, j, 0);
  for (int c = 0; c < 3; c++)
    for (long e = 0; e < tot; e++) F[c][e] = F[c][e] / 2;
  return 0;
}
