// k_solids.hip - the solid-body masks and lists built on the device, equal
// bit for bit and in order to the host builder (solids.hip), which restates
// FindInsideCells (src/solidbodies.f90:75-206), FindNeighbouringCells
// (src/immersedboundary.f90:1475-1506), InitImBoundaries with
// TVelIBPoint_Create and TScalFlIBPoint_Create (src/immersedboundary.f90:
// 690-1410, 1650-1720), GetSolidBodiesWM(_UVW) (src/immersedboundary.f90:
// 14-308), SetNullifiedPoints (src/initial.f90:4861-4962) and the scalar part
// of InitWMMasks (src/wallmodels.f90:911-1021).
//
// Every cell is a thread.  The type arrays: a cell takes the id of the last
// body whose per-axis half-space pairs all hold (one bit set per axis index,
// evaluated by the host: 3 x (n+6) x bodies plane tests), then free cells next
// to a solid one become -1 in place (a neighbour that changes goes from 0 to
// -1, and only "> 0" is read, so the order of the threads does not matter).
// Each list is then built in two passes over the cells in field order, which
// is the reference's k, j, i loop order: a count pass (points, and taps for
// the IB lists; the first error of a wall-model task by an atomic minimum of
// (cell, direction, code)), an inclusive prefix sum, and a fill pass that
// writes each cell's entries at its offset.  The per-point geometry is the
// host builder's own code (solid_geom.hpp).
#include <hip/hip_runtime.h>

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "case_setup.hpp"
#include "device_common.hpp"
#include "errors.hpp"
#include "solid_geom.hpp"

namespace elmm {

// Defined in solids.hip.
sg::Geom make_geom(const elmm_config &c, const Grid1D &g, const CaseData &cd);
int make_bodies(int ncub, const elmm_cuboid *cub, std::vector<sg::Body> &bodies);
double solid_delta(const Grid1D &g);
void body_thermal(const elmm_config &c, const sg::Body &B, int8_t &tc, double &tv, double &qf);

namespace {

struct DevError {
  std::string msg;
};

#define SCHK(x)                                                                            \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) throw DevError{std::string(#x) + ": " + hipGetErrorString(e_)}; \
  } while (0)

struct Cell {
  int i, j, k;
};

__device__ inline Cell decode(long p, long sy, int ny6, long sz) {
  return Cell{(int)(p % sy) - 2, (int)((p / sy) % ny6) - 2, (int)(p / sz) - 2};
}

// FindInsideCells: hx, hy, hz hold `words` words per index (-2..e+3) of the
// bodies whose half-space pair on that axis holds; the cell takes the last
// body (highest id) of their intersection.  Cells outside -2..e+3 are 0.
__global__ void k_solid_types(int32_t *T, long ntot, long sy, int ny6, long sz, int e0, int e1, int e2,
                              const uint32_t *hx, const uint32_t *hy, const uint32_t *hz, int words) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const Cell c = decode(p, sy, ny6, sz);
  int id = 0;
  if (c.i <= e0 + 3 && c.j <= e1 + 3 && c.k <= e2 + 3)
    for (int w = words - 1; w >= 0; w--) {
      const uint32_t m = hx[(c.i + 2) * words + w] & hy[(c.j + 2) * words + w] & hz[(c.k + 2) * words + w];
      if (m) {
        id = w * 32 + (31 - __builtin_clz(m)) + 1;
        break;
      }
    }
  T[p] = id;
}

// FindNeighbouringCells: free cells in -1..e+2 with a solid cell among their
// 27 neighbours become -1.
__global__ void k_solid_aux(int32_t *T, long ntot, long sy, int ny6, long sz, int e0, int e1, int e2) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const Cell c = decode(p, sy, ny6, sz);
  if (c.i < -1 || c.i > e0 + 2 || c.j < -1 || c.j > e1 + 2 || c.k < -1 || c.k > e2 + 2) return;
  if (T[p] != 0) return;
  bool solid = false;
  for (int z = -1; z <= 1; z++)
    for (int y = -1; y <= 1; y++)
      for (int x = -1; x <= 1; x++) solid |= T[p + x + y * sy + z * sz] > 0;
  if (solid) T[p] = -1;
}

enum Kind { K_IBV, K_IBS, K_WMP, K_WMUVW, K_NULL, K_SCFL };

struct ListArgs {
  int kind, dir;
  int lo[3], hi[3];             // index box of the reference's loop
  const int32_t *T;             // type array of the list's points
  const double *cx, *cy, *cz;   // coordinates of those points (lower bound -3)
  double mwd;                   // minimum wall distance
  long off;                     // scalar faces: offset of the +d neighbour
  int32_t *pcnt, *tcnt;         // count pass: points and taps per cell
  unsigned long long *errkey;   // count pass: first error of the task
  const int32_t *poff, *toff;   // fill pass: inclusive sums of the counts
  int32_t *ijk, *ptr, *src;     // fill pass outputs
  double *coef, *dist, *z0;
  int32_t *bid;                 // wall-model points: the body
  uint8_t *mask;                // scalar face mask (fill pass, every cell)
};

template <bool FILL>
__global__ void k_solid_list(sg::Geom G, const sg::Body *bodies, ListArgs a, long ntot, int ny6) {
  const long p = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (p >= ntot) return;
  const Cell c = decode(p, G.sy, ny6, G.sz);
  const bool in = c.i >= a.lo[0] && c.i <= a.hi[0] && c.j >= a.lo[1] && c.j <= a.hi[1] && c.k >= a.lo[2] &&
                  c.k <= a.hi[2];
  int np = 0, nt = 0, err = 0;
  sg::Tap taps[6];
  double dv[3] = {0, 0, 0}, z0 = 0;
  int bid = 0;
  bool face_solid = false;
  if (in) switch (a.kind) {
      case K_IBV:
        if (a.T[p] > 0 && sg::has_free_face_neighbour(G, a.T, c.i, c.j, c.k)) {
          np = 1;
          nt = sg::vel_ib_point(G, bodies, a.T, c.i, c.j, c.k, a.cx, a.cy, a.cz, taps);
        }
        break;
      case K_IBS:
        if (a.T[p] > 0 && sg::has_free_face_neighbour(G, a.T, c.i, c.j, c.k)) {
          np = 1;
          nt = sg::scalfl_ib_point(G, bodies, a.T, c.i, c.j, c.k, taps);
        }
        break;
      case K_WMP:
        if (a.T[p] < 0) {
          const int r = sg::wm_pr_point(G, bodies, a.T, c.i, c.j, c.k, a.mwd, dv, &z0, &bid);
          if (r < 0) err = -r;
          else np = r;
        }
        break;
      case K_WMUVW:
        if (a.T[p] < 0) {
          const int r =
              sg::wm_uvw_dir(G, bodies, a.T, c.i, c.j, c.k, a.dir, a.cx, a.cy, a.cz, a.mwd, dv, &z0, &bid);
          if (r < 0) err = -r;
          else np = r;
        }
        break;
      case K_NULL:
        np = sg::null_point(G, a.T, c.i, c.j, c.k);
        break;
      case K_SCFL: {
        const bool s0 = a.T[p] > 0, s1 = a.T[p + a.off] > 0;
        face_solid = s0 || s1;
        np = s0 != s1;
        break;
      }
    }
  if (!FILL) {
    a.pcnt[p] = np;
    if (a.tcnt) a.tcnt[p] = nt;
    if (err) atomicMin(a.errkey, ((unsigned long long)p * 8 + a.dir) * 8 + err);
    return;
  }
  if (a.kind == K_SCFL) a.mask[p] = face_solid ? 0 : 1;
  if (!np) return;
  const int q = a.poff[p] - 1;
  a.ijk[3 * q] = c.i;
  a.ijk[3 * q + 1] = c.j;
  a.ijk[3 * q + 2] = c.k;
  if (a.kind == K_IBV || a.kind == K_IBS) {
    const int t1 = a.toff[p], t0 = t1 - nt;
    for (int s = 0; s < nt; s++) {
      a.src[3 * (t0 + s)] = taps[s].i;
      a.src[3 * (t0 + s) + 1] = taps[s].j;
      a.src[3 * (t0 + s) + 2] = taps[s].k;
      a.coef[t0 + s] = taps[s].coef;
    }
    a.ptr[q] = t1;
  } else if (a.kind == K_WMP || a.kind == K_WMUVW) {
    a.dist[3 * q] = dv[0];
    a.dist[3 * q + 1] = dv[1];
    a.dist[3 * q + 2] = dv[2];
    a.z0[q] = z0;
    a.bid[q] = bid;
  }
}

// A device buffer carved into aligned pieces (one allocation per phase).
struct Slab {
  char *base = nullptr;
  size_t used = 0, cap = 0;
  bool sizing = true;  // first walk: only add up the sizes
  template <class T>
  T *take(size_t n) {
    const size_t at = (used + 255) & ~(size_t)255;
    used = at + std::max<size_t>(n, 1) * sizeof(T);
    return sizing ? nullptr : reinterpret_cast<T *>(base + at);
  }
  void allocate() {
    cap = used;
    SCHK(hipMalloc((void **)&base, std::max<size_t>(cap, 1)));
    used = 0;
    sizing = false;
  }
  ~Slab() {
    if (base) (void)hipFree(base);
  }
};

// One list: its kernel arguments and where its result goes on the host.
struct ListJob {
  ListArgs a;
  bool taps;
  IBList *ib = nullptr;
  WMList *wm = nullptr;
  std::vector<int32_t> *ijk = nullptr;
  int n = 0, nt = 0;
  std::vector<int32_t> ptr;  // IB end offsets before the leading 0
};

struct SetupClock {
  bool on = std::getenv("ELMM_SETUP_TRACE") != nullptr;
  std::chrono::steady_clock::time_point t = std::chrono::steady_clock::now();
  void mark(const char *what) {
    if (!on) return;
    const auto n = std::chrono::steady_clock::now();
    std::fprintf(stderr, "[elmm case/device] %-12s %8.2f ms\n", what,
                 std::chrono::duration<double, std::milli>(n - t).count());
    t = n;
  }
};

// Pinned host staging of the downloads, kept for later builds (one build at
// a time uses it; it grows to the largest case seen).
std::mutex g_stage_mu;
char *g_stage = nullptr;
size_t g_stage_cap = 0;

char *staging(size_t bytes) {
  if (bytes > g_stage_cap) {
    if (g_stage) SCHK(hipHostFree(g_stage));
    g_stage = nullptr;
    g_stage_cap = 0;
    SCHK(hipHostMalloc((void **)&g_stage, bytes));
    g_stage_cap = bytes;
  }
  return g_stage;
}

struct Stream {
  hipStream_t s = nullptr;
  int32_t *pinned = nullptr;
  ~Stream() {
    if (pinned) (void)hipHostFree(pinned);
    if (s) (void)hipStreamDestroy(s);
  }
};

int build_on_device(const elmm_config &c, const Grid1D &g, int ncub, const elmm_cuboid *cub, CaseData &cd) {
  SetupClock clock;
  std::vector<sg::Body> bodies;
  if (int rc = make_bodies(ncub, cub, bodies)) return rc;
  const sg::Geom Gh = make_geom(c, g, cd);
  const double delta = solid_delta(g);
  const long ntot = cd.ntot;
  if (ntot >= INT_MAX) return fail(ELMM_E_ARG, "grid too large for the device case builder");
  const int ny6 = c.ny + 6, grid = (int)((ntot + 255) / 256);
  int ext[4][3];
  for (int d = 0; d < 3; d++) {
    ext[0][d] = d == 0 ? c.nx : d == 1 ? c.ny : c.nz;
    for (int comp = 0; comp < 3; comp++) ext[comp + 1][d] = Gh.n[comp][d];
  }
  const std::vector<double> *cxh[4] = {&g.xPr, &g.xU, &g.xPr, &g.xPr}, *cyh[4] = {&g.yPr, &g.yPr, &g.yV, &g.yPr},
                            *czh[4] = {&g.zPr, &g.zPr, &g.zPr, &g.zW};
  // per-axis body bit sets of the four type arrays (host)
  const int words = (ncub + 31) / 32;
  std::vector<uint32_t> bits[12];
  for (int a = 0; a < 4; a++) {
    const std::vector<double> *co[3] = {cxh[a], cyh[a], czh[a]};
    for (int d = 0; d < 3; d++) {
      std::vector<uint32_t> &h = bits[3 * a + d];
      const int hi = ext[a][d] + 3;
      h.assign((size_t)(hi + 3) * words, 0u);
      for (int b = 0; b < ncub; b++)
        for (int i = -2; i <= hi; i++)
          if (sg::axis_hit(Gh, bodies[b], d, Grid1D::at(*co[d], i), delta))
            h[(size_t)(i + 2) * words + b / 32] |= 1u << (b % 32);
    }
  }

  Stream S;
  SCHK(hipStreamCreateWithFlags(&S.s, hipStreamNonBlocking));
  const hipStream_t st = S.s;
  size_t scan_bytes = 0;
  SCHK(hipcub::DeviceScan::InclusiveSum(nullptr, scan_bytes, (int32_t *)nullptr, (int32_t *)nullptr, (int)ntot, st));

  // fixed buffers: coordinates, bodies, bit sets, type arrays, counts,
  // offsets, scan space, scalar face masks, error keys (sized, then taken)
  Slab fix;
  sg::Geom G = Gh;
  double *crd[6];
  const std::vector<double> *crh[6] = {&g.xU, &g.yV, &g.zW, &g.xPr, &g.yPr, &g.zPr};
  sg::Body *dbodies = nullptr;
  uint32_t *dbits[12];
  int32_t *T[4], *pcnt, *tcnt, *poff, *toff;
  void *scan_tmp;
  uint8_t *mask[3];
  unsigned long long *errkey;
  for (int pass = 0; pass < 2; pass++) {
    for (int q = 0; q < 6; q++) crd[q] = fix.take<double>(crh[q]->size());
    dbodies = fix.take<sg::Body>(bodies.size());
    for (int q = 0; q < 12; q++) dbits[q] = fix.take<uint32_t>(bits[q].size());
    for (int a = 0; a < 4; a++) T[a] = fix.take<int32_t>(ntot);
    pcnt = fix.take<int32_t>(ntot);
    tcnt = fix.take<int32_t>(ntot);
    poff = fix.take<int32_t>(ntot);
    toff = fix.take<int32_t>(ntot);
    scan_tmp = fix.take<uint8_t>(scan_bytes);
    for (int d = 0; d < 3; d++) mask[d] = fix.take<uint8_t>(c.wall_model ? ntot : 1);
    errkey = fix.take<unsigned long long>(4);
    if (pass == 0) fix.allocate();
  }
  auto up = [&](void *d, const void *h, size_t bytes) {
    if (bytes) SCHK(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
  };
  for (int q = 0; q < 6; q++) up(crd[q], crh[q]->data(), crh[q]->size() * sizeof(double));
  up(dbodies, bodies.data(), bodies.size() * sizeof(sg::Body));
  for (int q = 0; q < 12; q++) up(dbits[q], bits[q].data(), bits[q].size() * sizeof(uint32_t));
  const unsigned long long none[4] = {ULLONG_MAX, ULLONG_MAX, ULLONG_MAX, ULLONG_MAX};
  up(errkey, none, sizeof none);
  G.xU = crd[0];
  G.yV = crd[1];
  G.zW = crd[2];
  G.xPr = crd[3];
  G.yPr = crd[4];
  G.zPr = crd[5];
  const double *cxd[4] = {G.xPr, G.xU, G.xPr, G.xPr}, *cyd[4] = {G.yPr, G.yPr, G.yV, G.yPr},
               *czd[4] = {G.zPr, G.zPr, G.zPr, G.zW};

  // type arrays
  for (int a = 0; a < 4; a++) {
    hipLaunchKernelGGL(k_solid_types, dim3(grid), dim3(256), 0, st, T[a], ntot, cd.sy, ny6, cd.sz, ext[a][0],
                       ext[a][1], ext[a][2], dbits[3 * a], dbits[3 * a + 1], dbits[3 * a + 2], words);
    SCHK(hipGetLastError());
    hipLaunchKernelGGL(k_solid_aux, dim3(grid), dim3(256), 0, st, T[a], ntot, cd.sy, ny6, cd.sz, ext[a][0],
                       ext[a][1], ext[a][2]);
    SCHK(hipGetLastError());
  }

  // the lists, in the host builder's order
  std::vector<ListJob> jobs;
  auto job = [&](int kind, int t, int nx, int ny, int nz) {
    ListJob J;
    J.a = ListArgs{};
    J.a.kind = kind;
    J.a.T = T[t];
    J.a.lo[0] = J.a.lo[1] = J.a.lo[2] = 1;
    J.a.hi[0] = nx;
    J.a.hi[1] = ny;
    J.a.hi[2] = nz;
    J.a.mwd = delta;
    if (t > 0) {
      J.a.cx = cxd[t];
      J.a.cy = cyd[t];
      J.a.cz = czd[t];
    }
    J.taps = kind == K_IBV || kind == K_IBS;
    return J;
  };
  for (int comp = 0; comp < 3; comp++) {
    ListJob J = job(K_IBV, comp + 1, Gh.n[comp][0], Gh.n[comp][1], Gh.n[comp][2]);
    J.ib = &cd.ib[comp];
    jobs.push_back(J);
  }
  {
    ListJob J = job(K_IBS, 0, c.nx, c.ny, c.nz);
    J.ib = &cd.ib[3];
    jobs.push_back(J);
  }
  if (c.wall_model) {
    ListJob J = job(K_WMP, 0, c.nx, c.ny, c.nz);
    J.a.errkey = errkey;
    J.wm = &cd.wmp;
    jobs.push_back(J);
    for (int comp = 0; comp < 3; comp++)
      for (int dir = 0; dir < 6; dir++) {
        ListJob U = job(K_WMUVW, comp + 1, Gh.n[comp][0], Gh.n[comp][1], Gh.n[comp][2]);
        U.a.dir = dir;
        U.a.errkey = errkey + 1 + comp;
        U.wm = &cd.wm[comp * 6 + dir];
        jobs.push_back(U);
      }
  }
  for (int comp = 0; comp < 3; comp++) {
    ListJob J = job(K_NULL, comp + 1, Gh.n[comp][0], Gh.n[comp][1], Gh.n[comp][2]);
    J.ijk = &cd.null_ijk[comp];
    jobs.push_back(J);
  }
  if (c.wall_model)
    for (int d = 0; d < 3; d++) {
      ListJob J = job(K_SCFL, 0, c.nx, c.ny, c.nz);
      J.a.lo[d] = 0;
      J.a.off = d == 0 ? 1 : d == 1 ? cd.sy : cd.sz;
      J.a.mask = mask[d];
      J.ijk = &cd.scfl[d];
      jobs.push_back(J);
    }

  // count passes: inclusive sums, totals to pinned memory
  const int nj = (int)jobs.size();
  SCHK(hipHostMalloc((void **)&S.pinned, 2 * nj * sizeof(int32_t)));
  auto count = [&](ListJob &J, int idx) {
    ListArgs a = J.a;
    a.pcnt = pcnt;
    a.tcnt = J.taps ? tcnt : nullptr;
    hipLaunchKernelGGL(k_solid_list<false>, dim3(grid), dim3(256), 0, st, G, dbodies, a, ntot, ny6);
    SCHK(hipGetLastError());
    SCHK(hipcub::DeviceScan::InclusiveSum(scan_tmp, scan_bytes, pcnt, poff, (int)ntot, st));
    if (J.taps) SCHK(hipcub::DeviceScan::InclusiveSum(scan_tmp, scan_bytes, tcnt, toff, (int)ntot, st));
    if (idx >= 0) {
      SCHK(hipMemcpyAsync(&S.pinned[2 * idx], poff + ntot - 1, sizeof(int32_t), hipMemcpyDeviceToHost, st));
      if (J.taps) SCHK(hipMemcpyAsync(&S.pinned[2 * idx + 1], toff + ntot - 1, sizeof(int32_t),
                                      hipMemcpyDeviceToHost, st));
    }
  };
  for (int q = 0; q < nj; q++) count(jobs[q], q);
  unsigned long long keys[4];
  SCHK(hipMemcpyAsync(keys, errkey, sizeof keys, hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  clock.mark("counts");
  // the first error in task order (Pr, then U, V, W), and within a task in
  // the serial order
  if (c.wall_model)
    for (int t = 0; t < 4; t++)
      if (keys[t] != ULLONG_MAX) return fail(ELMM_E_ARG, sg::solid_error((int)(keys[t] % 8)));

  // outputs of every list in one allocation, then the fill passes, the last
  // list first (its offsets are in place), every other one after its count
  // pass again (one offset buffer serves every list)
  Slab out;
  for (int pass = 0; pass < 2; pass++) {
    for (int q = 0; q < nj; q++) {
      ListJob &J = jobs[q];
      J.n = S.pinned[2 * q];
      J.nt = J.taps ? S.pinned[2 * q + 1] : 0;
      J.a.ijk = out.take<int32_t>(3 * (size_t)J.n);
      if (J.taps) {
        J.a.ptr = out.take<int32_t>(J.n);
        J.a.src = out.take<int32_t>(3 * (size_t)J.nt);
        J.a.coef = out.take<double>(J.nt);
      }
      if (J.wm) {
        J.a.dist = out.take<double>(3 * (size_t)J.n);
        J.a.z0 = out.take<double>(J.n);
        J.a.bid = out.take<int32_t>(J.n);
      }
    }
    if (pass == 0) out.allocate();
  }
  for (int q = nj - 1; q >= 0; q--) {
    ListJob &J = jobs[q];
    if (q < nj - 1) count(J, -1);  // the last job's offsets are still in place
    ListArgs a = J.a;
    a.poff = poff;
    a.toff = toff;
    hipLaunchKernelGGL(k_solid_list<true>, dim3(grid), dim3(256), 0, st, G, dbodies, a, ntot, ny6);
    SCHK(hipGetLastError());
  }

  // downloads: the type arrays, the scalar face masks and every list's
  // output, three contiguous regions, into the pinned staging buffer; then
  // copied into the case's arrays on threads
  const char *t0 = reinterpret_cast<const char *>(T[0]);
  const size_t tbytes = reinterpret_cast<const char *>(T[3] + ntot) - t0;
  const char *m0 = reinterpret_cast<const char *>(mask[0]);
  const size_t mbytes = c.wall_model ? reinterpret_cast<const char *>(mask[2] + ntot) - m0 : 0;
  std::lock_guard<std::mutex> stage_lock(g_stage_mu);
  char *stage = staging(tbytes + mbytes + out.cap);
  SCHK(hipMemcpyAsync(stage, t0, tbytes, hipMemcpyDeviceToHost, st));
  if (mbytes) SCHK(hipMemcpyAsync(stage + tbytes, m0, mbytes, hipMemcpyDeviceToHost, st));
  SCHK(hipMemcpyAsync(stage + tbytes + mbytes, out.base, out.cap, hipMemcpyDeviceToHost, st));
  SCHK(hipStreamSynchronize(st));
  clock.mark("download");
  const char *os = stage + tbytes + mbytes;
  auto from_out = [&](auto &v, const void *d, size_t n) {
    v.resize(n);
    if (n) std::memcpy(v.data(), os + (reinterpret_cast<const char *>(d) - out.base), n * sizeof(v[0]));
  };
  std::vector<int32_t> *Th[4] = {&cd.prtype, &cd.utype, &cd.vtype, &cd.wtype};
  run_tasks(8, [&](int task) {
    if (task < 4) {
      Th[task]->resize(ntot);
      std::memcpy(Th[task]->data(), stage + (reinterpret_cast<const char *>(T[task]) - t0), ntot * sizeof(int32_t));
    } else if (task < 7) {
      const int d = task - 4;
      if (c.wall_model) {
        cd.scfl_mask[d].resize(ntot);
        std::memcpy(cd.scfl_mask[d].data(), stage + tbytes + (reinterpret_cast<const char *>(mask[d]) - m0), ntot);
      }
    } else {
      for (ListJob &J : jobs) {
        from_out(J.ib ? J.ib->ijk : J.wm ? J.wm->ijk : *J.ijk, J.a.ijk, 3 * (size_t)J.n);
        if (J.ib) {
          from_out(J.ptr, J.a.ptr, J.n);
          from_out(J.ib->src, J.a.src, 3 * (size_t)J.nt);
          from_out(J.ib->coef, J.a.coef, J.nt);
          J.ib->ptr.assign(1, 0);
          J.ib->ptr.insert(J.ib->ptr.end(), J.ptr.begin(), J.ptr.end());
        }
        if (J.wm) {
          from_out(J.wm->dist, J.a.dist, 3 * (size_t)J.n);
          from_out(J.wm->z0, J.a.z0, J.n);
          std::vector<int32_t> bid;
          from_out(bid, J.a.bid, J.n);
          J.wm->tclass.resize(J.n);
          J.wm->tval.resize(J.n);
          J.wm->qval.assign(J.n, 0.0);
          J.wm->qflux.resize(J.n);
          for (int q = 0; q < J.n; q++)
            body_thermal(c, bodies[bid[q] - 1], J.wm->tclass[q], J.wm->tval[q], J.wm->qflux[q]);
        }
      }
    }
  });
  clock.mark("lists");
  return 0;
}

}  // namespace

int build_solids_device(const elmm_config &c, const Grid1D &g, int ncub, const elmm_cuboid *cub, CaseData &cd) {
  HostOpLock lock;
  try {
    return build_on_device(c, g, ncub, cub, cd);
  } catch (const DevError &e) {
    return fail(ELMM_E_HIP, "device case builder: " + e.msg);
  } catch (const std::bad_alloc &) {
    return fail(ELMM_E_NOMEM, "out of host memory in the device case builder");
  }
}

}  // namespace elmm
