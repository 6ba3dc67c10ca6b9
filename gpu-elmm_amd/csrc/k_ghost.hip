// k_ghost.hip - device-resident ghost-cell fills (replace the reference's
// MPI halo exchanges and serial ghost loops).
//
// The reference fills ghosts in passes: x faces, then y faces over the full
// x range, then z faces over the full x-y range, each pass reading values
// written by the earlier ones.  Every pass only copies, negates or applies an
// affine rule to one source value, so the final value of any ghost cell is a
// composition of at most three per-axis rules applied to one interior value.
// One launch therefore fills a whole ghost shell with each thread computing
// its cell's composite rule directly from interior data: no ordering between
// threads, no edge/corner special cases, one read and one write per cell.
#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

enum { BCN_NOSLIP = 1, BCN_FREESLIP = 2, BCN_PERIODIC = 3, BCN_DIRICHLET = 4, BCN_NEUMANN = 5 };

// Shell enumeration: region Z (k ghost, i,j in [zi0,zi1]x[zj0,zj1]),
// region Y (j ghost, k interior, i in [yi0,yi1]), region X (i ghost, j,k
// interior).  gl/gh: ghost layers below 1 and above n in every axis.
struct Shell {
  int n0, n1, n2;
  int gl, gh;
  int zi0, zi1, zj0, zj1;
  int yi0, yi1;
  long cz, cy, cx;  // region sizes
};

__host__ inline Shell make_shell(int n0, int n1, int n2, int gl, int gh, int zi0, int zi1, int zj0, int zj1,
                                 int yi0, int yi1) {
  Shell s{n0, n1, n2, gl, gh, zi0, zi1, zj0, zj1, yi0, yi1, 0, 0, 0};
  int g = gl + gh;
  s.cz = (long)(zi1 - zi0 + 1) * (zj1 - zj0 + 1) * g;
  s.cy = (long)(yi1 - yi0 + 1) * g * n2;
  s.cx = (long)g * n1 * n2;
  return s;
}

__device__ inline int ghost_index(int t, int gl, int n) { return t < gl ? t - gl + 1 : n + 1 + (t - gl); }

// 32-bit unsigned index arithmetic (a 64-bit division costs several times a
// 32-bit one).  A shell of a grid of at most 1024 x 1024 x 1024 cells has
// fewer than 2^26 cells.
__device__ inline bool decode(const Shell &s, long t_, int &i, int &j, int &k) {
  const unsigned g = (unsigned)(s.gl + s.gh);
  unsigned t = (unsigned)t_;
  if (t < (unsigned)s.cz) {
    const unsigned w = (unsigned)(s.zi1 - s.zi0 + 1), h = (unsigned)(s.zj1 - s.zj0 + 1);
    const unsigned r = t / w;
    i = s.zi0 + (int)(t - r * w);
    const unsigned kk = r / h;
    j = s.zj0 + (int)(r - kk * h);
    k = ghost_index((int)kk, s.gl, s.n2);
    return true;
  }
  t -= (unsigned)s.cz;
  if (t < (unsigned)s.cy) {
    const unsigned w = (unsigned)(s.yi1 - s.yi0 + 1);
    const unsigned r = t / w;
    i = s.yi0 + (int)(t - r * w);
    const unsigned kk = r / g;
    j = ghost_index((int)(r - kk * g), s.gl, s.n1);
    k = 1 + (int)kk;
    return true;
  }
  t -= (unsigned)s.cy;
  if (t < (unsigned)s.cx) {
    const unsigned r = t / g;
    i = ghost_index((int)(t - r * g), s.gl, s.n0);
    const unsigned kk = r / (unsigned)s.n1;
    j = 1 + (int)(r - kk * (unsigned)s.n1);
    k = 1 + (int)kk;
    return true;
  }
  return false;
}

// ---------------------------------------------------------------- BoundUVW
// Per-axis velocity rule (src/boundaries.f90:461-572, 783-894, 1122-1233,
// 1323-1434, 1533-1647, 1676-1789). Returns true for a zero value.
__device__ inline bool vel_rule(int m, int n, int blo, int bhi, bool normal, int &src, int &neg) {
  if (m <= 0) {
    if (blo == BCN_PERIODIC) { src = m + n; neg = 0; return false; }
    if (normal) {
      if (m == 0) return true;
      src = -m; neg = 1; return false;
    }
    if (blo == BCN_NOSLIP) { src = 1 - m; neg = 1; return false; }
    src = 1; neg = 0; return false;
  }
  if (bhi == BCN_PERIODIC) { src = m - n; neg = 0; return false; }
  if (normal) {
    if (m == n + 1) return true;
    src = 2 * n + 2 - m; neg = 1; return false;
  }
  if (bhi == BCN_NOSLIP) { src = 2 * n + 1 - m; neg = 1; return false; }
  src = n; neg = 0; return false;
}

// Composite rule of velocity ghost cell t of component c's shell: the cell
// (i, j, k) receives zero or (neg ? -1 : 1) times the interior value at
// (si, sj, sk).  Returns false past the end of the shell.
__device__ inline bool uvw_rule(const Geo &g, const Shell &s, int c, long t, int &i, int &j, int &k, int &si,
                                int &sj, int &sk, int &neg, bool &zero) {
  if (!decode(s, t, i, j, k)) return false;
  const int n[3] = {s.n0, s.n1, s.n2};
  neg = 0;
  zero = false;
  si = i;
  sj = j;
  sk = k;
  // outermost (last) pass first: z, then y, then x
  if (k < 1 || k > n[2]) {
    int nn;
    zero = vel_rule(k, n[2], g.bt[4], g.bt[5], c == 2, sk, nn);
    neg ^= zero ? 0 : nn;
  }
  if (!zero && (j < 1 || j > n[1])) {
    int nn;
    zero = vel_rule(j, n[1], g.bt[2], g.bt[3], c == 1, sj, nn);
    neg ^= zero ? 0 : nn;
  }
  if (!zero && (i < 1 || i > n[0])) {
    int nn;
    zero = vel_rule(i, n[0], g.bt[0], g.bt[1], c == 0, si, nn);
    neg ^= zero ? 0 : nn;
  }
  return true;
}

template <typename T>
__global__ void k_bound_uvw(Geo g, T *U, T *V, T *W, Shell su, Shell sv, Shell sw) {
  const int c = blockIdx.y;
  const Shell &s = c == 0 ? su : c == 1 ? sv : sw;
  T *F = c == 0 ? U : c == 1 ? V : W;
  int i, j, k, si, sj, sk, neg;
  bool zero;
  if (!uvw_rule(g, s, c, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k, si, sj, sk, neg, zero)) return;
  const T src = F[gix(g, si, sj, sk)];  // in the allocation also for a zero rule (load without a branch)
  const T v = zero ? T(0) : src;
  F[gix(g, i, j, k)] = neg ? -v : v;
}

// Ghost cells of component c whose source is a marked interior cell, as
// (ghost offset, source offset | neg << 62) pairs appended at list[2*count].
// Built once at set-up (the order of the entries is immaterial: every ghost
// cell appears at most once).
__global__ void k_ghost_map(Geo g, Shell s, int c, const uint8_t *__restrict__ mark, long long *__restrict__ list,
                            int cap, int *__restrict__ count) {
  int i, j, k, si, sj, sk, neg;
  bool zero;
  if (!uvw_rule(g, s, c, (long)blockIdx.x * blockDim.x + threadIdx.x, i, j, k, si, sj, sk, neg, zero)) return;
  if (zero) return;
  const long src = gix(g, si, sj, sk);
  if (!mark[src]) return;
  const int q = atomicAdd(count, 1);
  if (q < cap) {
    list[2 * q] = gix(g, i, j, k);
    list[2 * q + 1] = src | (neg ? (1LL << 62) : 0);
  }
}

__global__ void k_mark_cells(Geo g, const int32_t *__restrict__ ijk, int n, uint8_t *__restrict__ mark, uint8_t v) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) mark[gix(g, ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2])] = v;
}

// Re-applies the listed ghost rules (n[c] entries per component, U's
// first): after a change of the listed source cells only, this equals a full
// BoundUVW.
template <typename T>
__global__ void k_ghost_delta(T *U, T *V, T *W, const long long *__restrict__ list, int n0, int n1, int n2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n0 + n1 + n2) return;
  T *F = q < n0 ? U : q < n0 + n1 ? V : W;
  const long long dst = list[2 * q], e = list[2 * q + 1];
  const T v = F[e & ((1LL << 62) - 1)];
  F[dst] = (e >> 62) ? -v : v;
}

// ------------------------------------------------------ BoundViscosity
// src/scalarboundaries.f90:326-431: layers 0 and n+1, periodic or copy.
__device__ inline int visc_rule(int m, int n, bool periodic) {
  if (m == 0) return periodic ? n : 1;
  return periodic ? 1 : n;
}

template <typename T>
__global__ void k_bound_visc(Geo g, T *F, Shell s) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!decode(s, t, i, j, k)) return;
  int si = i, sj = j, sk = k;
  if (k < 1 || k > g.nz) sk = visc_rule(k, g.nz, g.bt[4] == BCN_PERIODIC);
  if (j < 1 || j > g.ny) sj = visc_rule(j, g.ny, g.bt[2] == BCN_PERIODIC);
  if (i < 1 || i > g.nx) si = visc_rule(i, g.nx, g.bt[0] == BCN_PERIODIC);
  F[gix(g, i, j, k)] = F[gix(g, si, sj, sk)];
}

// ------------------------------------------------------------ BoundScalar
// CommonBase, src/scalarboundaries.f90:18-279 (periodic, Neumann, Dirichlet;
// periodic z is rejected at set-up).  Inner rules (x) apply first.
struct ScalRule {
  int src;
  int kind;  // 0 copy, 1 add delta, 2 dirichlet
  double delta;
  double side;
};

__device__ inline ScalRule scal_rule(const Geo &g, int axis, int m) {
  const int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  const bool low = m <= 0;
  const int s = 2 * axis + (low ? 0 : 1);
  const int bt = g.sbt[s];
  const double side = g.side_scal[s];
  ScalRule r{0, 0, 0.0, side};
  if (bt == BCN_PERIODIC) {
    r.src = low ? m + n : m - n;
    return r;
  }
  if (bt == BCN_DIRICHLET) {
    r.kind = 2;
    r.src = low ? 1 - m : 2 * n + 1 - m;
    return r;
  }
  if (bt == 0) {  // TempBtype 0 (a Monin-Obukhov wall): the mirror
    r.src = low ? 1 - m : 2 * n + 1 - m;
    return r;
  }
  // Neumann: ghost = arr(1) - side*mult*d (low), arr(n) + side*mult*d (high)
  r.kind = 1;
  r.src = low ? 1 : n;
  int mult = low ? 1 - m : m - n;
  double d;
  if (axis == 0) {
    d = g.dxmin;
    d = mult == 1 ? side * d : (side * mult) * d;
  } else if (axis == 1) {
    d = g.dymin;
    d = mult == 1 ? side * d : (side * mult) * d;
  } else {
    if (low)
      d = mult == 1 ? g.dzW[0] : g.zPr[1] - g.zPr[1 - mult];
    else
      d = mult == 1 ? g.dzW[n + 1] : g.zPr[n + mult] - g.zPr[n];
    d = side * d;
  }
  r.delta = low ? -d : d;
  return r;
}

__device__ inline double apply_scal(const ScalRule &r, double v, bool low) {
  if (r.kind == 0) return v;
  if (r.kind == 2) return r.side - (v - r.side);
  // low: v - |term|, high: v + term (the reference's operation order)
  return low ? v - (-r.delta) : v + r.delta;
}

template <typename T>
__global__ void k_bound_scalar(Geo g, T *F, Shell s) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!decode(s, t, i, j, k)) return;
  ScalRule rz{}, ry{}, rx{};
  bool gz = k < 1 || k > g.nz, gy = j < 1 || j > g.ny, gx = i < 1 || i > g.nx;
  int si = i, sj = j, sk = k;
  if (gz) { rz = scal_rule(g, 2, k); sk = rz.src; }
  if (gy) { ry = scal_rule(g, 1, j); sj = ry.src; }
  if (gx) { rx = scal_rule(g, 0, i); si = rx.src; }
  double v = (double)F[gix(g, si, sj, sk)];
  if (gx) v = apply_scal(rx, v, i < 1);
  if (gy) v = apply_scal(ry, v, j < 1);
  if (gz) v = apply_scal(rz, v, k < 1);
  F[gix(g, i, j, k)] = (T)v;
}

// --------------------------------------------------------------- Bound_Pr
// src/boundaries.f90:1951-2090: face ghosts only; periodic copy or the
// extrapolation Pr(0)=Pr(3)-3*(Pr(2)-Pr(1)), Pr(-1)=Pr(2).
template <typename T>
__global__ void k_bound_pr(Geo g, T *P, Shell s) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (!decode(s, t, i, j, k)) return;
  int axis = (k < 1 || k > g.nz) ? 2 : (j < 1 || j > g.ny) ? 1 : 0;
  int m = axis == 0 ? i : axis == 1 ? j : k;
  int n = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  bool per = g.bt[2 * axis] == BCN_PERIODIC;
  auto at = [&](int q) -> T {
    int a = i, b = j, c = k;
    if (axis == 0) a = q; else if (axis == 1) b = q; else c = q;
    return P[gix(g, a, b, c)];
  };
  T v;
  if (per) {
    v = at(m <= 0 ? m + n : m - n);
  } else if (m == 0) {
    v = at(3) - 3 * (at(2) - at(1));
  } else if (m == -1) {
    v = at(2);
  } else if (m == n + 1) {
    v = at(n - 2) - 3 * (at(n - 1) - at(n));
  } else {
    v = at(n - 1);
  }
  P[gix(g, i, j, k)] = v;
}

// ---------------------------------------------------------------- launchers
static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

static Shell uvw_shell(const Geo &g, int c) {
  const int *n = g.n[c];
  return make_shell(n[0], n[1], n[2], 3, 3, -2, n[0] + 3, -2, n[1] + 3, -2, n[0] + 3);
}

template <typename T>
void launch_bound_uvw(const Geo &g, T *U, T *V, T *W, hipStream_t st) {
  Shell s[3];
  for (int c = 0; c < 3; c++) s[c] = uvw_shell(g, c);
  long mx = 0;
  for (int c = 0; c < 3; c++) mx = std::max(mx, s[c].cz + s[c].cy + s[c].cx);
  hipLaunchKernelGGL(k_bound_uvw<T>, dim3(nblk(mx, 256), 3), dim3(256), 0, st, g, U, V, W, s[0], s[1], s[2]);
}

long ghost_shell_cells(const Geo &g, int c) {
  const Shell s = uvw_shell(g, c);
  return s.cz + s.cy + s.cx;
}

int ghost_delta_build(const Geo &g, int c, const int32_t *ijk, int n, uint8_t *mark, long long *list, int cap,
                      int *count, hipStream_t st) {
  const Shell s = uvw_shell(g, c);
  int h = 0;
  if (hipMemcpyAsync(count, &h, sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess) return -1;
  if (n) hipLaunchKernelGGL(k_mark_cells, dim3(nblk(n, 256)), dim3(256), 0, st, g, ijk, n, mark, (uint8_t)1);
  hipLaunchKernelGGL(k_ghost_map, dim3(nblk(s.cz + s.cy + s.cx, 256)), dim3(256), 0, st, g, s, c, mark, list, cap,
                     count);
  if (n) hipLaunchKernelGGL(k_mark_cells, dim3(nblk(n, 256)), dim3(256), 0, st, g, ijk, n, mark, (uint8_t)0);
  if (hipMemcpyAsync(&h, count, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return -1;
  if (hipStreamSynchronize(st) != hipSuccess) return -1;
  return h;
}

template <typename T>
void launch_ghost_delta(T *U, T *V, T *W, const long long *list, const int n[3], hipStream_t st) {
  const int tot = n[0] + n[1] + n[2];
  if (tot) hipLaunchKernelGGL(k_ghost_delta<T>, dim3(nblk(tot, 256)), dim3(256), 0, st, U, V, W, list, n[0], n[1], n[2]);
}

template <typename T>
void launch_bound_visc(const Geo &g, T *F, hipStream_t st) {
  Shell s = make_shell(g.nx, g.ny, g.nz, 1, 1, 0, g.nx + 1, 0, g.ny + 1, 0, g.nx + 1);
  hipLaunchKernelGGL(k_bound_visc<T>, dim3(nblk(s.cz + s.cy + s.cx, 256)), dim3(256), 0, st, g, F, s);
}

template <typename T>
void launch_bound_scalar(const Geo &g, T *F, hipStream_t st) {
  Shell s = make_shell(g.nx, g.ny, g.nz, 3, 3, -1, g.nx + 2, -1, g.ny + 2, -1, g.nx + 2);
  hipLaunchKernelGGL(k_bound_scalar<T>, dim3(nblk(s.cz + s.cy + s.cx, 256)), dim3(256), 0, st, g, F, s);
}

template <typename T>
void launch_bound_pr(const Geo &g, T *P, hipStream_t st) {
  Shell s = make_shell(g.nx, g.ny, g.nz, 2, 2, 1, g.nx, 1, g.ny, 1, g.nx);
  hipLaunchKernelGGL(k_bound_pr<T>, dim3(nblk(s.cz + s.cy + s.cx, 256)), dim3(256), 0, st, g, P, s);
}

template void launch_ghost_delta<double>(double *, double *, double *, const long long *, const int[3], hipStream_t);
template void launch_ghost_delta<float>(float *, float *, float *, const long long *, const int[3], hipStream_t);
template void launch_bound_uvw<double>(const Geo &, double *, double *, double *, hipStream_t);
template void launch_bound_uvw<float>(const Geo &, float *, float *, float *, hipStream_t);
template void launch_bound_visc<double>(const Geo &, double *, hipStream_t);
template void launch_bound_visc<float>(const Geo &, float *, hipStream_t);
template void launch_bound_scalar<double>(const Geo &, double *, hipStream_t);
template void launch_bound_scalar<float>(const Geo &, float *, hipStream_t);
template void launch_bound_pr<double>(const Geo &, double *, hipStream_t);
template void launch_bound_pr<float>(const Geo &, float *, hipStream_t);

}  // namespace elmm
