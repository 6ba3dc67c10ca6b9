// k_pressure.hip - pressure projection on the device: divergence -> FFT
// Poisson solve -> velocity correction + pressure update + reference fix.
//
// Replaces PrePoisson/PostPoisson/FixPressureToReference
// (src/pressure.f90:272-723, 728-991, 994-1229) and the external FFT
// library behind Poiss_PoisSolver (src/poisson.f90:76-210), solving the
// FD2 operator of Residue_variable_z (src/poisson.f90:23-60):
//   periodic x, y : discrete Fourier modes, eigenvalue -4 sin^2(pi m/n)/d^2
//   z (walls)     : tridiagonal per mode with the staggered-Neumann rows,
//                   solved by Thomas elimination, one thread per mode.
// Passes (each one read + one write of the spectral array in HBM):
//   1 divergence + x FFT of each (j,k) row  (RHS never goes to HBM)
//   2 y FFT of tiles of 8 x-modes
//   3 z tridiagonal solve of every (x-mode, y-mode) pair
//   4 inverse y FFT
//   5 inverse x FFT, real part -> Phi
#include <cmath>
#include <cstdlib>
#include <type_traits>

#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Spectral arithmetic in R (double: FP64 form; float: the inner solves of the
// precision form).  C2<R> is double2 / float2.
__device__ inline double2 mk2(double x, double y) { return make_double2(x, y); }
__device__ inline float2 mk2(float x, float y) { return make_float2(x, y); }
template <typename V>
__device__ inline V cmul(V a, V b) {
  return mk2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename V>
__device__ inline V cadd(V a, V b) {
  return mk2(a.x + b.x, a.y + b.y);
}
template <typename V>
__device__ inline V csub(V a, V b) {
  return mk2(a.x - b.x, a.y - b.y);
}

__device__ inline int bitrev(int v, int logn) { return logn ? (int)(__brev((unsigned)v) >> (32 - logn)) : 0; }

// LDS slot of element m of a line: one pad slot per 8 elements, plus one per
// 64 with padmode 1.  fspan: slots one line occupies.
__device__ inline int fpad(int m, int padmode) { return m + (m >> 3) + (padmode ? (m >> 6) : 0); }
__host__ __device__ inline int fspan(int n, int padmode) {
  const int m = n - 1;
  return m + (m >> 3) + (padmode ? (m >> 6) : 0) + 1;
}
__host__ __device__ inline int twlen(int n) { return n / 2; }

// Radix-2 decimation-in-time FFT of CNT independent lines of length 2^LOGN
// by NT threads (the block size).  The stages run three at a time on
// 8-element sets held in registers (the sets {base + (e << s0)} are closed
// under stages s0..s0+2), so every butterfly is the same operation on the
// same operands as a stage-by-stage radix-2 sweep over bit-reversed input.
//   first round  : each set is loaded straight from global memory
//                  (load(line, natural index), all eight elements before
//                  fin(line, index, loaded value) turns each into the
//                  transform's input), so the input never takes a
//                  bit-reversed trip through LDS.  With PermFirst the sets
//                  are dealt to threads in bit-reversed order, which makes
//                  the natural indices of one element slot consecutive
//                  across consecutive threads (coalesced loads along a line);
//   middle rounds: in place in LDS, line l at buf[l*stride + fpad(pos)];
//   last round   : the outputs (natural order) go from registers straight to
//                  global memory (store(line, index, value)).
// LineMajor: consecutive threads take consecutive sets of one line (rows
// contiguous along the transform, the x passes); otherwise consecutive
// threads take the same set of consecutive lines (lines side by side in
// memory, the y pass).  tw_g (global) and tw_s (LDS, written before the
// first round) hold the same table tw[t] = exp(-2 pi i t / n); inverse uses
// the conjugate.
template <int LOGN, int CNT, int NT, bool LineMajor, bool PermFirst, typename V, typename Load, typename Fin,
          typename Store>
__device__ __forceinline__ void fft_lines(V *buf, int stride, int padmode, const V *__restrict__ tw_g, const V *tw_s,
                                          bool inverse, Load load, Fin fin, Store store) {
  constexpr int N = 1 << LOGN;
  if constexpr (LOGN == 0) {
    for (int l = threadIdx.x; l < CNT; l += NT) store(l, 0, fin(l, 0, load(l, 0)));  // length 1: identity
  } else {
#pragma unroll
    for (int s0 = 0; s0 < LOGN; s0 += 3) {
      const int nb = LOGN - s0 < 3 ? LOGN - s0 : 3;
      const int ne = 1 << nb;
      const int per_line = N >> nb;
      const int nsets = per_line * CNT;
      const bool first = s0 == 0, last = s0 + nb >= LOGN;
      const V *tw = first ? tw_g : tw_s;
#pragma unroll
      for (int it = 0; it < (nsets + NT - 1) / NT; it++) {
        const int q = (int)threadIdx.x + it * NT;
        if (nsets % NT == 0 || q < nsets) {
          const int l = LineMajor ? q / per_line : q % CNT;
          int r = LineMajor ? q % per_line : q / CNT;
          if (PermFirst && first) r = bitrev(r, LOGN - nb);
          const int lo = r & ((1 << s0) - 1), hi = r >> s0;
          const int base = (hi << (s0 + nb)) | lo;
          V x[8];
          if (first) {
            decltype(load(0, 0)) raw[8];
#pragma unroll
            for (int e = 0; e < 8; e++)
              if (e < ne) raw[e] = load(l, bitrev(base + (e << s0), LOGN));
#pragma unroll
            for (int e = 0; e < 8; e++)
              if (e < ne) x[e] = fin(l, bitrev(base + (e << s0), LOGN), raw[e]);
          } else {
#pragma unroll
            for (int e = 0; e < 8; e++)
              if (e < ne) x[e] = buf[l * stride + fpad(base + (e << s0), padmode)];
          }
#pragma unroll
          for (int b = 0; b < 3; b++) {
            if (b >= nb) break;
            const int half = 1 << (s0 + b), tstep = N / (2 * half);
#pragma unroll
            for (int e = 0; e < 8; e++) {
              if (e >= ne || (e >> b) & 1) continue;
              const int e1 = e | (1 << b);
              const int pos = (base + (e << s0)) & (half - 1);
              V w = tw[pos * tstep];
              if (inverse) w.y = -w.y;
              const V a = x[e], c = cmul(x[e1], w);
              x[e] = cadd(a, c);
              x[e1] = csub(a, c);
            }
          }
#pragma unroll
          for (int e = 0; e < 8; e++) {
            if (e >= ne) break;
            const int pos = base + (e << s0);
            if (last)
              store(l, pos, x[e]);
            else
              buf[l * stride + fpad(pos, padmode)] = x[e];
          }
        }
      }
      if (!last) __syncthreads();
    }
  }
}

// Lines per block of the x passes: 2048 elements per block (4 rows of 512;
// 2 rows measured 5 % slower, 8 rows 45 % slower); threads: one per
// 8-element set of the rows, 64..256.
__host__ __device__ constexpr int xrows(int logn) { return logn >= 11 ? 1 : (2048 >> logn); }
__host__ __device__ constexpr int xthreads(int logn, int rows) {
  return rows * (logn >= 3 ? 1 << (logn - 3) : 1) < 64    ? 64
         : rows * (logn >= 3 ? 1 << (logn - 3) : 1) > 256 ? 256
                                                          : rows * (logn >= 3 ? 1 << (logn - 3) : 1);
}
constexpr int kFftThreads = 256;

// twiddle table of length n/2 from global memory into LDS
template <typename V>
__device__ inline void load_twiddles(V *dst, const V *__restrict__ src, int n) {
  for (int t = threadIdx.x; t < n / 2; t += blockDim.x) dst[t] = src[t];
}

// Pass 1: RHS = ((dU/dx + dV/dy + dW/dz) - Q) / coef for the rows of this
// block, then forward FFT in x.  One block handles ROWS rows.
// The divergence is evaluated in FP64 whatever the field type T.
// Residual of the solved FD2 operator at one interior cell, in FP64:
// r = RHS - A Phi with A the operator the spectral solve inverts (periodic x
// and y by index wrap; z rows a_k (Phi_{k-1} - Phi_k) + c_k (Phi_{k+1} -
// Phi_k), a_1 = c_nz = 0).
__device__ inline double fd2_residual(const Geo &g, const double *__restrict__ Phi, const double *__restrict__ RHS,
                                      const double *__restrict__ az, const double *__restrict__ cz, int i, int j,
                                      int k, long p) {
  const double rx = 1 / (g.dxmin * g.dxmin), ry = 1 / (g.dymin * g.dymin);
  const double ph = Phi[p];
  const double pw = Phi[i > 1 ? p - 1 : p - 1 + g.nx], pe = Phi[i < g.nx ? p + 1 : p + 1 - g.nx];
  const double ps = Phi[j > 1 ? p - g.sy : p - g.sy + (long)g.ny * g.sy];
  const double pn = Phi[j < g.ny ? p + g.sy : p + g.sy - (long)g.ny * g.sy];
  double a = ((pw - ph) + (pe - ph)) * rx + ((ps - ph) + (pn - ph)) * ry;
  if (k > 1) a = a + az[k - 1] * (Phi[p - g.sz] - ph);
  if (k < g.nz) a = a + cz[k - 1] * (Phi[p + g.sz] - ph);
  return RHS[p] - a;
}

// Right-hand side of the row, by MODE: kRhsDiv (+ kRhsQ, + kRhsOut) the
// divergence of (U, V, W) (minus Q) over coef, stored to rhs_out with
// kRhsOut (first solve); kRhsIn a stored right-hand side; kRhsResid the FP64
// residual rhs_in - A phi (refinement sweeps).  The variants are separate
// instantiations so that the element loads carry no branches and issue
// together.  Q is non-zero on few cells: qmask[row] has bit s set when
// elements 64s..64s+63 of the row may hold a non-zero Q, and only those are
// read (elsewhere r - 0 = r).  Rows past the end load row nrows-1 and are not
// stored.  ny is a power of two: row = (j-1) + ny (k-1).
// kRhsFast: uniform z and reciprocals of dx, dy, dz that satisfy div_rn's
// condition (the quotients need no fallback branch).
enum { kRhsDiv = 0, kRhsQ = 1, kRhsOut = 2, kRhsIn = 4, kRhsResid = 8, kRhsFast = 16 };

template <typename T, typename R, int LOGN, int MODE, int ROWS>
__global__ void __launch_bounds__(kFftThreads)
    k_fft_x_fwd(Geo g, PoissonPlan<R> pp, const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
                const double *__restrict__ Q, const uint32_t *__restrict__ qmask, const double *__restrict__ rhs_in,
                const double *__restrict__ phi, double dt2_rec, double *__restrict__ rhs_out, int logny,
                int padmode) {
  constexpr int N = 1 << LOGN, NT = xthreads(LOGN, ROWS);
  extern __shared__ __align__(16) unsigned char sh_raw[];
  C2<R> *tw = reinterpret_cast<C2<R> *>(sh_raw);
  C2<R> *sh = tw + twlen(N);
  load_twiddles(tw, pp.twx, N);
  const int ls = fspan(N, padmode), M = pp.M;
  const int row0 = blockIdx.x * ROWS;
  const int nrows = pp.ny * pp.nz;
  auto cell = [&](int l, int m, int &row) {
    row = min(row0 + l, nrows - 1);
    return gix(g, m + 1, (row & (pp.ny - 1)) + 1, (row >> logny) + 1);
  };
  auto load = [&](int l, int m) -> double {
    int row;
    const long p = cell(l, m, row);
    const int k = (row >> logny) + 1;
    if constexpr (MODE == kRhsResid) {
      return fd2_residual(g, phi, rhs_in, pp.az, pp.cz, m + 1, (row & (pp.ny - 1)) + 1, k, p);
    } else if constexpr (MODE == kRhsIn) {
      return rhs_in[p];
    } else {
      const double du = (double)U[p] - (double)U[p - 1], dv = (double)V[p] - (double)V[p - g.sy];
      const double dw = (double)W[p] - (double)W[p - g.sz];
      if constexpr ((MODE & kRhsFast) != 0)
        return div_rn_nz(du, g.dxmin, g.rdx) + div_rn_nz(dv, g.dymin, g.rdy) + div_rn_nz(dw, g.dzmin, g.rdz);
      else
        return div_rn(du, g.dxmin, g.rdx) + div_rn(dv, g.dymin, g.rdy) +
               (g.variable_z ? dw / g.dzPr[k] : div_rn(dw, g.dzmin, g.rdz));
    }
  };
  auto fin = [&](int l, int m, double r) -> C2<R> {
    if constexpr ((MODE & (kRhsResid | kRhsIn)) == 0) {
      int row;
      const long p = cell(l, m, row);
      if constexpr ((MODE & kRhsQ) != 0) {
        if ((qmask[row] >> (m >> 6)) & 1u) r = r - Q[p];
      }
      r = r * dt2_rec;
      if constexpr ((MODE & kRhsOut) != 0) {
        if (row0 + l < nrows) rhs_out[p] = r;
      }
    }
    return mk2((R)r, (R)0);
  };
  auto store = [&](int l, int m, C2<R> v) {
    const int row = row0 + l;
    if (row < nrows && m < M) pp.X[(long)row * pp.Mp + m] = v;
  };
  fft_lines<LOGN, ROWS, NT, true, true>(sh, ls, padmode, pp.twx, tw, false, load, fin, store);
}

// LDS line stride of the y pass: odd, so that the TX lines touched by one
// global row segment (consecutive lanes, one element each) fall in distinct
// banks.
__host__ __device__ inline int ystride(int ny, int padmode) { return fspan(ny, padmode) | 1; }

// Passes 2 and 4: FFT along y of TX consecutive x-modes for one z-plane.
template <int TX, typename R, int LOGN>
__global__ void __launch_bounds__(kFftThreads) k_fft_y(PoissonPlan<R> pp, int inverse, int padmode) {
  constexpr int N = 1 << LOGN;
  extern __shared__ __align__(16) unsigned char sh_raw[];
  C2<R> *tw = reinterpret_cast<C2<R> *>(sh_raw);
  C2<R> *sh = tw + twlen(N);
  load_twiddles(tw, pp.twy, N);
  const int M = pp.M, Mp = pp.Mp, ls = ystride(N, padmode);
  const int m0 = blockIdx.x * TX, k = blockIdx.y;
  C2<R> *X = pp.X + (long)k * N * Mp;
  auto load = [&](int c, int j) -> C2<R> {  // branch-free: modes past M load M-1
    const int m = m0 + c;
    const C2<R> v = X[(long)j * Mp + min(m, M - 1)];
    return m < M ? v : mk2((R)0, (R)0);
  };
  auto store = [&](int c, int j, C2<R> v) {
    const int m = m0 + c;
    if (m < M) X[(long)j * Mp + m] = v;
  };
  fft_lines<LOGN, TX, kFftThreads, false, false>(sh, ls, padmode, pp.twy, tw, inverse != 0, load,
                                                 [](int, int, C2<R> v) { return v; }, store);
}

// Pass 3: per (m, j) mode, a_k x_{k-1} + (lam - a_k - c_k) x_k + c_k x_{k+1}
// = d_k (a_1 = c_nz = 0), one thread per mode, Thomas elimination in FP64 for
// both forms (the lowest modes' tridiagonals are too ill-conditioned for an
// FP32 elimination inside a one-sweep refinement).  Every mode but the
// horizontal mean has lam < 0, so the rows are strictly diagonally dominant
// and need no pivoting.  The forward state (c', d') is checkpointed in LDS at
// the end of every SEG levels; the back substitution recomputes one segment
// at a time into registers from its checkpoint, so the spectral array is
// read twice and written once, coalesced across the modes of a wavefront.
// The singular mode (m = j = 0) pins x_1 = 0 and is integrated exactly.
// LDS (sized at launch): the checkpoints, 3 x ceil(nz/ZSEG) x 64 values, and
// the mean mode's partial sums, 2 x nz values.
constexpr int ZSEG = 16;
constexpr int kSolveZMaxNz = 1024;  // 112 KiB of LDS at the maximum

__host__ __device__ constexpr size_t solve_z_lds(int nz) {
  return (size_t)(3 * ((nz + ZSEG - 1) / ZSEG) * 64 + 2 * nz) * sizeof(double);
}
static_assert(solve_z_lds(kSolveZMaxNz) <= 160 * 1024, "the z solve's LDS at the largest nz exceeds a CU's 160 KiB");

template <typename R, bool PIPE>
__global__ void __launch_bounds__(64) k_solve_z(PoissonPlan<R> pp) {
  extern __shared__ double zlds[];
  const int nseg = (pp.nz + ZSEG - 1) / ZSEG;
  const int t = threadIdx.x;
  // checkpoint a (0: c', 1: d' real, 2: d' imaginary) of segment s
  auto ck = [&](int a, int s) -> double & { return zlds[(a * nseg + s) * 64 + t]; };
  double *gm0 = zlds + 3 * nseg * 64, *gm1 = gm0 + pp.nz;
  // one thread per slot of a padded level (the padding threads return), so
  // that a wavefront's loads are one contiguous segment (measured 5 %
  // faster than a compact mode numbering)
  const long plane = (long)pp.ny * pp.Mp;
  const long q = (long)blockIdx.x * 64 + t;
  if (q >= plane) return;
  const int m = (int)(q % pp.Mp), j = (int)(q / pp.Mp);
  if (m >= pp.M) return;
  const int nz = pp.nz;
  C2<R> *X = pp.X + q;
  if (q == 0) {
    // Horizontal-mean mode with x_1 pinned to 0 (rows 2..nz of the Neumann
    // problem): G_{k-1} = G_k - d_k dzPr(k) from G_nz = 0 downwards, then
    // x_{k+1} = x_k + G_k dzW(k).  Both sums are compensated (Kahan), so the
    // mode is exact to rounding instead of carrying the nz^2 growth of an
    // elimination on the singular operator.
    // (loops unrolled so that the independent loads of several levels are in
    // flight together: this one thread bounds the kernel on small grids)
    double Gr = 0, Gi = 0, cr = 0, ci = 0;
#pragma unroll 8
    for (int k = nz; k >= 2; k--) {
      const double dz = pp.dzPr[k];
      const C2<R> d = X[(long)(k - 1) * plane];
      double tt = -(double)d.x * dz, y = tt - cr, s2 = Gr + y;
      cr = (s2 - Gr) - y;
      Gr = s2;
      tt = -(double)d.y * dz;
      y = tt - ci;
      s2 = Gi + y;
      ci = (s2 - Gi) - y;
      Gi = s2;
      gm0[k - 2] = Gr;
      gm1[k - 2] = Gi;
    }
    double xr = 0, xi = 0;
    cr = ci = 0;
    X[0] = mk2((R)0, (R)0);
#pragma unroll 8
    for (int k = 1; k <= nz - 1; k++) {
      const double dz = pp.dzW[k];
      double tt = gm0[k - 1] * dz, y = tt - cr, s2 = xr + y;
      cr = (s2 - xr) - y;
      xr = s2;
      tt = gm1[k - 1] * dz;
      y = tt - ci;
      s2 = xi + y;
      ci = (s2 - xi) - y;
      xi = s2;
      X[(long)k * plane] = mk2((R)xr, (R)xi);
    }
    return;
  }
  const double lam = pp.lamx[m] + pp.lamy[j];
  // forward elimination; den = b - a c', every division a multiplication by
  // 1/den.  PIPE (grids with several wavefronts per CU): the levels are
  // loaded one segment ahead of the elimination (the load index clamped into
  // the column), so that a wavefront keeps a segment of loads in flight while
  // it works through the dependent chain.  Without PIPE (few modes, one
  // wavefront per CU, measured faster there) each level is loaded where it
  // is used, the loop unrolled by 8.  Same operations in the same order.
  double cp = 0, dr = 0, di = 0;
  C2<R> cur[ZSEG], nxt[ZSEG];
  if constexpr (PIPE) {
#pragma unroll
    for (int r = 0; r < ZSEG; r++) cur[r] = X[(long)min(r, nz - 1) * plane];
    for (int k0 = 0; k0 < nz; k0 += ZSEG) {
#pragma unroll
      for (int r = 0; r < ZSEG; r++) nxt[r] = X[(long)min(k0 + ZSEG + r, nz - 1) * plane];
#pragma unroll
      for (int r = 0; r < ZSEG; r++) {
        const int k = k0 + r;
        if (k < nz) {
          const double a = pp.az[k], c = pp.cz[k];
          const double b = lam - a - c;
          const C2<R> x = cur[r];
          const double inv = 1.0 / (b - a * cp);
          cp = c * inv;
          dr = ((double)x.x - a * dr) * inv;
          di = ((double)x.y - a * di) * inv;
        }
      }
      // state at the end of the segment (levels k0 .. k0+ZSEG-1)
      ck(0, k0 / ZSEG) = cp;
      ck(1, k0 / ZSEG) = dr;
      ck(2, k0 / ZSEG) = di;
#pragma unroll
      for (int r = 0; r < ZSEG; r++) cur[r] = nxt[r];
    }
  } else {
#pragma unroll 8
    for (int k = 0; k < nz; k++) {
      const double a = pp.az[k], c = pp.cz[k];
      const double b = lam - a - c;
      const C2<R> x = X[(long)k * plane];
      const double inv = 1.0 / (b - a * cp);
      cp = c * inv;
      dr = ((double)x.x - a * dr) * inv;
      di = ((double)x.y - a * di) * inv;
      if (k % ZSEG == ZSEG - 1) {
        ck(0, k / ZSEG) = cp;
        ck(1, k / ZSEG) = dr;
        ck(2, k / ZSEG) = di;
      }
    }
  }
  // back substitution x_k = d'_k - c'_k x_{k+1}, segment by segment (PIPE:
  // the next, lower segment's levels loaded while this one is worked
  // through)
  double xr = 0, xi = 0;
  const int stop = (nz - 1) / ZSEG;
  if constexpr (PIPE) {
#pragma unroll
    for (int r = 0; r < ZSEG; r++) cur[r] = X[(long)min(stop * ZSEG + r, nz - 1) * plane];
  }
  for (int s = stop; s >= 0; s--) {
    const int k0 = s * ZSEG;
    double scp[ZSEG], sdr[ZSEG], sdi[ZSEG];
    if constexpr (PIPE) {
#pragma unroll
      for (int r = 0; r < ZSEG; r++) nxt[r] = X[(long)max(k0 - ZSEG + r, 0) * plane];
    }
    cp = s ? ck(0, s - 1) : 0.0;
    dr = s ? ck(1, s - 1) : 0.0;
    di = s ? ck(2, s - 1) : 0.0;
#pragma unroll
    for (int r = 0; r < ZSEG; r++) {
      const int k = k0 + r;
      if (k < nz) {
        const double a = pp.az[k], c = pp.cz[k];
        const double b = lam - a - c;
        const C2<R> x = PIPE ? cur[r] : X[(long)k * plane];
        const double inv = 1.0 / (b - a * cp);
        cp = c * inv;
        dr = ((double)x.x - a * dr) * inv;
        di = ((double)x.y - a * di) * inv;
      }
      scp[r] = cp;
      sdr[r] = dr;
      sdi[r] = di;
    }
#pragma unroll
    for (int r = ZSEG - 1; r >= 0; r--) {
      const int k = k0 + r;
      if (k < nz) {
        xr = sdr[r] - scp[r] * xr;
        xi = sdi[r] - scp[r] * xi;
        X[(long)k * plane] = mk2((R)xr, (R)xi);
      }
    }
    if constexpr (PIPE) {
#pragma unroll
      for (int r = 0; r < ZSEG; r++) cur[r] = nxt[r];
    }
  }
}

// Pass 5: inverse x FFT from the half spectrum (Hermitian symmetry), real
// part scaled by 1/(nx*ny) -> Phi interior (FP64); with accumulate the
// result is added to Phi (iterative-refinement correction).
template <typename R, int LOGN, int ROWS>
__global__ void __launch_bounds__(kFftThreads)
    k_fft_x_inv(Geo g, PoissonPlan<R> pp, double *__restrict__ Phi, int accumulate, int logny, int padmode) {
  constexpr int N = 1 << LOGN, NT = xthreads(LOGN, ROWS);
  extern __shared__ __align__(16) unsigned char sh_raw[];
  C2<R> *tw = reinterpret_cast<C2<R> *>(sh_raw);
  C2<R> *sh = tw + twlen(N);
  load_twiddles(tw, pp.twx, N);
  const int ls = fspan(N, padmode), M = pp.M;
  const int row0 = blockIdx.x * ROWS;
  const int nrows = pp.ny * pp.nz;
  const R scale = (R)1.0 / ((R)N * pp.ny);
  auto load = [&](int l, int m) -> C2<R> {  // branch-free; rows past the end load row nrows-1
    const int row = min(row0 + l, nrows - 1);
    C2<R> v = pp.X[(long)row * pp.Mp + (m < M ? m : N - m)];
    if (m >= M) v.y = -v.y;
    return v;
  };
  auto store = [&](int l, int m, C2<R> v) {
    const int row = row0 + l;
    if (row < nrows) {
      const int j = (row & (pp.ny - 1)) + 1, k = (row >> logny) + 1;
      const long q = gix(g, m + 1, j, k);
      const double val = (double)(v.x * scale);
      Phi[q] = accumulate ? Phi[q] + val : val;
    }
  };
  fft_lines<LOGN, ROWS, NT, true, true>(sh, ls, padmode, pp.twx, tw, true, load, [](int, int, C2<R> v) { return v; },
                                        store);
}

static int ilog2(int n) {
  int l = 0;
  while ((1 << l) < n) l++;
  return l;
}

// LDS padding of the lines (fpad's padmode), measured per pass on the
// 512-point lines: the forward x pass is fastest with one pad slot per 8
// elements, the inverse x pass with one more per 64 (-9 %), the y pass is the
// same with either.
constexpr int kPadXFwd = 0, kPadXInv = 1, kPadY = 0;

// Dynamic LDS above 64 KiB must be opted into per kernel (gfx950 has 160 KiB
// per workgroup).
static void allow_lds(const void *fn, size_t bytes) {
  if (bytes > 65536) (void)lds_limit_raise(fn, bytes);
}

// Runs f(std::integral_constant<int, L>{}) for L = logn (0..10: the engine
// admits power-of-two nx, ny <= 1024).
template <typename F>
static void with_logn(int logn, F f) {
  switch (logn) {
    case 0: f(std::integral_constant<int, 0>{}); break;
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 9: f(std::integral_constant<int, 9>{}); break;
    default: f(std::integral_constant<int, 10>{}); break;
  }
}

template <typename T, typename R>
void launch_fft_x_fwd(const Geo &g, const PoissonPlan<R> &pp, const T *U, const T *V, const T *W, const double *Q,
                      const uint32_t *qmask, const double *rhs, const double *phi, double dt2_rec, double *rhs_out,
                      hipStream_t st) {
  if (!fft_pow2(pp.nx) || !fft_pow2(pp.ny)) {
    launch_fft_x_fwd_mr<T, R>(g, pp, U, V, W, Q, qmask, rhs, phi, dt2_rec, rhs_out, st);
    return;
  }
  const int pm = kPadXFwd, nrows = pp.ny * pp.nz;
  const bool fast = !g.variable_z && g.rdx != 0 && g.rdy != 0 && g.rdz != 0;
  const int mode =
      phi ? kRhsResid : rhs ? kRhsIn : (Q ? kRhsQ : 0) | (rhs_out ? kRhsOut : 0) | (fast ? kRhsFast : 0);
  with_logn(ilog2(pp.nx), [&](auto L) {
    constexpr int LOGN = decltype(L)::value, ROWS = xrows(LOGN);
    const size_t sh = ((size_t)ROWS * fspan(1 << LOGN, pm) + twlen(1 << LOGN)) * sizeof(C2<R>);
    auto go = [&](auto MD) {
      constexpr int MODE = decltype(MD)::value;
      allow_lds((const void *)k_fft_x_fwd<T, R, LOGN, MODE, ROWS>, sh);
      hipLaunchKernelGGL((k_fft_x_fwd<T, R, LOGN, MODE, ROWS>), dim3(nblk(nrows, ROWS)), dim3(xthreads(LOGN, ROWS)),
                         sh, st, g, pp, U, V, W, Q, qmask, rhs, phi, dt2_rec, rhs_out, ilog2(pp.ny), pm);
    };
    switch (mode) {
      case kRhsResid: go(std::integral_constant<int, kRhsResid>{}); break;
      case kRhsIn: go(std::integral_constant<int, kRhsIn>{}); break;
      case kRhsQ: go(std::integral_constant<int, kRhsQ>{}); break;
      case kRhsOut: go(std::integral_constant<int, kRhsOut>{}); break;
      case kRhsQ | kRhsOut: go(std::integral_constant<int, kRhsQ | kRhsOut>{}); break;
      case kRhsFast: go(std::integral_constant<int, kRhsFast>{}); break;
      case kRhsQ | kRhsFast: go(std::integral_constant<int, kRhsQ | kRhsFast>{}); break;
      case kRhsOut | kRhsFast: go(std::integral_constant<int, kRhsOut | kRhsFast>{}); break;
      case kRhsQ | kRhsOut | kRhsFast: go(std::integral_constant<int, kRhsQ | kRhsOut | kRhsFast>{}); break;
      default: go(std::integral_constant<int, kRhsDiv>{}); break;
    }
  });
}

// Modes per block of the y pass: 8 (128-byte FP64 row segments; 4 measured
// 30 % slower, 16 slower still: its LDS tile halves the blocks resident per
// CU).
constexpr int kFftYModes = 8;

template <typename R>
void launch_fft_y(const PoissonPlan<R> &pp, int inverse, hipStream_t st) {
  if (!fft_pow2(pp.ny)) {
    launch_fft_y_mr<R>(pp, inverse, st);
    return;
  }
  const int pm = kPadY;
  with_logn(ilog2(pp.ny), [&](auto L) {
    constexpr int LOGN = decltype(L)::value, TX = kFftYModes;
    const size_t sh = ((size_t)TX * ystride(1 << LOGN, pm) + twlen(1 << LOGN)) * sizeof(C2<R>);
    allow_lds((const void *)k_fft_y<TX, R, LOGN>, sh);
    hipLaunchKernelGGL((k_fft_y<TX, R, LOGN>), dim3(nblk(pp.M, TX), pp.nz), dim3(kFftThreads), sh, st, pp, inverse,
                       pm);
  });
}

template <typename R>
void launch_poisson_solve_z(const Geo &, const PoissonPlan<R> &pp, hipStream_t st) {
  // nz <= kSolveZMaxNz is checked when the plan is set up
  const long modes = (long)pp.Mp * pp.ny;
  const unsigned blocks = (unsigned)((modes + 63) / 64);
  const size_t sh = solve_z_lds(pp.nz);
  // one wavefront per block: more than 1024 fill several per CU
  if (pp.zpipe == 1 || (pp.zpipe < 0 && blocks > 1024)) {
    allow_lds((const void *)k_solve_z<R, true>, sh);
    hipLaunchKernelGGL((k_solve_z<R, true>), dim3(blocks), dim3(64), sh, st, pp);
  } else {
    allow_lds((const void *)k_solve_z<R, false>, sh);
    hipLaunchKernelGGL((k_solve_z<R, false>), dim3(blocks), dim3(64), sh, st, pp);
  }
}

template <typename R>
void launch_fft_x_inv(const Geo &g, const PoissonPlan<R> &pp, double *Phi, int accumulate, hipStream_t st) {
  if (!fft_pow2(pp.nx) || !fft_pow2(pp.ny)) {
    launch_fft_x_inv_mr<R>(g, pp, Phi, accumulate, st);
    return;
  }
  const int pm = kPadXInv, nrows = pp.ny * pp.nz;
  with_logn(ilog2(pp.nx), [&](auto L) {
    constexpr int LOGN = decltype(L)::value, ROWS = xrows(LOGN);
    const size_t sh = ((size_t)ROWS * fspan(1 << LOGN, pm) + twlen(1 << LOGN)) * sizeof(C2<R>);
    allow_lds((const void *)k_fft_x_inv<R, LOGN, ROWS>, sh);
    hipLaunchKernelGGL((k_fft_x_inv<R, LOGN, ROWS>), dim3(nblk(nrows, ROWS)), dim3(xthreads(LOGN, ROWS)), sh, st, g,
                       pp, Phi, accumulate, ilog2(pp.ny), pm);
  });
}

// ------------------------------------------------------------ reductions
// Block partial sums/maxima, combined by the consumer in block order so the
// result does not depend on scheduling.
__device__ inline double block_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += red[w];
  return s;
}
// NaN inputs count as +inf, so a maximum over a field holding a NaN is not
// finite (fmax alone would drop the NaN).
__device__ inline double block_max(double v) {
  if (v != v) v = INFINITY;
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  __shared__ double red[16];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s = fmax(s, red[w]);
  return s;
}

// Sum over the top layer of (Pr + Phi): FixPressureToReference with the
// default reference boundary To (src/pressure.f90:994-1229).
template <typename T>
__global__ void k_top_sum(Geo g, const T *__restrict__ Pr, const double *__restrict__ Phi,
                          double *__restrict__ partial) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (t < (long)g.nx * g.ny) {
    const int i = (int)(t % g.nx) + 1, j = (int)(t / g.nx) + 1;
    const long p = gix(g, i, j, g.nz);
    v = (double)Pr[p] + Phi[p];
  }
  const double s = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Block sums combined in block order by one thread: partial[nblocks] = mean.
__global__ void k_top_final(Geo g, double *__restrict__ partial, int nblocks) {
  extern __shared__ double buf[];
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) buf[b] = partial[b];
  __syncthreads();
  if (threadIdx.x != 0) return;
  // in block order; eight loads issued ahead of their additions
  double s = 0;
  int b = 0;
  for (; b + 8 <= nblocks; b += 8) {
    double t[8];
#pragma unroll
    for (int q = 0; q < 8; q++) t[q] = buf[b + q];
#pragma unroll
    for (int q = 0; q < 8; q++) s += t[q];
  }
  for (; b < nblocks; b++) s += buf[b];
  partial[nblocks] = s / (double)(g.nx * g.ny);
}

template <typename T>
void launch_top_sum(const Geo &g, const T *Pr, const double *Phi, double *partial, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_top_sum<T>, dim3(nblocks), dim3(256), 0, st, g, Pr, Phi, partial);
  hipLaunchKernelGGL(k_top_final, dim3(1), dim3(256), (size_t)nblocks * sizeof(double), st, g, partial, nblocks);
}

// PostPoisson (2nd order) + Pr = Pr + Phi - ref:
// U -= coef*(Phi(i+1)-Phi(i))/dx etc. (src/pressure.f90:789-853).  x and y
// are periodic, so Phi(nx+1) = Phi(1) is read by index wrap (Bound_Phi).
// The correction is evaluated in FP64 and rounded to the field type T.
template <typename T>
__global__ void __launch_bounds__(256)
    k_projection(Geo g, T *__restrict__ U, T *__restrict__ V, T *__restrict__ W, T *__restrict__ Pr,
                 const double *__restrict__ Phi, const double *__restrict__ partial, int nblocks, double dt2) {
  const double ref = partial[nblocks];
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long p = gix(g, i, j, k);
  // every operand is loaded unconditionally (all in the allocation: Phi(k+1)
  // at k = nz is its ghost level), only the stores depend on the component
  // extents, so the loads of a thread issue together
  const double ph = Phi[p];
  const double phx = Phi[i < g.nx ? p + 1 : p + 1 - g.nx];
  const double phy = Phi[j < g.ny ? p + g.sy : p + g.sy - (long)g.ny * g.sy];
  const double phz = Phi[p + g.sz];
  const T u = U[p], v = V[p], w = W[p], pr0 = Pr[p];
  const double Au = dt2 / g.dxmin, Av = dt2 / g.dymin, Aw = dt2 / g.dzmin;
  if (i <= g.n[0][0] && j <= g.n[0][1] && k <= g.n[0][2]) U[p] = (T)((double)u - Au * (phx - ph));
  if (i <= g.n[1][0] && j <= g.n[1][1] && k <= g.n[1][2]) V[p] = (T)((double)v - Av * (phy - ph));
  if (i <= g.n[2][0] && j <= g.n[2][1] && k <= g.n[2][2]) {
    if (g.variable_z)
      W[p] = (T)((double)w - dt2 * (phz - ph) / g.dzW[k]);
    else
      W[p] = (T)((double)w - Aw * (phz - ph));
  }
  double pr = (double)pr0 + ph;
  if (ref != 0) pr = pr - ref;
  Pr[p] = (T)pr;
}

template <typename T>
void launch_projection(const Geo &g, T *U, T *V, T *W, T *Pr, const double *Phi, const double *partial, int nblocks,
                       double dt2, hipStream_t st) {
  dim3 b(64, 4, 1), gr(nblk(g.nx, 64), nblk(g.ny, 4), g.nz);
  hipLaunchKernelGGL(k_projection<T>, gr, b, 0, st, g, U, V, W, Pr, Phi, partial, nblocks, dt2);
}

// max |div U - Q| (check_divergence, src/pressure.f90:937-981)
template <typename T>
__global__ void k_max_div(Geo g, const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
                          const double *__restrict__ Q, double *__restrict__ partial) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (t < (long)g.nx * g.ny * g.nz) {
    const int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long)g.nx * g.ny)) + 1;
    const long p = gix(g, i, j, k);
    const double dz = g.variable_z ? g.dzPr[k] : g.dzmin;
    double d = ((double)U[p] - (double)U[p - 1]) / (g.dxmin) + ((double)V[p] - (double)V[p - g.sy]) / (g.dymin) +
               ((double)W[p] - (double)W[p - g.sz]) / (dz);
    if (Q) d = d - Q[p];
    v = fabs(d);
  }
  const double m = block_max(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

template <typename T>
void launch_max_div(const Geo &g, const T *U, const T *V, const T *W, const double *Q, double *partial, int nblocks,
                    hipStream_t st) {
  hipLaunchKernelGGL(k_max_div<T>, dim3(nblocks), dim3(256), 0, st, g, U, V, W, Q, partial);
}

// Maximum of n block maxima (a NaN counts as +inf) into out[0]; one block.
__global__ void __launch_bounds__(256) k_max_final(const double *__restrict__ partial, int n, double *__restrict__ out) {
  double v = 0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) {
    const double x = partial[b];
    v = x != x ? INFINITY : fmax(v, x);
  }
  const double m = block_max(v);
  if (threadIdx.x == 0) out[0] = m;
}

void launch_max_final(const double *partial, int n, double *out, hipStream_t st) {
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, st, partial, n, out);
}

// Residue_variable_z, src/poisson.f90:23-60 (Phi ghosts: periodic x/y by
// index wrap, mirror in z)
__global__ void k_residual(Geo g, const double *__restrict__ Phi, const double *__restrict__ RHS,
                           double *__restrict__ partial) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (t < (long)g.nx * g.ny * g.nz) {
    const int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long)g.nx * g.ny)) + 1;
    const long p = gix(g, i, j, k);
    const double Aw = 1 / (g.dxmin * g.dxmin), Ae = Aw, As = 1 / (g.dymin * g.dymin), An = As;
    const double Ab = 1 / ((g.zPr[k] - g.zPr[k - 1]) * (g.zW[k] - g.zW[k - 1]));
    const double At = 1 / ((g.zPr[k + 1] - g.zPr[k]) * (g.zW[k] - g.zW[k - 1]));
    const double Ap = Aw + Ae + As + An + Ab + At;
    const double pw = Phi[i > 1 ? p - 1 : p - 1 + g.nx], pe = Phi[i < g.nx ? p + 1 : p + 1 - g.nx];
    const double ps = Phi[j > 1 ? p - g.sy : p - g.sy + (long)g.ny * g.sy];
    const double pn = Phi[j < g.ny ? p + g.sy : p + g.sy - (long)g.ny * g.sy];
    const double pb = Phi[k > 1 ? p - g.sz : p], pt = Phi[k < g.nz ? p + g.sz : p];
    double q = 0;
    q = q + pw * Aw;
    q = q + pe * Ae;
    q = q + ps * As;
    q = q + pn * An;
    q = q + pb * Ab;
    q = q + pt * At;
    q = q - RHS[p];
    v = fabs(-q + Ap * Phi[p]);
  }
  const double m = block_max(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

void launch_residual(const Geo &g, const double *Phi, const double *RHS, double *partial, int nblocks,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_residual, dim3(nblocks), dim3(256), 0, st, g, Phi, RHS, partial);
}

// Maximum FP64 residual of the solved operator (diagnostics of the
// precision form): block maxima of |r| go to partial, r to r_out if given.
__global__ void k_refine_residual(Geo g, const double *__restrict__ Phi, const double *__restrict__ RHS,
                                  const double *__restrict__ az, const double *__restrict__ cz,
                                  double *__restrict__ r_out, double *__restrict__ partial) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (t < (long)g.nx * g.ny * g.nz) {
    const int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long)g.nx * g.ny)) + 1;
    const long p = gix(g, i, j, k);
    const double r = fd2_residual(g, Phi, RHS, az, cz, i, j, k, p);
    if (r_out) r_out[p] = r;
    v = fabs(r);
  }
  const double m = block_max(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

void launch_refine_residual(const Geo &g, const double *Phi, const double *RHS, const double *az, const double *cz,
                            double *r_out, double *partial, int nblocks, hipStream_t st) {
  hipLaunchKernelGGL(k_refine_residual, dim3(nblocks), dim3(256), 0, st, g, Phi, RHS, az, cz, r_out, partial);
}

#define ELMM_INST_PRESSURE_R(R)                                                                                    \
  template void launch_fft_y<R>(const PoissonPlan<R> &, int, hipStream_t);                                        \
  template void launch_poisson_solve_z<R>(const Geo &, const PoissonPlan<R> &, hipStream_t);                       \
  template void launch_fft_x_inv<R>(const Geo &, const PoissonPlan<R> &, double *, int, hipStream_t);
#define ELMM_INST_PRESSURE_T(T)                                                                                    \
  template void launch_top_sum<T>(const Geo &, const T *, const double *, double *, int, hipStream_t);             \
  template void launch_projection<T>(const Geo &, T *, T *, T *, T *, const double *, const double *, int, double,  \
                                     hipStream_t);                                                                 \
  template void launch_max_div<T>(const Geo &, const T *, const T *, const T *, const double *, double *, int,      \
                                  hipStream_t);
ELMM_INST_PRESSURE_R(double)
ELMM_INST_PRESSURE_R(float)
ELMM_INST_PRESSURE_T(double)
ELMM_INST_PRESSURE_T(float)
#define ELMM_INST_XFWD(T, R)                                                                                       \
  template void launch_fft_x_fwd<T, R>(const Geo &, const PoissonPlan<R> &, const T *, const T *, const T *,        \
                                       const double *, const uint32_t *, const double *, const double *, double,     \
                                       double *, hipStream_t);
ELMM_INST_XFWD(double, double)
ELMM_INST_XFWD(float, double)
ELMM_INST_XFWD(float, float)

}  // namespace elmm
