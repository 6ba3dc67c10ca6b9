// k_momentum.hip - subgrid viscosity, wall-model fluxes, fused momentum
// right-hand side + RK3 combine + pressure gradient, ordered wall-model
// corrections, immersed-boundary gathers.
//
// Fusion (one HBM pass per component instead of the reference's separate
// passes): Convection's advection (CD4divU, src/advection.f90:334-491),
// explicit diffusion (MomentumDiffusion_nobranch_2ord,
// src/momentum_diffusion.f90:169-288), the RK combine (src/dynamics.f90:
// 266-295, 450-478), PressureGrad (src/dynamics.f90:58-199) and
// add(U2,U) (src/tsteps.f90:350-355).  The arithmetic of each value is the
// reference's, operation by operation (built with -ffp-contract=off).
#include <algorithm>
#include <cmath>

#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------- Smagorinsky
// NuSmag/StrainIJ/Strainu, src/subgrid.f90:55-67, 170-209.  T is the field
// and arithmetic type (double: the reference's FP64; float: precision form).
// Uf(di, dj, dk) etc. return the field at (i+di, j+dj, k+dk); both SGS
// kernels evaluate the same expressions in the same order.
// RCP (FP64): the quotients as correctly rounded products with the host's
// reciprocals (div_rn_nz), equal to the divisions.
// FP32 level reciprocals of the strain divisors: RN(1 / dzPr(k)) and
// RN(1 / (2 (zPr(k+1) - zPr(k-1)))) in single precision (a table per level
// in the z-march, the same values computed in place by the per-cell kernel).
__device__ inline float sgs_rdzf(const Geo &g, int k) { return 1.f / (float)g.dzPr[k]; }
__device__ inline float sgs_rz2f(const Geo &g, int k) {
  return 1.f / (2 * ((float)g.zPr[k + 1] - (float)g.zPr[k - 1]));
}

template <typename T, bool RCP = false, typename FU, typename FV, typename FW>
__device__ inline T smag_visc(const Geo &g, FU Uf, FV Vf, FW Wf, int k, T width, T nu, T c_smag,
                              const SgsWidth &wr, float rdzk = 0.f, float rz2k = 0.f) {
  const T dx = (T)g.dxmin, dy = (T)g.dymin;
  if constexpr (sizeof(T) == 4) {
    // precision form: products with reciprocals (the level's two from
    // rdzk, rz2k) and the hardware square root; within a few ulp of the
    // quotient form
    const float rdx = g.rdxf != 0.f ? g.rdxf : 1.f / dx, rdy = g.rdyf != 0.f ? g.rdyf : 1.f / dy;
    const float rdz = rdzk, rz2 = rz2k;
    const float r4x = 0.25f * rdx, r4y = 0.25f * rdy;
    float D00 = (Uf(0, 0, 0) - Uf(-1, 0, 0)) * rdx;
    float D11 = (Vf(0, 0, 0) - Vf(0, -1, 0)) * rdy;
    float D22 = (Wf(0, 0, 0) - Wf(0, 0, -1)) * rdz;
    float D01 = (Uf(0, 1, 0) + Uf(-1, 1, 0) - Uf(0, -1, 0) - Uf(-1, -1, 0)) * r4y;
    float D02 = (Uf(0, 0, 1) + Uf(-1, 0, 1) - Uf(0, 0, -1) - Uf(-1, 0, -1)) * rz2;
    float D10 = (Vf(1, 0, 0) + Vf(1, -1, 0) - Vf(-1, 0, 0) - Vf(-1, -1, 0)) * r4x;
    float D12 = (Vf(0, 0, 1) + Vf(0, -1, 1) - Vf(0, 0, -1) - Vf(0, -1, -1)) * rz2;
    float D20 = (Wf(1, 0, 0) + Wf(1, 0, -1) - Wf(-1, 0, 0) - Wf(-1, 0, -1)) * r4x;
    float D21 = (Wf(0, 1, 0) + Wf(0, 1, -1) - Wf(0, -1, 0) - Wf(0, -1, -1)) * r4y;
    const float S01 = (D01 + D10) * 0.5f, S02 = (D02 + D20) * 0.5f, S12 = (D12 + D21) * 0.5f;
    const float s = D00 * D00 + D11 * D11 + D22 * D22 + 2.f * (S01 * S01 + S02 * S02 + S12 * S12);
    const float Sbar = __builtin_amdgcn_sqrtf(2.f * s);
    const float wc = width * c_smag;
    float v = Sbar * (wc * wc);
    if (nu > 0) v = v + nu;
    return v;
  }
  const T zp1 = (T)g.zPr[k + 1], zm1 = (T)g.zPr[k - 1];
  // (4 d)'s reciprocal is RN(1/d) / 4 exactly
  auto q = [&](T x, T d, double r) {
    if constexpr (RCP)
      return div_rn_nz<T>(x, d, (T)r);
    else
      return x / d;
  };
  T D00 = q(Uf(0, 0, 0) - Uf(-1, 0, 0), dx, wr.rdx);
  T D11 = q(Vf(0, 0, 0) - Vf(0, -1, 0), dy, wr.rdy);
  T D22 = q(Wf(0, 0, 0) - Wf(0, 0, -1), (T)g.dzPr[k], RCP ? wr.rdz[k] : 0.0);
  T D01 = q(Uf(0, 1, 0) + Uf(-1, 1, 0) - Uf(0, -1, 0) - Uf(-1, -1, 0), 4 * dy, 0.25 * wr.rdy);
  T D02 = q(Uf(0, 0, 1) + Uf(-1, 0, 1) - Uf(0, 0, -1) - Uf(-1, 0, -1), 2 * (zp1 - zm1), RCP ? wr.rz2[k] : 0.0);
  T D10 = q(Vf(1, 0, 0) + Vf(1, -1, 0) - Vf(-1, 0, 0) - Vf(-1, -1, 0), 4 * dx, 0.25 * wr.rdx);
  T D12 = q(Vf(0, 0, 1) + Vf(0, -1, 1) - Vf(0, 0, -1) - Vf(0, -1, -1), 2 * (zp1 - zm1), RCP ? wr.rz2[k] : 0.0);
  T D20 = q(Wf(1, 0, 0) + Wf(1, 0, -1) - Wf(-1, 0, 0) - Wf(-1, 0, -1), 4 * dx, 0.25 * wr.rdx);
  T D21 = q(Wf(0, 1, 0) + Wf(0, 1, -1) - Wf(0, -1, 0) - Wf(0, -1, -1), 4 * dy, 0.25 * wr.rdy);
  const T D[3][3] = {{D00, D01, D02}, {D10, D11, D12}, {D20, D21, D22}};
  T s = 0;
#pragma unroll
  for (int jj = 0; jj < 3; jj++)
#pragma unroll
    for (int ii = 0; ii < 3; ii++) {
      T Sij = (D[ii][jj] + D[jj][ii]) / 2;
      s = s + Sij * Sij;
    }
  T Sbar = sqrt((T)2.0 * s);
  T wc = width * c_smag;
  T v = Sbar * (wc * wc);
  if (nu > 0) v = v + nu;
  return v;
}

// Filter width of column (i, j) at level k: table from the host libm (the
// cube root of a static geometric factor per column class and level).
// TabOnly: the table is known to be present (no pow code in the kernel).
// column class of the width table
__device__ inline int smag_class(const Geo &g, const SgsWidth &w, int i, int j) {
  return w.idx ? w.idx[(i - 1) + (long)g.nx * (j - 1)] : 0;
}
template <typename T, bool TabOnly = false>
__device__ inline T smag_width(const Geo &g, const SgsWidth &w, double filter_ratio, int i, int j, int k) {
  if (TabOnly || w.tab) {
    const int cls = w.idx ? w.idx[(i - 1) + (long)g.nx * (j - 1)] : 0;
    return (T)w.tab[cls * w.ld + k];
  }
  return (T)(filter_ratio * pow(g.dxPr[i] * g.dyPr[j] * g.dzPr[k], 1.0 / 3.0));
}

// TimeStepLength's U_dx_max term of one cell, src/dynamics.f90:556-594
// (NaN -> +inf); u1 = U(i), u0 = U(i-1), v1/v0 over j, w1/w0 over k.
__device__ inline double cfl_cell(const Geo &g, double u1, double u0, double v1, double v0, double w1, double w0,
                                  int i, int j, int k) {
  double a1 = fabs(u1), a0 = fabs(u0);
  double b1 = fabs(v1), b0 = fabs(v0);
  double c1 = fabs(w1), c0 = fabs(w0);
  double q;
  if (g.variable_z) {
    q = (a1 > a0 ? a1 : a0) / g.dxPr[i];
    q = q + (b1 > b0 ? b1 : b0) / g.dyPr[j];
    q = q + (c1 > c0 ? c1 : c0) / g.dzPr[k];
  } else {
    // uniform spacing: correctly rounded reciprocal products (the same
    // quotients as the divisions) where the host's reciprocals qualify
    q = div_rn(a1 > a0 ? a1 : a0, g.dxmin, g.rdx);
    q = q + div_rn(b1 > b0 ? b1 : b0, g.dymin, g.rdy);
    q = q + div_rn(c1 > c0 ? c1 : c0, g.dzmin, g.rdz);
  }
  return (q != q) ? INFINITY : q;
}

// Block maximum of v (blockDim.x a multiple of 64, at most 1024 threads);
// thread 0 holds the result.
__device__ inline double block_max_cfl(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  __shared__ double red[16];
  const int lane = threadIdx.x + blockDim.x * threadIdx.y;
  if ((lane & 63) == 0) red[lane >> 6] = v;
  __syncthreads();
  double m = 0;
  if (lane == 0)
    for (int w = 0; w < (int)(blockDim.x * blockDim.y) / 64; w++) m = fmax(m, red[w]);
  return m;
}

// One thread per cell (the stretched-z path and ELMM_SGS=cellwise).
template <typename T>
__global__ void __launch_bounds__(256) k_sgs(Geo g, const T *__restrict__ U, const T *__restrict__ V,
                                             const T *__restrict__ W, T *__restrict__ Visc, double nu_,
                                             double c_smag_, double filter_ratio, SgsWidth width_t,
                                             int sgs_model, double *__restrict__ cfl_partial) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const T nu = (T)nu_, c_smag = (T)c_smag_;
  double cfl = 0;
  if (i <= g.nx && j <= g.ny) {
    const long p = gix(g, i, j, k);
    const long sy = g.sy, sz = g.sz;
    auto Uf = [&](int di, int dj, int dk) { return U[p + di + dj * sy + dk * sz]; };
    auto Vf = [&](int di, int dj, int dk) { return V[p + di + dj * sy + dk * sz]; };
    auto Wf = [&](int di, int dj, int dk) { return W[p + di + dj * sy + dk * sz]; };
    Visc[p] = sgs_model == 1 ? smag_visc<T>(g, Uf, Vf, Wf, k, smag_width<T>(g, width_t, filter_ratio, i, j, k), nu,
                                            c_smag, width_t, sizeof(T) == 4 ? sgs_rdzf(g, k) : 0.f,
                                            sizeof(T) == 4 ? sgs_rz2f(g, k) : 0.f)
                             : nu;
    if (cfl_partial) cfl = cfl_cell(g, U[p], U[p - 1], V[p], V[p - sy], W[p], W[p - sz], i, j, k);
  }
  if (cfl_partial) {
    const double m = block_max_cfl(cfl);
    if (threadIdx.x == 0 && threadIdx.y == 0)
      cfl_partial[(long)blockIdx.z * gridDim.y * gridDim.x + blockIdx.y * gridDim.x + blockIdx.x] = m;
  }
}

// z-marching form: a block owns a TX x TY tile of columns and walks k up a
// chunk of levels.  LDS holds rings of (TX+2) x (TY+2) planes (halo 1) of U,
// V and W, R levels each, level L in slot (L - kbeg + 1) mod R; the z loop is
// unrolled by R so that every slot is a constant.  The planes of level k+2
// (U, V) and k+1 (W) are loaded into registers while level k is evaluated
// and then stored into their slots: with R = 4 (4-byte fields) no thread
// reads those slots at level k and one barrier per level suffices; with
// R = 3 (8-byte fields, whose LDS then keeps three blocks per CU) the stores
// wait for a second barrier.  Global loads per cell: one plane of U, V and W
// plus the halo ring, instead of 21 stencil values.
// RPT: rows per thread (the tile is TX x TY*RPT columns).  Launch bounds: 4
// waves per SIMD for 4-byte fields (128 VGPRs: at 5 waves it spills and runs
// 181 instead of 166 us on 512x512x128), 3 for 8-byte fields.
template <typename T, int TX, int TY, int RPT, bool RCP>
__global__ void __launch_bounds__(TX *TY, sizeof(T) == 4 ? 4 : 3)
    k_sgs_zm(Geo g, const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W, T *__restrict__ Visc,
             double nu_, double c_smag_, double filter_ratio, SgsWidth width_t, int sgs_model,
             double *__restrict__ cfl_partial, int zchunk) {
  constexpr int NT = TX * TY, RX = TX + 2, RC = RX * (TY * RPT + 2), PER = (RC + NT - 1) / NT;
  constexpr int R = sizeof(T) == 4 ? 4 : 3;
  constexpr int SU = 0, SV = R, SW = 2 * R;
  __shared__ T ring[3 * R * RC];
  const int t = threadIdx.x, tx = t % TX, ty = t / TX;
  const int i0 = blockIdx.x * TX + 1, j0 = blockIdx.y * TY * RPT + 1;
  const int kbeg = blockIdx.z * zchunk + 1, kend = min(g.nz, kbeg + zchunk - 1);
  const int i = i0 + tx;
  const T nu = (T)nu_, c_smag = (T)c_smag_;
  // region cells of this thread (halo 1 around the tile), clamped into the
  // allocation (values of clamped cells only reach cells outside the
  // domain); a thread past the region loads and stores the region's last
  // cell again, the value its owner stores
  long roff[PER];
  int ridx[PER];
#pragma unroll
  for (int q = 0; q < PER; q++) {
    const int r = min(t + q * NT, RC - 1);
    ridx[q] = r;
    const int ri = min(i0 - 1 + r % RX, g.nx + 3), rj = min(j0 - 1 + r / RX, g.ny + 3);
    roff[q] = gix(g, ri, rj, 0);
  }
  auto slot = [](int base, int rel) { return base + (rel + 2 * R) % R; };
  auto put = [&](int s, const T *src, int k) {
#pragma unroll
    for (int q = 0; q < PER; q++) ring[s * RC + ridx[q]] = src[roff[q] + (long)k * g.sz];
  };
  put(slot(SU, -1), U, kbeg - 1);
  put(slot(SU, 0), U, kbeg);
  put(slot(SU, 1), U, kbeg + 1);
  put(slot(SV, -1), V, kbeg - 1);
  put(slot(SV, 0), V, kbeg);
  put(slot(SV, 1), V, kbeg + 1);
  put(slot(SW, -1), W, kbeg - 1);
  put(slot(SW, 0), W, kbeg);
  // filter-width class of the thread's columns (constant over k)
  int cls[RPT];
#pragma unroll
  for (int rr = 0; rr < RPT; rr++) {
    const int j = min(j0 + ty + rr * TY, g.ny);
    cls[rr] = smag_class(g, width_t, min(i, g.nx), j);
  }
  __syncthreads();
  double cfl = 0;
  for (int k0 = kbeg; k0 <= kend; k0 += R) {
#pragma unroll
    for (int u = 0; u < R; u++) {
      const int k = k0 + u;
      if (k > kend) break;
      // next planes (levels inside the allocation; unused after the chunk)
      T pu[PER], pv[PER], pw[PER];
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const long o = roff[q] + (long)(k + 2) * g.sz;
        pu[q] = U[o];
        pv[q] = V[o];
        pw[q] = W[o - g.sz];
      }
      const float rdzk = sizeof(T) == 4 ? width_t.rdzf[k] : 0.f, rz2k = sizeof(T) == 4 ? width_t.rz2f[k] : 0.f;
      const T *Ul = ring + slot(SU, u - 1) * RC, *Uc = ring + slot(SU, u) * RC, *Uh = ring + slot(SU, u + 1) * RC;
      const T *Vl = ring + slot(SV, u - 1) * RC, *Vc = ring + slot(SV, u) * RC, *Vh = ring + slot(SV, u + 1) * RC;
      const T *Wl = ring + slot(SW, u - 1) * RC, *Wc = ring + slot(SW, u) * RC;
#pragma unroll
      for (int rr = 0; rr < RPT; rr++) {
        const int j = j0 + ty + rr * TY;
        if (i <= g.nx && j <= g.ny) {
          const int c = (tx + 1) + RX * (ty + rr * TY + 1);
          auto Uf = [&](int di, int dj, int dk) { return (dk < 0 ? Ul : dk > 0 ? Uh : Uc)[c + di + RX * dj]; };
          auto Vf = [&](int di, int dj, int dk) { return (dk < 0 ? Vl : dk > 0 ? Vh : Vc)[c + di + RX * dj]; };
          auto Wf = [&](int di, int dj, int dk) { return (dk < 0 ? Wl : Wc)[c + di + RX * dj]; };
          Visc[gix(g, i, j, k)] =
              sgs_model == 1 ? smag_visc<T, RCP>(g, Uf, Vf, Wf, k, (T)width_t.tab[cls[rr] * width_t.ld + k], nu,
                                                 c_smag, width_t, rdzk, rz2k)
                             : nu;
          if (cfl_partial)
            cfl = fmax(cfl, cfl_cell(g, Uc[c], Uc[c - 1], Vc[c], Vc[c - RX], Wc[c], Wl[c], i, j, k));
        }
      }
      if constexpr (R == 3) __syncthreads();
#pragma unroll
      for (int q = 0; q < PER; q++) {
        ring[slot(SU, u + 2) * RC + ridx[q]] = pu[q];
        ring[slot(SV, u + 2) * RC + ridx[q]] = pv[q];
        ring[slot(SW, u + 1) * RC + ridx[q]] = pw[q];
      }
      __syncthreads();
    }
  }
  if (cfl_partial) {
    const double m = block_max_cfl(cfl);
    if (t == 0) cfl_partial[((long)blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = m;
  }
}

// 64 x 4 threads, two rows each (a 64 x 8 tile: 2 % faster than one row per
// thread; 32 x 8 and 64 x 8 threads measured slower), 32 levels per block (16
// measured equal, 64 12 % slower in FP32: fewer blocks to hide the per-level
// load latency)
// z chunk of 26 levels (measured 237 -> 222 us against 32 levels with the
// earlier 5-blocks-per-CU form; ELMM_SGS_ZCHUNK overrides)
constexpr int kSgsTX = 64, kSgsTY = 4, kSgsRPT = 2, kSgsZChunk = 26;
static int sgs_zchunk(const Geo &g) { return g.sgs_zchunk > 0 ? g.sgs_zchunk : kSgsZChunk; }

static bool sgs_cellwise(const Geo &g, const SgsWidth &w) { return g.sgs_cellwise || g.variable_z || !w.tab; }

int sgs_partials(const Geo &g, const SgsWidth &w) {
  if (sgs_cellwise(g, w)) return (int)(nblk(g.nx, 64) * nblk(g.ny, 4) * g.nz);
  return (int)(nblk(g.nx, kSgsTX) * nblk(g.ny, kSgsTY * kSgsRPT) * nblk(g.nz, sgs_zchunk(g)));
}

template <typename T, int RPT>
static void launch_sgs_zm(const Geo &g, const T *U, const T *V, const T *W, T *Visc, double nu, double c_smag,
                          double filter_ratio, SgsWidth width, int sgs_model, double *cfl_partial, hipStream_t st) {
  dim3 gr(nblk(g.nx, kSgsTX), nblk(g.ny, kSgsTY * RPT), nblk(g.nz, sgs_zchunk(g)));
  if (sizeof(T) == 8 && width.rcp())
    hipLaunchKernelGGL((k_sgs_zm<T, kSgsTX, kSgsTY, RPT, true>), gr, dim3(kSgsTX * kSgsTY), 0, st, g, U, V, W, Visc,
                       nu, c_smag, filter_ratio, width, sgs_model, cfl_partial, sgs_zchunk(g));
  else
    hipLaunchKernelGGL((k_sgs_zm<T, kSgsTX, kSgsTY, RPT, false>), gr, dim3(kSgsTX * kSgsTY), 0, st, g, U, V, W, Visc,
                       nu, c_smag, filter_ratio, width, sgs_model, cfl_partial, sgs_zchunk(g));
}

template <typename T>
void launch_sgs(const Geo &g, const T *U, const T *V, const T *W, T *Visc, double nu, double c_smag,
                double filter_ratio, SgsWidth width, int sgs_model, double *cfl_partial, int, hipStream_t st) {
  if (sgs_cellwise(g, width)) {
    dim3 b(64, 4, 1), gr(nblk(g.nx, 64), nblk(g.ny, 4), g.nz);
    hipLaunchKernelGGL(k_sgs<T>, gr, b, 0, st, g, U, V, W, Visc, nu, c_smag, filter_ratio, width, sgs_model,
                       cfl_partial);
    return;
  }
  launch_sgs_zm<T, kSgsRPT>(g, U, V, W, Visc, nu, c_smag, filter_ratio, width, sgs_model, cfl_partial, st);
}

// ------------------------------------------------------------- wall model
// vel_and_dist, WMRoughUstar, WMFlatUstar, fluxes:
// src/wallmodels.f90:1190-1240, 1312-1391, 2159-2337, 2356-2399
// Natural logarithm of a positive normal x with the same operations as the
// oracle (so CPU and GPU agree bit for bit; the device and host libm logs
// may differ by an ulp): x = 2^k m, m in [sqrt(2)/2, sqrt(2)), f = m - 1,
// s = f / (2 + f), log(1 + f) = 2 atanh(s) = f - (hfsq - s (hfsq + R(s^2))),
// R(z) = sum_{n=1..11} 2 z^n / (2n + 1); log x = k ln2_hi - ((hfsq -
// (s (hfsq + R) + k ln2_lo)) - f), k ln2_hi exact (31-bit ln2_hi).  The
// reduction of P. T. P. Tang (ACM TOMS 16, 1990), Taylor polynomial.
__device__ inline double det_log(double x) {
  unsigned long long b = (unsigned long long)__double_as_longlong(x);
  int k = (int)((b >> 52) & 0x7ff) - 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m = __longlong_as_double((long long)b);
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    k = k + 1;
  }
  const double f = m - 1.0;
  const double s = f / (2.0 + f);
  const double z = s * s;
  double R = 2.0 / 23.0;
  R = 2.0 / 21.0 + z * R;
  R = 2.0 / 19.0 + z * R;
  R = 2.0 / 17.0 + z * R;
  R = 2.0 / 15.0 + z * R;
  R = 2.0 / 13.0 + z * R;
  R = 2.0 / 11.0 + z * R;
  R = 2.0 / 9.0 + z * R;
  R = 2.0 / 7.0 + z * R;
  R = 2.0 / 5.0 + z * R;
  R = 2.0 / 3.0 + z * R;
  R = z * R;
  const double hfsq = 0.5 * f * f;
  const double dk = (double)k;
  const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
  return dk * ln2_hi - ((hfsq - (s * (hfsq + R) + dk * ln2_lo)) - f);
}

__device__ inline double wm_flat_ustar(double ustar, double vel, double dist, double nu) {
  const double eps = 1e-4;
  if (nu <= 0) return 0;
  double ustar_lam = sqrt(nu * vel / dist);
  if ((dist * ustar_lam / nu) < 5.0) return ustar_lam;
  for (int i = 1;; i++) {
    double ustar2 = ustar;
    double z_pl = dist * ustar2 / nu;
    if (z_pl <= 5.0)
      ustar = sqrt(nu * vel / dist);
    else if (z_pl < 30.0)
      ustar = vel / (5 * det_log(z_pl) - (double)3.05f);
    else
      ustar = vel / (det_log(z_pl) / 0.4 + 5.5);
    if (fabs(ustar - ustar2) / fabs(ustar) < eps) break;
    if (i >= 50) {
      ustar = 0;
      break;
    }
  }
  return ustar;
}

__device__ inline double wm_rough_ustar(double ustar, double vel, double dist, double z0, double nu,
                                       double logdz0) {
  const double yplcrit = 11.225;
  if (nu > 0) {
    ustar = sqrt(nu * vel / dist);
    if (dist * ustar / nu < yplcrit * 1.1) return ustar;
  }
  if (dist <= z0 * 2.0) {
    ustar = nu > 0 ? wm_flat_ustar(ustar, vel, dist, nu) : 0.0;
  } else {
    ustar = vel * 0.41 / logdz0;  // log(dist/z0), precomputed per point
  }
  return ustar;
}

// exp(x) with fixed operations, the same as the oracle's (host and device
// libm differ by an ulp): x = k ln2 + r, k = rint(x / ln2), r = (x - k
// ln2_hi) - k ln2_lo (ln2_hi of 33 significant bits), |r| <= ln2/2; exp(r)
// by its Taylor polynomial to r^13/13! in Horner form; times 2^k (exact).
// The reduction of W. J. Cody and W. Waite (1980).
__device__ inline double det_exp(double x) {
  const double inv_ln2 = 0x1.71547652b82fep+0;
  const double ln2_hi = 0x1.62e42fee00000p-1, ln2_lo = 0x1.a39ef35793c76p-33;
  const double k = rint(x * inv_ln2);
  const double r = (x - k * ln2_hi) - k * ln2_lo;
  double p = 1.0 / 6227020800.0;
  p = 1.0 / 479001600.0 + r * p;
  p = 1.0 / 39916800.0 + r * p;
  p = 1.0 / 3628800.0 + r * p;
  p = 1.0 / 362880.0 + r * p;
  p = 1.0 / 40320.0 + r * p;
  p = 1.0 / 5040.0 + r * p;
  p = 1.0 / 720.0 + r * p;
  p = 1.0 / 120.0 + r * p;
  p = 1.0 / 24.0 + r * p;
  p = 1.0 / 6.0 + r * p;
  p = 0.5 + r * p;
  p = 1.0 + r * p;
  p = 1.0 + r * p;
  return ldexp(p, (int)k);
}

// atan(x) with fixed operations (the oracle's): 1/a above 1, the pi/6 shift
// above tan(pi/12), then the Taylor series to t^31/31.
__device__ inline double det_atan(double x) {
  const double pi2 = 0x1.921fb54442d18p+0, pi6 = 0x1.0c152382d7365p-1;
  const double sqrt3 = 0x1.bb67ae8584caap+0, tan12 = 0x1.126145e9ecd58p-2;
  double a = fabs(x), off = 0;
  bool inv = false;
  if (a > 1) {
    a = 1 / a;
    inv = true;
  }
  if (a > tan12) {
    a = (a * sqrt3 - 1) / (a + sqrt3);
    off = pi6;
  }
  const double z = a * a;
  double p = -1.0 / 31.0;
  p = 1.0 / 29.0 + z * p;
  p = -1.0 / 27.0 + z * p;
  p = 1.0 / 25.0 + z * p;
  p = -1.0 / 23.0 + z * p;
  p = 1.0 / 21.0 + z * p;
  p = -1.0 / 19.0 + z * p;
  p = 1.0 / 17.0 + z * p;
  p = -1.0 / 15.0 + z * p;
  p = 1.0 / 13.0 + z * p;
  p = -1.0 / 11.0 + z * p;
  p = 1.0 / 9.0 + z * p;
  p = -1.0 / 7.0 + z * p;
  p = 1.0 / 5.0 + z * p;
  p = -1.0 / 3.0 + z * p;
  p = 1.0 + z * p;
  double r = off + a * p;
  if (inv) r = pi2 - r;
  return x < 0 ? -r : r;
}

// x**y (x >= 0, y > 0) as exp(y log x); 0 for x = 0
__device__ inline double det_pow(double x, double y) { return x == 0 ? 0.0 : det_exp(y * det_log(x)); }

// Monin-Obukhov laws, src/wallmodels.f90:1546-1716 (pi = acos(-1.0_knd),
// k_U = 0.4, k_T = 0.47; the flux law's eps = 1e-3, 1E-5, 1E-4 and the
// Dirichlet law's 1.1 are default-real literals)
__device__ inline double mo_psi_unstable(double x) {
  const double c = (1 + x) / 2.0;
  return (det_log(((1 + x * x) / 2.0) * (c * c)) - 2.0 * det_atan(x)) + 0x1.921fb54442d18p+1 / 2;
}

__device__ inline double psim_mo(double zeta) {
  return zeta < 0 ? mo_psi_unstable(det_pow(1 - 15.0 * zeta, 0.25)) : -(4.8 * zeta);
}

__device__ inline double psim_mo_mod(double zeta, double kU) {
  return zeta < 0 ? mo_psi_unstable(det_pow(1 - (6.0 * zeta) * kU, 0.25)) : -(3 * det_pow(zeta * kU, 5.0 / 6.0));
}

__device__ inline double psih_mo_mod(double zeta, double kU) {
  if (zeta < 0) {
    const double x = det_pow(1 - (6.0 * zeta) * kU, 0.25);
    return 2.0 * det_log((1 + x * x) / 2.0);
  }
  return -(2.5 * det_pow(zeta * kU, 4.0 / 5.0));
}

__device__ inline double mo_flux_ustar(double vel, double dist, double ustar, double z0, double tflux, double tref,
                                       double g, double nu) {
  const double eps = (double)1e-3f, yplcrit = 11.225;
  if (dist <= z0) {
    if (nu > 0)
      return (dist * ustar / nu) < yplcrit ? sqrt(nu * vel / dist)
                                           : vel / (det_log(fabs(ustar * dist / nu)) / 0.4 + 5.2);
    return 0;
  }
  double zL = 0, Psi = 0;
  for (int i = 1;; i++) {
    const double ustar2 = ustar;
    double m = dist / z0 - Psi;
    if (!(m > (double)1e-5f)) m = (double)1e-5f;
    double t = vel * 0.4 / det_log(m);
    if (!(t > 0)) t = 0;
    ustar = ustar + (t - ustar) / 2;
    if (ustar < (double)1e-4f) {
      zL = -10000;
    } else {
      const double zL2 = (dist * ((0.4 * (g / tref)) * tflux)) / (-((ustar * ustar) * ustar));
      zL = zL + (zL2 - zL) / 2;
    }
    Psi = psim_mo(zL);
    double den = fabs(ustar);
    if (!(den > 1e-3)) den = 1e-3;
    if (fabs(ustar - ustar2) / den < eps) break;
    if (i >= 50) {
      ustar = 0;
      break;
    }
  }
  return ustar;
}

// WM_MO_Dirichlet_ustar_tfl with z0H = z0 (logz0 = log(dist/z0), host libm)
__device__ inline void mo_dirichlet(double &ustar, double &tflux, double dist, double z0, double logz0, double vel,
                                    double tempdif, double nu, double kappa, double g, double tref) {
  const double eps = 1e-3, yplcrit = 11.225, kU = 0.4, kT = 0.47;
  ustar = wm_rough_ustar(ustar, vel, dist, z0, nu, logz0);
  double dist_plus = 1.7976931348623157e308;
  if (nu > 0) {
    const double ustar_lam = sqrt(nu * vel / dist);
    dist_plus = dist * ustar_lam / nu;
  }
  if (dist <= z0 || dist_plus < (double)1.1f * yplcrit) {
    tflux = -((kappa * tempdif) / dist);
    return;
  }
  double zL = 0, zL0 = -10000;
  int i = 0;
  for (;;) {
    i = i + 1;
    zL = -(((dist * g) * tflux) / ((kU * tref) * ((ustar * ustar) * ustar)));
    const double psi_m = psim_mo_mod(zL, kU), psi_h = psih_mo_mod(zL, kU);
    ustar = (vel * kU) / (logz0 - psi_m);
    tflux = -(((tempdif * ustar) * kT) / (logz0 - psi_h));
    double den = fabs(zL);
    if (!(den > 1e-4)) den = 1e-4;
    if (i > 1 && fabs(zL - zL0) / den < eps) break;
    if (i >= 50 || zL > 10000) break;
    zL0 = zL;
  }
  if (i >= 50 || zL > 10000) {
    ustar = sqrt(nu * vel / dist);
    tflux = -((kappa * tempdif) / dist);
  }
}

// WM_MO_Dirichlet_ustar_tfl_mfl (src/wallmodels.f90:1721-1784): the virtual
// heat flux in the Obukhov length, the moisture flux with k_Q = k_T; laminar
// branch dist <= z0 only; the fallback keeps the moisture flux
__device__ inline void mo_dirichlet_moist(double &ustar, double &tflux, double &qflux, double dist, double z0,
                                          double logz0, double vel, double temp, double surf_temp, double moist,
                                          double surf_moist, double nu, double kappa, double g, double tref) {
  const double eps = 1e-3, kU = 0.4, kT = 0.47, kQ = 0.47, e_coef = 0.61;
  const double tempdif = temp - surf_temp, moistdif = moist - surf_moist;
  ustar = wm_rough_ustar(ustar, vel, dist, z0, nu, logz0);
  if (dist <= z0) {
    tflux = -((kappa * tempdif) / dist);
    qflux = -((kappa * moistdif) / dist);
    return;
  }
  double zL = 0, zL0 = -10000;
  int i = 0;
  for (;;) {
    i = i + 1;
    const double virt = tflux * (1 + e_coef * moist) + (e_coef * temp) * qflux;
    zL = -(((dist * g) * virt) / ((kU * tref) * ((ustar * ustar) * ustar)));
    const double psi_m = psim_mo_mod(zL, kU), psi_h = psih_mo_mod(zL, kU), psi_e = psi_h;
    ustar = (vel * kU) / (logz0 - psi_m);
    tflux = -(((tempdif * ustar) * kT) / (logz0 - psi_h));
    qflux = -(((moistdif * ustar) * kQ) / (logz0 - psi_e));
    double den = fabs(zL);
    if (!(den > 1e-4)) den = 1e-4;
    if (i > 1 && fabs(zL - zL0) / den < eps) break;
    if (i >= 50 || zL > 10000) break;
    zL0 = zL;
  }
  if (i >= 50 || zL > 10000) {
    ustar = sqrt(nu * vel / dist);
    tflux = -((kappa * tempdif) / dist);
  }
}

template <typename T>
__global__ void k_wm_fluxes(Geo g, const T *__restrict__ U_, const T *__restrict__ V_, const T *__restrict__ W_,
                            const T *__restrict__ Temp, const T *__restrict__ Moist, WMPoints wm, double nu,
                            int variable_z) {
  // surface lists: evaluated in FP64 whatever the field type
  auto U = [&](long q) { return (double)U_[q]; };
  auto V = [&](long q) { return (double)V_[q]; };
  auto W = [&](long q) { return (double)W_[q]; };
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= wm.n) return;
  const int l = wm.list[p];
  const int comp = l / 6, dir = l % 6;
  const int xi = wm.ijk[3 * p], yj = wm.ijk[3 * p + 1], zk = wm.ijk[3 * p + 2];
  const long q = gix(g, xi, yj, zk);
  const long sy = g.sy, sz = g.sz;
  double v[3];
  if (comp == 0) {
    v[0] = U(q);
    v[1] = (V(q + 1) + V(q) + V(q + 1 - sy) + V(q - sy)) / 4.0;
    v[2] = (W(q + 1) + W(q) + W(q + 1 - sz) + W(q - sz)) / 4.0;
  } else if (comp == 1) {
    v[0] = (U(q + sy) + U(q) + U(q - 1 + sy) + U(q - 1)) / 4.0;
    v[1] = V(q);
    v[2] = (W(q + sy) + W(q) + W(q + sy - sz) + W(q - sz)) / 4.0;
  } else {
    v[0] = (U(q + sz) + U(q) + U(q - 1 + sz) + U(q - 1)) / 4.0;
    v[1] = (V(q + sz) + V(q) + V(q - sy + sz) + V(q - sy)) / 4.0;
    v[2] = W(q);
  }
  const double d0 = wm.dist[3 * p], d1 = wm.dist[3 * p + 1], d2 = wm.dist[3 * p + 2];
  const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  double t0 = v[0] - 0.0, t1 = v[1] - 0.0, t2 = v[2] - 0.0;
  const double dot = t0 * d0 + t1 * d1 + t2 * d2;
  const double dd = dist * dist;
  t0 = t0 - dot * d0 / dd;
  t1 = t1 - dot * d1 / dd;
  t2 = t2 - dot * d2 / dd;
  const double vel = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
  double ustar = wm.ustar[p];
  const double z0 = wm.z0[p];
  if (z0 > 0 && wm.mo.buoy && wm.mo.tclass[p] == ELMM_WM_SURFACE_TEMP) {
    // WM_MO_DirichletStress with the temperature at the velocity point
    // (local_value, src/wallmodels.f90:2340-2353)
    const long off = comp == 0 ? 1 : comp == 1 ? sy : sz;
    const double temp = ((double)Temp[q + off] + (double)Temp[q]) / 2;
    double tf = wm.mo.tflux[p];
    if (wm.mo.moist) {  // WM_MO_DirichletStress_moist
      double qf = wm.mo.qflux[p];
      const double mq = ((double)Moist[q + off] + (double)Moist[q]) / 2;
      mo_dirichlet_moist(ustar, tf, qf, dist, z0, wm.logdz0[p], vel, temp, wm.mo.tval[p], mq, wm.mo.qval[p], nu,
                         wm.mo.kappa, wm.mo.g, wm.mo.tref);
      wm.mo.qflux[p] = qf;
    } else {
      mo_dirichlet(ustar, tf, dist, z0, wm.logdz0[p], vel, temp - wm.mo.tval[p], nu, wm.mo.kappa, wm.mo.g,
                   wm.mo.tref);
    }
    wm.mo.tflux[p] = tf;
  } else if (z0 > 0 && wm.mo.buoy && wm.mo.tflux[p] > 0) {
    // WM_MO_FluxStress
    ustar = vel != 0 ? mo_flux_ustar(vel, dist, ustar, z0, wm.mo.tflux[p], wm.mo.tref, wm.mo.g, nu) : 0.0;
  } else if (vel != 0) {
    ustar = z0 > 0 ? wm_rough_ustar(ustar, vel, dist, z0, nu, wm.logdz0[p]) : wm_flat_ustar(ustar, vel, dist, nu);
  }
  if (ustar < 0) ustar = 0;
  wm.ustar[p] = ustar;
  const double eps = 1.4901161193847656e-08;  // sqrt(epsilon(1._knd))
  // fluxp: the face flux per volume of the cell beyond the face, fluxm: of
  // the cell before it (src/wallmodels.f90:2285-2321); equal on uniform z
  double fluxp = 0, fluxm = 0;
  if (vel > eps) {
    const double tc = comp == 0 ? t0 : comp == 1 ? t1 : t2;
    if (variable_z && dir >= 4) {
      const int kp = dir == 4 ? zk : zk + 1;
      fluxp = fluxm = (tc / vel) * (ustar * ustar);
      fluxp = comp < 2 ? fluxp / g.dzPr[kp] : fluxp / g.dzW[kp];
      fluxm = comp < 2 ? fluxm / g.dzPr[kp - 1] : fluxm / g.dzW[kp - 1];
    } else {
      const double drec = dir < 2 ? 1 / g.dxmin : dir < 4 ? 1 / g.dymin : 1 / g.dzmin;
      fluxp = fluxm = (tc / vel) * (ustar * ustar) * drec;
    }
    if ((dir & 1) == 0) {
      fluxp = -fluxp;
      fluxm = -fluxm;
    }
  }
  wm.fluxp[p] = fluxp;
  wm.fluxm[p] = fluxm;
}

template <typename T>
void launch_wm_fluxes(const Geo &g, const T *U, const T *V, const T *W, const T *Temp, const T *Moist, WMPoints wm,
                      double nu, int variable_z, hipStream_t st) {
  if (wm.n == 0) return;
  hipLaunchKernelGGL(k_wm_fluxes<T>, dim3(nblk(wm.n, 256)), dim3(256), 0, st, g, U, V, W, Temp, Moist, wm, nu,
                     variable_z);
}

// ComputeViscsWM (src/wallmodels.f90:2050-2152; neutral, wall velocity 0,
// wallmodeltype 1): the persistent WMPoints%ustar of every scalar wall-model
// point from the cell-centred two-point velocity (2069-2071), by
// WMRoughVisc (1362-1379) where z0 > 0, else WMFlatVisc (1268-1304).  With
// Visc (the initial-condition tail) the point's wall-model viscosity is
// stored too; in SubgridStresses SubgridModel overwrites it, so the step
// keeps only u*.  FP64 whatever the field type.
template <typename T>
__global__ void k_wm_ustar(Geo g, const T *__restrict__ U_, const T *__restrict__ V_, const T *__restrict__ W_,
                           const T *__restrict__ Temp, const T *__restrict__ Moist, WMUstarPoints wp, double nu,
                           T *__restrict__ Visc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= wp.n) return;
  const int xi = wp.ijk[3 * p], yj = wp.ijk[3 * p + 1], zk = wp.ijk[3 * p + 2];
  const long q = gix(g, xi, yj, zk);
  const double v0 = ((double)U_[q] + (double)U_[q - 1]) / 2.0;
  const double v1 = ((double)V_[q] + (double)V_[q - g.sy]) / 2.0;
  const double v2 = ((double)W_[q] + (double)W_[q - g.sz]) / 2.0;
  const double d0 = wp.dist[3 * p], d1 = wp.dist[3 * p + 1], d2 = wp.dist[3 * p + 2];
  const double z0 = wp.z0[p];
  double ustar = wp.ustar[p], visc;
  if (z0 <= 0 && nu <= 0) {
    ustar = 0;
    visc = 0;
  } else {
    const double dist = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double t0 = v0 - 0.0, t1 = v1 - 0.0, t2 = v2 - 0.0;
    const double dot = t0 * d0 + t1 * d1 + t2 * d2;
    const double dd = dist * dist;
    t0 = t0 - dot * d0 / dd;
    t1 = t1 - dot * d1 / dd;
    t2 = t2 - dot * d2 / dd;
    const double vel = sqrt(t0 * t0 + t1 * t1 + t2 * t2);
    if (z0 > 0 && wp.mo.buoy && wp.mo.tclass[p] == ELMM_WM_SURFACE_TEMP) {
      // WM_MO_Dirichlet (src/wallmodels.f90:1858-1884)
      double tf = wp.mo.tflux[p];
      if (wp.mo.moist) {  // WM_MO_Dirichlet_moist (src/wallmodels.f90:1887-1919)
        double qf = wp.mo.qflux[p];
        mo_dirichlet_moist(ustar, tf, qf, dist, z0, wp.logdz0[p], vel, (double)Temp[q], wp.mo.tval[p],
                           (double)Moist[q], wp.mo.qval[p], nu, wp.mo.kappa, wp.mo.g, wp.mo.tref);
        wp.mo.qflux[p] = qf;
      } else {
        mo_dirichlet(ustar, tf, dist, z0, wp.logdz0[p], vel, (double)Temp[q] - wp.mo.tval[p], nu, wp.mo.kappa,
                     wp.mo.g, wp.mo.tref);
      }
      wp.mo.tflux[p] = tf;
      if (ustar < 0) ustar = 0;
      if (vel > 0 || tf != 0) {
        const double a = ustar * ustar * dist / vel;
        visc = a > nu ? a : nu;
      } else {
        visc = nu > 0 ? nu : 0.0;
      }
    } else if (z0 > 0 && wp.mo.buoy && wp.mo.tflux[p] > 0) {
      // WM_MO_FLUX (src/wallmodels.f90:1792-1832; gravity replaced by the
      // cosine of the wall normal with the vertical, as the reference does)
      if (vel != 0) {
        const double g_cos = fabs(d2) / dist;
        ustar = mo_flux_ustar(vel, dist, ustar, z0, wp.mo.tflux[p], wp.mo.tref, g_cos, nu);
        if (ustar < 0) ustar = 0;
        visc = ustar * ustar * dist / vel > nu ? ustar * ustar * dist / vel : (nu > 0 ? nu : 0.0);
      } else {
        ustar = 0;
        visc = nu > 0 ? nu : 0.0;
      }
    } else if (z0 > 0) {
      if (vel != 0) ustar = wm_rough_ustar(ustar, vel, dist, z0, nu, wp.logdz0[p]);
      if (ustar < 0) ustar = 0;
      if (vel > 0 && ustar * ustar * dist / vel > nu)
        visc = ustar * ustar * dist / vel;
      else
        visc = nu > 0 ? nu : 0.0;
    } else {
      if (vel != 0) ustar = wm_flat_ustar(ustar, vel, dist, nu);
      if (ustar < 0) ustar = 0;
      if (vel > 0)
        visc = dist * ustar / nu > 1 ? ustar * ustar * dist / vel : nu;
      else
        visc = nu;
    }
  }
  wp.ustar[p] = ustar;
  if (Visc) Visc[q] = (T)visc;
  // TemperatureExtra's surface source (src/scalars.f90:322-326)
  if (wp.heat) wp.heat[p] = wp.mo.tflux[p] * wp.af[p];
  // MoistureExtra's (src/scalars.f90:350-354)
  if (wp.mheat) wp.mheat[p] = wp.mo.qflux[p] * wp.af[p];
}

template <typename T>
void launch_wm_ustar(const Geo &g, const T *U, const T *V, const T *W, const T *Temp, const T *Moist,
                     WMUstarPoints wp, double nu, T *Visc, hipStream_t st) {
  if (wp.n == 0) return;
  hipLaunchKernelGGL(k_wm_ustar<T>, dim3(nblk(wp.n, 256)), dim3(256), 0, st, g, U, V, W, Temp, Moist, wp, nu, Visc);
}

// ------------------------------------------------------ CD4 advection
template <typename T>
__device__ inline T mask4(T a, T b, T c, T d) { return ((-a + 9 * b) + 9 * c) + -d; }

// Advection of component C at linear index p (CD4divU/V/W,
// src/advection.f90:334-834).  cls: 0 4th order, 1 2nd order, 2 none.
template <typename T, int C>
__device__ inline double cd4(const Geo &g, const T *__restrict__ U, const T *__restrict__ V,
                             const T *__restrict__ W, long p, int cls) {
  const long sy = g.sy, sz = g.sz;
  T acc = 0, d;
  if (cls == 2) return acc * (T)(-1.0 / 256);
  // F: advected component; A: advecting velocity per direction (U, V, W)
  const T *F = C == 0 ? U : C == 1 ? V : W;
  const long st[3] = {1, sy, sz};
  const T inv[3] = {(T)g.dxmin, (T)g.dymin, (T)g.dzmin};
  const T rinv[3] = {sizeof(T) == 4 ? (T)g.rdxf : (T)g.rdx, sizeof(T) == 4 ? (T)g.rdyf : (T)g.rdy,
                     sizeof(T) == 4 ? (T)g.rdzf : (T)g.rdz};
#pragma unroll
  for (int dd = 0; dd < 3; dd++) {
    const long s = st[dd];
    if (cls == 0) {
      T uv1b, uv1a, uv3b, uv3a;
      if (dd == C) {
        // self-advection: UV1(m)=Uint(m)*(F(m-1)+F(m)), Uint over F(m-2..m+1)
        T Fi;
        Fi = mask4<T>(F[p - s], F[p], F[p + s], F[p + 2 * s]);
        uv1b = Fi * (F[p] + F[p + s]);
        Fi = mask4<T>(F[p - 2 * s], F[p - s], F[p], F[p + s]);
        uv1a = Fi * (F[p - s] + F[p]);
        Fi = mask4<T>(F[p], F[p + s], F[p + 2 * s], F[p + 3 * s]);
        uv3b = Fi * (F[p] + F[p + 3 * s]);
        Fi = mask4<T>(F[p - 3 * s], F[p - 2 * s], F[p - s], F[p]);
        uv3a = Fi * (F[p - 3 * s] + F[p]);
      } else {
        // cross term: advecting component A = (U,V,W)[dd] interpolated along
        // axis C: Aint(m) = mask4<T>(A(m, C-1..C+2)); UV1(m) = Aint(m)*(F(m)+F(m+1))
        const T *A = dd == 0 ? U : dd == 1 ? V : W;
        const long c = st[C];
#define AI(o) mask4<T>(A[p + (o) * s - c], A[p + (o) * s], A[p + (o) * s + c], A[p + (o) * s + 2 * c])
        uv1b = AI(0) * (F[p] + F[p + s]);
        uv1a = AI(-1) * (F[p - s] + F[p]);
        uv3b = AI(1) * (F[p] + F[p + 3 * s]);
        uv3a = AI(-2) * (F[p - 3 * s] + F[p]);
#undef AI
      }
      d = 9 * (uv1b - uv1a);
      d = d - div_rn_nz<T>(uv3b - uv3a, (T)3, recip3<T>());
      acc = acc + div_rn<T>(d, inv[dd], rinv[dd]);
    } else {
      // 2nd order near boundaries (coef2ord = 64)
      if (dd == C) {
        d = 64 * ((F[p + s] + F[p]) * (F[p + s] + F[p]) - (F[p] + F[p - s]) * (F[p] + F[p - s]));
      } else {
        const T *A = dd == 0 ? U : dd == 1 ? V : W;
        const long c = st[C];
        // (F(+s)+F)*(A(+c)+A) - (F+F(-s))*(A(+c-s)+A(-s)); the reference's
        // W/y term lists the last pair as (V(j-1)+V(j-1,k+1)): a+b = b+a.
        d = 64 * ((F[p + s] + F[p]) * (A[p + c] + A[p]) - (F[p] + F[p - s]) * (A[p + c - s] + A[p - s]));
      }
      acc = acc + div_rn<T>(d, inv[dd], rinv[dd]);
    }
  }
  return acc * (T)(-1.0 / 256);
}

// ------------------------------------------------------ diffusion fluxes
// Flux on the upper face of cell p in direction d for component C
// (MomentumDiffusion_nobranch_2ord, src/momentum_diffusion.f90:169-288).
template <typename T, int C>
__device__ inline T dflux(const Geo &g, const T *__restrict__ nu, const T *__restrict__ F, long p,
                               int d) {
  const long sy = g.sy, sz = g.sz;
  const T rx = (T)1.0 / ((T)g.dxmin * (T)g.dxmin), ry = (T)1.0 / ((T)g.dymin * (T)g.dymin),
          rz = (T)1.0 / ((T)g.dzmin * (T)g.dzmin);
  const T q4 = (T)0.25;
  if (C == 0) {
    if (d == 0) return nu[p + 1] * (F[p + 1] - F[p]) * rx;
    if (d == 1) return q4 * (nu[p + 1 + sy] + nu[p + 1] + nu[p + sy] + nu[p]) * (F[p + sy] - F[p]) * ry;
    return q4 * (nu[p + 1 + sz] + nu[p + 1] + nu[p + sz] + nu[p]) * (F[p + sz] - F[p]) * rz;
  } else if (C == 1) {
    if (d == 0) return q4 * (nu[p + 1 + sy] + nu[p + 1] + nu[p + sy] + nu[p]) * (F[p + 1] - F[p]) * rx;
    if (d == 1) return nu[p + sy] * (F[p + sy] - F[p]) * ry;
    return q4 * (nu[p + sy + sz] + nu[p + sy] + nu[p + sz] + nu[p]) * (F[p + sz] - F[p]) * rz;
  } else {
    if (d == 0) return q4 * (nu[p + 1 + sz] + nu[p + 1] + nu[p + sz] + nu[p]) * (F[p + 1] - F[p]) * rx;
    if (d == 1) return q4 * (nu[p + sy + sz] + nu[p + sz] + nu[p + sy] + nu[p]) * (F[p + sy] - F[p]) * ry;
    return nu[p + sz] * (F[p + sz] - F[p]) * rz;
  }
}

template <typename T, int C>
__device__ inline T diffusion(const Geo &g, const T *__restrict__ nu, const T *__restrict__ F, long p,
                                   T a) {
  a = a + dflux<T, C>(g, nu, F, p, 0);
  a = a - dflux<T, C>(g, nu, F, p - 1, 0);
  a = a + dflux<T, C>(g, nu, F, p, 1);
  a = a - dflux<T, C>(g, nu, F, p - g.sy, 1);
  a = a + dflux<T, C>(g, nu, F, p, 2);
  a = a - dflux<T, C>(g, nu, F, p - g.sz, 2);
  return a;
}

// ------------------------------------------- stretched z (CD2, 2nd order)
// CDU/V/W_variable_z (src/advection_variable_z.f90:15-321): the divergence
// form minus the advective form, times 1/2; z differences over dzPr/dzW and
// zPr/zW, x and y over dxmin/dymin.
template <typename T, int C>
__device__ inline T cd2vz(const Geo &g, const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
                          long p, int k) {
  const long sy = g.sy, sz = g.sz;
  const T ax4 = (T)(0.25 / g.dxmin), ay4 = (T)(0.25 / g.dymin), four = 4;
  T d;
  if (C == 0) {
    d = -(((ax4 * (U[p + 1] + U[p])) * (U[p + 1] + U[p]) - (ax4 * (U[p] + U[p - 1])) * (U[p] + U[p - 1])) +
          ((ay4 * (U[p + sy] + U[p])) * (V[p + 1] + V[p]) - (ay4 * (U[p] + U[p - sy])) * (V[p + 1 - sy] + V[p - sy])) +
          ((U[p + sz] + U[p]) * (W[p + 1] + W[p]) - (U[p] + U[p - sz]) * (W[p + 1 - sz] + W[p - sz])) /
              (T)g.dzPr[k] / four);
    const T ax = (T)(0.5 / g.dxmin), ay = (T)(0.125 / g.dymin);
    const T va = ((V[p] + V[p + 1]) + V[p - sy]) + V[p + 1 - sy];
    const T wa = (((W[p] + W[p + 1]) + W[p - sz]) + W[p + 1 - sz]) / four;
    d = d - (((ax * (U[p + 1] - U[p - 1])) * U[p] + (ay * (U[p + sy] - U[p - sy])) * va) +
             ((U[p + sz] - U[p - sz]) / (T)(g.zPr[k + 1] - g.zPr[k - 1])) * wa);
  } else if (C == 1) {
    d = -(((ay4 * (V[p + sy] + V[p])) * (V[p + sy] + V[p]) - (ay4 * (V[p] + V[p - sy])) * (V[p] + V[p - sy])) +
          ((ax4 * (V[p + 1] + V[p])) * (U[p + sy] + U[p]) - (ax4 * (V[p] + V[p - 1])) * (U[p - 1 + sy] + U[p - 1])) +
          ((V[p + sz] + V[p]) * (W[p + sy] + W[p]) - (V[p] + V[p - sz]) * (W[p + sy - sz] + W[p - sz])) /
              (T)g.dzPr[k] / four);
    const T ax = (T)(0.125 / g.dxmin), ay = (T)(0.5 / g.dymin);
    const T ua = ((U[p] + U[p + sy]) + U[p - 1]) + U[p - 1 + sy];
    const T wa = (((W[p] + W[p + sy]) + W[p - sz]) + W[p + sy - sz]) / four;
    d = d - (((ax * (V[p + 1] - V[p - 1])) * ua + (ay * (V[p + sy] - V[p - sy])) * V[p]) +
             ((V[p + sz] - V[p - sz]) / (T)(g.zPr[k + 1] - g.zPr[k - 1])) * wa);
  } else {
    d = -(((W[p + sz] + W[p]) * (W[p + sz] + W[p]) - (W[p] + W[p - sz]) * (W[p] + W[p - sz])) / (T)g.dzW[k] / four +
          ((ay4 * (W[p + sy] + W[p])) * (V[p + sz] + V[p]) - (ay4 * (W[p] + W[p - sy])) * (V[p - sy + sz] + V[p - sy])) +
          ((ax4 * (W[p + 1] + W[p])) * (U[p + sz] + U[p]) - (ax4 * (W[p] + W[p - 1])) * (U[p - 1 + sz] + U[p - 1])));
    const T ax = (T)(0.125 / g.dxmin), ay = (T)(0.125 / g.dymin);
    const T ua = ((U[p] + U[p + sz]) + U[p - 1]) + U[p - 1 + sz];
    const T va = ((V[p] + V[p + sz]) + V[p - sy]) + V[p - sy + sz];
    d = d - (((ax * (W[p + 1] - W[p - 1])) * ua + (ay * (W[p + sy] - W[p - sy])) * va) +
             ((W[p + sz] - W[p - sz]) / (T)(g.zW[k + 1] - g.zW[k - 1])) * W[p]);
  }
  return d * (T)0.5;
}

// z flux on the upper face of cell p (level f) in
// MomentumDiffusion_variable_z_nobranch_2ord (src/momentum_diffusion.f90:
// 298-420): U, V: nu4/4 * dF / dzW(f); W: nu(f+1) * dW / dzPr(f+1).
template <typename T, int C>
__device__ inline T zflux_vz(const Geo &g, const T *__restrict__ nu, const T *__restrict__ F, long p, int f) {
  const long sy = g.sy, sz = g.sz;
  const T q4 = (T)0.25;
  if (C == 0) return q4 * (nu[p + 1 + sz] + nu[p + 1] + nu[p + sz] + nu[p]) * (F[p + sz] - F[p]) / (T)g.dzW[f];
  if (C == 1) return q4 * (nu[p + sy + sz] + nu[p + sy] + nu[p + sz] + nu[p]) * (F[p + sz] - F[p]) / (T)g.dzW[f];
  return nu[p + sz] * (F[p + sz] - F[p]) / (T)g.dzPr[f + 1];
}

template <typename T, int C>
__device__ inline T diffusion_vz(const Geo &g, const T *__restrict__ nu, const T *__restrict__ F, long p, int k,
                                 T a) {
  a = a + dflux<T, C>(g, nu, F, p, 0);
  a = a - dflux<T, C>(g, nu, F, p - 1, 0);
  a = a + dflux<T, C>(g, nu, F, p, 1);
  a = a - dflux<T, C>(g, nu, F, p - g.sy, 1);
  T tmp = zflux_vz<T, C>(g, nu, F, p, k);
  tmp = tmp - zflux_vz<T, C>(g, nu, F, p - g.sz, k - 1);
  tmp = tmp / (T)(C < 2 ? g.dzPr[k] : g.dzW[k]);
  return a + tmp;
}

// Wall-model replacement of a z face on a stretched grid
// (wmfluxes-variable_z-nobranch-{U,V,W}-inc.f90): the face above level f
// seen from the written cell at level t.
template <typename T, int C>
__device__ inline T zterm_vz(const Geo &g, const T *__restrict__ nu, const T *__restrict__ F, long p, int f, int t) {
  const long sy = g.sy, sz = g.sz;
  const T four = 4;
  if (C == 0)
    return (nu[p + 1 + sz] + nu[p + 1] + nu[p + sz] + nu[p]) * (F[p + sz] - F[p]) / (T)g.dzW[f] / four /
           (T)g.dzPr[t];
  if (C == 1)
    return (nu[p + sy + sz] + nu[p + sy] + nu[p + sz] + nu[p]) * (F[p + sz] - F[p]) / (T)g.dzW[f] / four /
           (T)g.dzPr[t];
  return nu[p + sz] * (F[p + sz] - F[p]) / (T)g.dzPr[f + 1] / (T)g.dzW[t];
}

// RK combine + PressureGrad + add(U2,U) for one cell
template <typename T, int C>
__device__ inline void finish(const Geo &g, const StageConst &sc, const T *__restrict__ F,
                              const T *__restrict__ Pr, T *__restrict__ Ustar, T *__restrict__ U2,
                              long p, int k, T w, T us_prev) {
  T u2 = sc.stage > 0 ? us_prev * (T)sc.rho * (T)sc.dt : (T)0.0;
  u2 = u2 + w * (T)sc.beta * (T)sc.dt;
  if (C == 0 && sc.prgx != 0) u2 = u2 + (T)sc.A * (T)sc.prgx;
  if (C == 1 && sc.prgy != 0) u2 = u2 + (T)sc.A * (T)sc.prgy;
  if (C == 0) u2 = u2 + (T)sc.Ax * (Pr[p + 1] - Pr[p]);
  if (C == 1) u2 = u2 + (T)sc.Ay * (Pr[p + g.sy] - Pr[p]);
  if (C == 2) {
    if (g.variable_z)
      u2 = u2 - (T)sc.coef * (Pr[p + g.sz] - Pr[p]) / (T)g.dzW[k];
    else
      u2 = u2 + (T)sc.Az * (Pr[p + g.sz] - Pr[p]);
  }
  u2 = u2 + F[p];
  Ustar[p] = w;
  U2[p] = u2;
}

// bulk advection + diffusion of component C at cell p: CD4 on uniform grids
// (VZ = 0), variable-z CD2 and diffusion on stretched z grids (VZ = 1)
template <typename T, int C, int VZ>
__device__ inline T momentum_rhs(const Geo &g, const T *__restrict__ U, const T *__restrict__ V,
                                 const T *__restrict__ W, const T *__restrict__ Visc, long p, int k, uint8_t cd) {
  const T *F = C == 0 ? U : C == 1 ? V : W;
  if (VZ) {
    T w = cd2vz<T, C>(g, U, V, W, p, k);
    return diffusion_vz<T, C>(g, Visc, F, p, k, w);
  }
  T w = cd4<T, C>(g, U, V, W, p, cd & CODE_ADV_MASK);
  return diffusion<T, C>(g, Visc, F, p, w);
}

template <typename T, int C, int VZ>
__global__ void __launch_bounds__(256)
    k_momentum(Geo g, const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
               const T *__restrict__ Visc, const T *__restrict__ Pr, T *__restrict__ Ustar,
               T *__restrict__ U2, const uint8_t *__restrict__ code, StageConst sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.n[C][0] || j > g.n[C][1]) return;
  const long p = gix(g, i, j, k);
  const uint8_t cd = code[p];
  const T *F = C == 0 ? U : C == 1 ? V : W;
  const T w = momentum_rhs<T, C, VZ>(g, U, V, W, Visc, p, k, cd);
  const T us_prev = Ustar[p];
  if (cd & CODE_CORRECTED) {
    // hand-over to the ordered-correction kernel: the bulk value in Ustar,
    // the previous stage's Ustar in U2
    Ustar[p] = w;
    U2[p] = us_prev;
  } else {
    finish<T, C>(g, sc, F, Pr, Ustar, U2, p, k, w, us_prev);
  }
}

template <typename T, int VZ>
void launch_momentum_vz(const Geo &g, int comp, const T *U, const T *V, const T *W, const T *Visc, const T *Pr,
                        T *Ustar, T *U2, const uint8_t *code, StageConst sc, hipStream_t st) {
  const int *n = g.n[comp];
  dim3 b(64, 4, 1), gr(nblk(n[0], 64), nblk(n[1], 4), n[2]);
  if (comp == 0)
    hipLaunchKernelGGL((k_momentum<T, 0, VZ>), gr, b, 0, st, g, U, V, W, Visc, Pr, Ustar, U2, code, sc);
  else if (comp == 1)
    hipLaunchKernelGGL((k_momentum<T, 1, VZ>), gr, b, 0, st, g, U, V, W, Visc, Pr, Ustar, U2, code, sc);
  else
    hipLaunchKernelGGL((k_momentum<T, 2, VZ>), gr, b, 0, st, g, U, V, W, Visc, Pr, Ustar, U2, code, sc);
}

template <typename T>
void launch_momentum(const Geo &g, int comp, const T *U, const T *V, const T *W, const T *Visc, const T *Pr, T *Ustar,
                     T *U2, const uint8_t *code, StageConst sc, hipStream_t st) {
  if (g.variable_z)
    launch_momentum_vz<T, 1>(g, comp, U, V, W, Visc, Pr, Ustar, U2, code, sc, st);
  else
    launch_momentum_vz<T, 0>(g, comp, U, V, W, Visc, Pr, Ustar, U2, code, sc, st);
}

// Cells with wall-model face-flux replacements (wmfluxes-nobranch-inc.f90:
// 49-208; wmfluxes-variable_z-nobranch-{U,V,W}-inc.f90 on stretched z): the
// bulk value (evaluated by the bulk momentum kernel and passed through
// Ustar, the previous Ustar through U2), then the terms in the reference's
// scatter order.
// term = {axis | neg<<2, WM point index, face-cell linear index, face level}
// A negative term (the cell before the face) carries fluxm, a positive one
// fluxp.
template <typename T, int C, int VZ>
__device__ inline void corr_cell(const Geo &g, const T *__restrict__ U, const T *__restrict__ V,
                                 const T *__restrict__ W, const T *__restrict__ Visc, const T *__restrict__ Pr,
                                 T *__restrict__ Ustar, T *__restrict__ U2, const StageConst &sc,
                                 const int32_t *__restrict__ cell, const int32_t *__restrict__ ptr,
                                 const int32_t *__restrict__ term, const double *__restrict__ fluxp,
                                 const double *__restrict__ fluxm, int c) {
  const TermList tl{0, cell, ptr, term};
  const int i = tl.cell[3 * c], j = tl.cell[3 * c + 1], k = tl.cell[3 * c + 2];
  const long p = gix(g, i, j, k);
  const T *F = C == 0 ? U : C == 1 ? V : W;
  // the bulk kernel left momentum_rhs of this cell in Ustar and the previous
  // stage's Ustar in U2
  T w = Ustar[p];
  const T us_prev = U2[p];
  for (int t = tl.ptr[c]; t < tl.ptr[c + 1]; t++) {
    const int32_t *tt = &tl.term[4 * t];
    const int ax = tt[0] & 3;
    const bool neg = (tt[0] >> 2) & 1;
    const T fl = VZ && ax == 2 ? zterm_vz<T, C>(g, Visc, F, (long)tt[2], tt[3], k)
                               : dflux<T, C>(g, Visc, F, (long)tt[2], ax);
    const T fp = (T)(neg ? fluxm[tt[1]] : fluxp[tt[1]]);
    if (neg) {
      w = w - fl;
      w = w - fp;
    } else {
      w = w + fl;
      w = w + fp;
    }
  }
  finish<T, C>(g, sc, F, Pr, Ustar, U2, p, k, w, us_prev);
}

template <typename T, int C, int VZ>
__global__ void k_momentum_corr(Geo g, const T *__restrict__ U, const T *__restrict__ V,
                                const T *__restrict__ W, const T *__restrict__ Visc,
                                const T *__restrict__ Pr, T *__restrict__ Ustar, T *__restrict__ U2,
                                const uint8_t *__restrict__ code, StageConst sc, TermList tl,
                                const double *__restrict__ fluxp, const double *__restrict__ fluxm) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= tl.n) return;
  corr_cell<T, C, VZ>(g, U, V, W, Visc, Pr, Ustar, U2, sc, tl.cell, tl.ptr, tl.term, fluxp, fluxm, c);
}

// The three components' corrections in one launch (they touch disjoint
// arrays: component c reads the stage-start field and writes its own Ustar
// and U2), so their short dependent load chains overlap.
template <typename T>
struct Corr3 {
  TermList t0, t1, t2;
  T *us0, *us1, *us2, *u20, *u21, *u22;
};

template <typename T, int VZ>
__global__ void k_momentum_corr3(Geo g, const T *__restrict__ U, const T *__restrict__ V, const T *__restrict__ W,
                                 const T *__restrict__ Visc, const T *__restrict__ Pr, Corr3<T> a, StageConst sc,
                                 const double *__restrict__ fluxp, const double *__restrict__ fluxm) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int n0 = a.t0.n, n1 = a.t1.n;
  if (c < n0) {
    corr_cell<T, 0, VZ>(g, U, V, W, Visc, Pr, a.us0, a.u20, sc, a.t0.cell, a.t0.ptr, a.t0.term, fluxp, fluxm, c);
  } else if ((c -= n0) < n1) {
    corr_cell<T, 1, VZ>(g, U, V, W, Visc, Pr, a.us1, a.u21, sc, a.t1.cell, a.t1.ptr, a.t1.term, fluxp, fluxm, c);
  } else if ((c -= n1) < a.t2.n) {
    corr_cell<T, 2, VZ>(g, U, V, W, Visc, Pr, a.us2, a.u22, sc, a.t2.cell, a.t2.ptr, a.t2.term, fluxp, fluxm, c);
  }
}

template <typename T>
void launch_momentum_corr3(const Geo &g, const T *U, const T *V, const T *W, const T *Visc, const T *Pr,
                           T *const Ustar[3], T *const U2[3], StageConst sc, const TermList tl[3],
                           const double *fluxp, const double *fluxm, hipStream_t st) {
  const int tot = tl[0].n + tl[1].n + tl[2].n;
  if (tot == 0) return;
  const Corr3<T> a{tl[0], tl[1], tl[2], Ustar[0], Ustar[1], Ustar[2], U2[0], U2[1], U2[2]};
  if (g.variable_z)
    hipLaunchKernelGGL((k_momentum_corr3<T, 1>), dim3(nblk(tot, 128)), dim3(128), 0, st, g, U, V, W, Visc, Pr, a, sc,
                       fluxp, fluxm);
  else
    hipLaunchKernelGGL((k_momentum_corr3<T, 0>), dim3(nblk(tot, 128)), dim3(128), 0, st, g, U, V, W, Visc, Pr, a, sc,
                       fluxp, fluxm);
}

template <typename T, int VZ>
void launch_momentum_corr_vz(const Geo &g, int comp, const T *U, const T *V, const T *W, const T *Visc, const T *Pr,
                             T *Ustar, T *U2, const uint8_t *code, StageConst sc, TermList tl, const double *fluxp,
                             const double *fluxm, hipStream_t st) {
  dim3 b(128), gr(nblk(tl.n, 128));
  if (comp == 0)
    hipLaunchKernelGGL((k_momentum_corr<T, 0, VZ>), gr, b, 0, st, g, U, V, W, Visc, Pr, Ustar, U2, code, sc, tl,
                       fluxp, fluxm);
  else if (comp == 1)
    hipLaunchKernelGGL((k_momentum_corr<T, 1, VZ>), gr, b, 0, st, g, U, V, W, Visc, Pr, Ustar, U2, code, sc, tl,
                       fluxp, fluxm);
  else
    hipLaunchKernelGGL((k_momentum_corr<T, 2, VZ>), gr, b, 0, st, g, U, V, W, Visc, Pr, Ustar, U2, code, sc, tl,
                       fluxp, fluxm);
}

template <typename T>
void launch_momentum_corr(const Geo &g, int comp, const T *U, const T *V, const T *W, const T *Visc, const T *Pr,
                          T *Ustar, T *U2, const uint8_t *code, StageConst sc, TermList tl, const double *fluxp,
                          const double *fluxm, hipStream_t st) {
  if (tl.n == 0) return;
  if (g.variable_z)
    launch_momentum_corr_vz<T, 1>(g, comp, U, V, W, Visc, Pr, Ustar, U2, code, sc, tl, fluxp, fluxm, st);
  else
    launch_momentum_corr_vz<T, 0>(g, comp, U, V, W, Visc, Pr, Ustar, U2, code, sc, tl, fluxp, fluxm, st);
}

// ------------------------------------------------------------- IB gathers
// TIBPoint_Interpolate, src/immersedboundary.f90:637-651; IBMomentum is
// gather-all-then-scatter (src/tsteps.f90:368-418).  The components are
// independent (each interpolates its own field), so one launch gathers the
// points of all three and a second scatters them.
// The interpolation is evaluated in FP64 (coefficients are FP64) and the
// result rounded to the field type.
struct IB3 {
  IBDev c0, c1, c2;
  int n0, n01, n012;  // prefix counts
};

// Point p of the concatenated lists: its component's list (selected field by
// field, so the kernel arguments are not copied to private memory) and its
// index q in that list.
struct IBSel {
  const int32_t *tgt, *ptr, *src;
  const double *coef;
  int q;
};
__device__ inline IBSel ib_pick(const IB3 &ib, int p) {
  const bool a0 = p < ib.n0, a1 = p < ib.n01;
  return IBSel{a0 ? ib.c0.tgt : a1 ? ib.c1.tgt : ib.c2.tgt, a0 ? ib.c0.ptr : a1 ? ib.c1.ptr : ib.c2.ptr,
               a0 ? ib.c0.src : a1 ? ib.c1.src : ib.c2.src, a0 ? ib.c0.coef : a1 ? ib.c1.coef : ib.c2.coef,
               a0 ? p : a1 ? p - ib.n0 : p - ib.n01};
}

template <typename T>
__global__ void k_ib_interp(Geo g, IB3 ib, const T *__restrict__ U, const T *__restrict__ V,
                            const T *__restrict__ W, double *__restrict__ tmp) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ib.n012) return;
  const IBSel d = ib_pick(ib, p);
  const T *F = p < ib.n0 ? U : p < ib.n01 ? V : W;
  double s = 0;
  for (int t = d.ptr[d.q]; t < d.ptr[d.q + 1]; t++)
    s = s + d.coef[t] * (double)F[gix(g, d.src[3 * t], d.src[3 * t + 1], d.src[3 * t + 2])];
  tmp[p] = s;
}
template <typename T>
__global__ void k_ib_scatter(Geo g, IB3 ib, const double *__restrict__ tmp, T *__restrict__ U, T *__restrict__ V,
                             T *__restrict__ W) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= ib.n012) return;
  const IBSel d = ib_pick(ib, p);
  T *F = p < ib.n0 ? U : p < ib.n01 ? V : W;
  F[gix(g, d.tgt[3 * d.q], d.tgt[3 * d.q + 1], d.tgt[3 * d.q + 2])] = (T)tmp[p];
}
template <typename T>
void launch_ib_momentum(const Geo &g, const IBDev ib[3], T *U, T *V, T *W, double *tmp, hipStream_t st) {
  IB3 b{ib[0], ib[1], ib[2], ib[0].n, ib[0].n + ib[1].n, ib[0].n + ib[1].n + ib[2].n};
  if (!b.n012) return;
  hipLaunchKernelGGL(k_ib_interp<T>, dim3(nblk(b.n012, 256)), dim3(256), 0, st, g, b, U, V, W, tmp);
  hipLaunchKernelGGL(k_ib_scatter<T>, dim3(nblk(b.n012, 256)), dim3(256), 0, st, g, b, tmp, U, V, W);
}

// ScalFlIBPoints viscosity loop (src/dynamics.f90:899-902) is sequential:
// a point may read a value written by an earlier point.  The host groups the
// points into dependency levels; points of one level are independent.
template <typename T>
__global__ void k_ib_visc_level(Geo g, IBDev ib, T *__restrict__ Visc, int p0, int p1) {
  const int p = p0 + blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= p1) return;
  double s = 0;
  for (int t = ib.ptr[p]; t < ib.ptr[p + 1]; t++)
    s = s + ib.coef[t] * (double)Visc[gix(g, ib.src[3 * t], ib.src[3 * t + 1], ib.src[3 * t + 2])];
  Visc[gix(g, ib.tgt[3 * p], ib.tgt[3 * p + 1], ib.tgt[3 * p + 2])] = (T)s;
}
template <typename T>
void launch_ib_visc_serial(const Geo &g, IBDev ib, T *Visc, const int32_t *level_ptr, int nlevels, hipStream_t st) {
  for (int l = 0; l < nlevels; l++) {
    int p0 = level_ptr[l], p1 = level_ptr[l + 1];
    if (p1 > p0)
      hipLaunchKernelGGL(k_ib_visc_level<T>, dim3(nblk(p1 - p0, 256)), dim3(256), 0, st, g, ib, Visc, p0, p1);
  }
}

// NullInterior, src/tsteps.f90:543-567 (the three components in one launch)
template <typename T>
__global__ void k_null(Geo g, int n0, int n1, int n2, const int32_t *__restrict__ a, const int32_t *__restrict__ b,
                       const int32_t *__restrict__ c, T *__restrict__ U, T *__restrict__ V, T *__restrict__ W) {
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int32_t *ijk = a;
  T *F = U;
  if (p >= n0) {
    p -= n0;
    ijk = b;
    F = V;
    if (p >= n1) {
      p -= n1;
      ijk = c;
      F = W;
      if (p >= n2) return;
    }
  }
  F[gix(g, ijk[3 * p], ijk[3 * p + 1], ijk[3 * p + 2])] = 0;
}
template <typename T>
void launch_null(const Geo &g, const int n[3], const int32_t *const ijk[3], T *U, T *V, T *W, hipStream_t st) {
  const int tot = n[0] + n[1] + n[2];
  if (tot)
    hipLaunchKernelGGL(k_null<T>, dim3(nblk(tot, 256)), dim3(256), 0, st, g, n[0], n[1], n[2], ijk[0], ijk[1], ijk[2],
                       U, V, W);
}

template <typename T>
__global__ void k_mass_sources(Geo g, TermList tl, const T *__restrict__ U, const T *__restrict__ V,
                               const T *__restrict__ W, double *__restrict__ Q) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= tl.n) return;
  double q = 0;
  for (int t = tl.ptr[c]; t < tl.ptr[c + 1]; t++) {
    const int32_t *tt = &tl.term[4 * t];
    const int comp = tt[0] & 3;
    const bool neg = (tt[0] >> 2) & 1;
    const T *F = comp == 0 ? U : comp == 1 ? V : W;
    const double *dl = comp == 0 ? g.dxPr : comp == 1 ? g.dyPr : g.dzPr;
    const double val = (double)F[(long)tt[1]] / dl[tt[2]];
    q = neg ? q - val : q + val;
  }
  Q[gix(g, tl.cell[3 * c], tl.cell[3 * c + 1], tl.cell[3 * c + 2])] = q;
}

// Bound_Q, src/boundaries.f90:2094-2192: the ghost layers of Q fold into
// the interior in three passes, x (rows j, k in 1..n), then y (columns i, k),
// then z (columns i, j).  No pass writes a ghost cell, and a pass reads
// interior cells only where it writes them, so the final value of an
// interior boundary cell is its own value with the three axes' ghost terms
// added in pass order; one thread per boundary cell applies them in that
// order (the same additions, the same results).
__device__ inline double fold_axis(double v, int m, int n, int periodic, double g0, double g1) {
  // g0 = ghost 0, g1 = ghost n+1 of the axis
  if (periodic) {
    if (m == n) v = g0 + v;
    if (m == 1) v = v + g1;
  } else {
    if (m == 1) v = v + g0;
    if (m == n) v = v + g1;
  }
  return v;
}

// (32-bit unsigned index arithmetic: a boundary shell of the admitted grids
// has far fewer than 2^32 cells)
__global__ void k_fold_q(Geo g, double *__restrict__ Q, int px, int py, int pz) {
  const unsigned t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const unsigned fx = nx > 1 ? 2 : 1, fy = ny > 1 ? 2 : 1, fz = nz > 1 ? 2 : 1;
  const unsigned ix = (unsigned)max(nx - 2, 0), iy = (unsigned)max(ny - 2, 0);
  const unsigned ca = fx * ny * nz, cb = ix * fy * nz, cc = ix * iy * fz;
  int i, j, k;
  if (t < ca) {  // x faces, every j, k
    const unsigned r = t / fx, q = r / (unsigned)ny;
    i = t - r * fx ? nx : 1;
    j = 1 + (int)(r - q * ny);
    k = 1 + (int)q;
  } else if (t < ca + cb) {  // y faces, i inside
    const unsigned u = t - ca, r = u / fy, q = r / ix;
    j = u - r * fy ? ny : 1;
    i = 2 + (int)(r - q * ix);
    k = 1 + (int)q;
  } else if (t < ca + cb + cc) {  // z faces, i and j inside
    const unsigned u = t - ca - cb, r = u / fz, q = r / ix;
    k = u - r * fz ? nz : 1;
    i = 2 + (int)(r - q * ix);
    j = 2 + (int)q;
  } else {
    return;
  }
  const long p = gix(g, i, j, k);
  double v = Q[p];
  if (i == 1 || i == nx) v = fold_axis(v, i, nx, px, Q[gix(g, 0, j, k)], Q[gix(g, nx + 1, j, k)]);
  if (j == 1 || j == ny) v = fold_axis(v, j, ny, py, Q[gix(g, i, 0, k)], Q[gix(g, i, ny + 1, k)]);
  if (k == 1 || k == nz) v = fold_axis(v, k, nz, pz, Q[gix(g, i, j, 0)], Q[gix(g, i, j, nz + 1)]);
  Q[p] = v;
}

template <typename T>
void launch_mass_sources(const Geo &g, TermList tl, const T *U, const T *V, const T *W, double *Q, hipStream_t st) {
  if (!tl.n) return;
  hipLaunchKernelGGL(k_mass_sources<T>, dim3(nblk(tl.n, 256)), dim3(256), 0, st, g, tl, U, V, W, Q);
  const int fx = g.nx > 1 ? 2 : 1, fy = g.ny > 1 ? 2 : 1, fz = g.nz > 1 ? 2 : 1;
  const long n = (long)fx * g.ny * g.nz + (long)std::max(g.nx - 2, 0) * fy * g.nz +
                 (long)std::max(g.nx - 2, 0) * std::max(g.ny - 2, 0) * fz;
  hipLaunchKernelGGL(k_fold_q, dim3(nblk(n, 256)), dim3(256), 0, st, g, Q, g.bt[0] == 3, g.bt[2] == 3,
                     g.bt[4] == 3);
}

// explicit instantiations: FP64 (reference arithmetic) and FP32 (precision form)
#define ELMM_INST_MOMENTUM(T)                                                                                       \
  template void launch_sgs<T>(const Geo &, const T *, const T *, const T *, T *, double, double, double,           \
                              SgsWidth, int, double *, int, hipStream_t);                                          \
  template void launch_wm_fluxes<T>(const Geo &, const T *, const T *, const T *, const T *, const T *, WMPoints,  \
                                    double, int, hipStream_t);                                                     \
  template void launch_wm_ustar<T>(const Geo &, const T *, const T *, const T *, const T *, const T *,             \
                                   WMUstarPoints, double, T *, hipStream_t);                                       \
  template void launch_momentum<T>(const Geo &, int, const T *, const T *, const T *, const T *, const T *, T *,   \
                                   T *, const uint8_t *, StageConst, hipStream_t);                                 \
  template void launch_momentum_corr3<T>(const Geo &, const T *, const T *, const T *, const T *, const T *,      \
                                         T *const[3], T *const[3], StageConst, const TermList[3], const double *,  \
                                         const double *, hipStream_t);                                             \
  template void launch_momentum_corr<T>(const Geo &, int, const T *, const T *, const T *, const T *, const T *,   \
                                        T *, T *, const uint8_t *, StageConst, TermList, const double *,           \
                                        const double *, hipStream_t);                                              \
  template void launch_ib_momentum<T>(const Geo &, const IBDev[3], T *, T *, T *, double *, hipStream_t);         \
  template void launch_ib_visc_serial<T>(const Geo &, IBDev, T *, const int32_t *, int, hipStream_t);              \
  template void launch_null<T>(const Geo &, const int[3], const int32_t *const[3], T *, T *, T *, hipStream_t);    \
  template void launch_mass_sources<T>(const Geo &, TermList, const T *, const T *, const T *, double *,           \
                                       hipStream_t);
ELMM_INST_MOMENTUM(double)
ELMM_INST_MOMENTUM(float)

}  // namespace elmm
