// k_scalar.hip - passive-scalar RK3 stage as one fused pass per cell:
// kappa-scheme advection with flux limiter (KappaScalar,
// src/scalar_advection.f90:102-269), zeroing in solids (src/scalars.f90:
// 261-269), explicit diffusion (ScalarDiffusion_nobranch,
// src/scalar_diffusion.f90:404-449), RK combine and update
// (src/scalars.f90:279-283).
//
// The reference scatters each face flux into the two cells it separates
// (x, y passes in increasing index order; z pass over even faces first, then
// odd faces).  Here every cell gathers its two fluxes per direction and adds
// them in exactly that order, so the result is the reference's value bit for
// bit and the pass needs no atomics and no Scal2 zero-fill.
#include <cmath>

#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

static inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// Real = field and arithmetic type (double: FP64 form, float: precision form)
// FP64 (the reference's arithmetic): the quotient correctly rounded and the
// reference's comparisons.  FP32 (the precision form): the ratio as a product
// with the hardware reciprocal (within 2 ulp) and min/max instructions (equal
// to the comparisons for every non-NaN ratio).
template <typename Real>
__device__ inline Real flux_limiter(Real r) {
  if constexpr (sizeof(Real) == 4) {
    return fmaxf(0.0f, fminf(2 * r, fminf(2.0f, (1 + 2 * r) * (1.0f / 3.0f))));
  } else {
    Real a = div_rn_nz<Real>(1 + 2 * r, (Real)3, recip3<Real>());
    Real b = (Real)2.0 < a ? (Real)2.0 : a;
    Real c = 2 * r < b ? 2 * r : b;
    return (Real)0.0 > c ? (Real)0.0 : c;
  }
}
template <typename Real>
__device__ inline Real ratio(Real a, Real b) {
  if constexpr (sizeof(Real) == 4)
    return a * __builtin_amdgcn_rcpf(b);
  else
    return a / b;
}

// Flux through the face between s0 and s1 with face velocity u, from the four
// values s_m1 = S(-1), s0 = S(0), s1 = S(+1), s2 = S(+2) along the direction.
template <typename Real>
__device__ inline Real kappa_flux_v(Real u, Real s_m1, Real s0, Real s1, Real s2) {
  const Real eps = (Real)1e-8f;  // default-real literal in the reference
  Real sr, sl;
  if (u > 0) {
    sr = s1 - s0;
    sl = s0 - s_m1;
  } else {
    sr = s0 - s1;
    sl = s1 - s2;
  }
  const Real sg = copysign((Real)1.0, sl);
  const Real slope = flux_limiter<Real>(ratio<Real>(sr + eps * sg, sl + eps * sg));
  if (u > 0) return u * (s0 + (s0 - s_m1) * slope / (Real)2.0);
  return u * (s1 + (s1 - s2) * slope / (Real)2.0);
}

// Flux through the upper face of cell p along stride s with face velocity u.
template <typename Real>
__device__ inline Real kappa_flux(const Real *__restrict__ S, long p, long s, Real u) {
  const Real eps = (Real)1e-8f;  // default-real literal in the reference
  Real sr, sl;
  if (u > 0) {
    sr = S[p + s] - S[p];
    sl = S[p] - S[p - s];
  } else {
    sr = S[p] - S[p + s];
    sl = S[p + s] - S[p + 2 * s];
  }
  const Real sg = copysign((Real)1.0, sl);
  const Real slope = flux_limiter<Real>(ratio<Real>(sr + eps * sg, sl + eps * sg));
  if (u > 0) return u * (S[p] + (S[p] - S[p - s]) * slope / (Real)2.0);
  return u * (S[p + s] + (S[p + s] - S[p + 2 * s]) * slope / (Real)2.0);
}

template <typename Real>
__device__ inline Real scalar_cell(const Geo &g, const Real *__restrict__ U, const Real *__restrict__ V,
                                   const Real *__restrict__ W, const Real *__restrict__ T,
                                   const Real *__restrict__ S, const uint8_t *__restrict__ mx,
                                     const uint8_t *__restrict__ my, const uint8_t *__restrict__ mz, long p, int k,
                                     uint8_t code) {
  const long sy = g.sy, sz = g.sz;
  const Real Ax = 1 / (Real)g.dxmin, Ay = 1 / (Real)g.dymin, Az = 1 / (Real)g.dzmin;
  Real a = 0;
  // x pass: face i-1 adds, face i subtracts
  if (mx[p - 1]) a = a + Ax * kappa_flux(S, p - 1, 1, U[p - 1]);
  if (mx[p]) a = a - Ax * kappa_flux(S, p, 1, U[p]);
  if (my[p - sy]) a = a + Ay * kappa_flux(S, p - sy, sy, V[p - sy]);
  if (my[p]) a = a - Ay * kappa_flux(S, p, sy, V[p]);
  // z pass: faces of even index are processed before faces of odd index
  const bool mlo = mz[p - sz] != 0, mhi = mz[p] != 0;
  if ((k & 1) == 0) {
    if (mhi) a = a - Az * kappa_flux(S, p, sz, W[p]);
    if (mlo) a = a + Az * kappa_flux(S, p - sz, sz, W[p - sz]);
  } else {
    if (mlo) a = a + Az * kappa_flux(S, p - sz, sz, W[p - sz]);
    if (mhi) a = a - Az * kappa_flux(S, p, sz, W[p]);
  }
  if (code & CODE_SOLID) a = 0;
  const Real Dx = 1 / (2 * ((Real)g.dxmin * (Real)g.dxmin)), Dy = 1 / (2 * ((Real)g.dymin * (Real)g.dymin)),
             Dz = 1 / (2 * ((Real)g.dzmin * (Real)g.dzmin));
  a = a + (T[p + 1] + T[p]) * (S[p + 1] - S[p]) * Dx;
  a = a - (T[p] + T[p - 1]) * (S[p] - S[p - 1]) * Dx;
  a = a + (T[p + sy] + T[p]) * (S[p + sy] - S[p]) * Dy;
  a = a - (T[p] + T[p - sy]) * (S[p] - S[p - sy]) * Dy;
  a = a + (T[p + sz] + T[p]) * (S[p + sz] - S[p]) * Dz;
  a = a - (T[p] + T[p - sz]) * (S[p] - S[p - sz]) * Dz;
  return a;
}

template <typename Real>
__device__ inline void scalar_finish(const StageConst &sc, const Real *__restrict__ S, Real *__restrict__ Snew,
                                     Real *__restrict__ Sadv, long p, Real a) {
  Real s2 = sc.stage > 0 ? Sadv[p] * (Real)sc.rho_dt : (Real)0.0;
  s2 = s2 + a * (Real)sc.beta_dt;
  Sadv[p] = a;
  Snew[p] = S[p] + s2;
}

template <typename Real>
__global__ void __launch_bounds__(256)
    k_scalar(Geo g, const Real *__restrict__ U, const Real *__restrict__ V, const Real *__restrict__ W,
             const Real *__restrict__ T, const Real *__restrict__ S, Real *__restrict__ Snew,
             Real *__restrict__ Sadv, const uint8_t *__restrict__ code, const uint8_t *__restrict__ mx,
             const uint8_t *__restrict__ my, const uint8_t *__restrict__ mz, StageConst sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long p = gix(g, i, j, k);
  const uint8_t cd = code[p];
  Real a = scalar_cell<Real>(g, U, V, W, T, S, mx, my, mz, p, k, cd);
  if (cd & CODE_CORRECTED)
    Snew[p] = a;  // bulk value for the ordered-correction kernel
  else
    scalar_finish<Real>(sc, S, Snew, Sadv, p, a);
}

template <typename Real>
void launch_scalar(const Geo &g, const Real *U, const Real *V, const Real *W, const Real *TDiff, const Real *S,
                   Real *Snew, Real *Sadv, const uint8_t *code, const uint8_t *mx, const uint8_t *my,
                   const uint8_t *mz, StageConst sc, hipStream_t st) {
  dim3 b(64, 4, 1), gr(nblk(g.nx, 64), nblk(g.ny, 4), g.nz);
  hipLaunchKernelGGL(k_scalar<Real>, gr, b, 0, st, g, U, V, W, TDiff, S, Snew, Sadv, code, mx, my, mz, sc);
}

// ------------------------------------------------------------------------
// z-marching form of k_scalar for uniform grids: a block owns a TXS x TYS
// tile of columns; S at level k (halo 2) and TDiff at level k (halo 1) are
// LDS planes, double-buffered with the next level fetched into registers
// while the current one is computed; the column values S(k-2..k+2),
// TDiff(k-1..k+1), W(k-1..k) and the z masks live in registers.  Every face
// flux is evaluated once: each thread computes the upper x and y faces of its
// cell (edge threads also the tile's lower faces) into LDS, and the upper z
// face of level k is the lower z face of level k+1.  Per cell the operations
// are those of scalar_cell + scalar_finish in the same order.
constexpr int TXS = 64, TYS = 4, NTS = TXS * TYS;
constexpr int SPW = TXS + 4, SPH = TYS + 4, SPS = SPW * SPH;   // S plane, halo 2
constexpr int TPW = TXS + 2, TPH = TYS + 2, TPS = TPW * TPH;   // TDiff plane, halo 1
constexpr int SLD = (SPS + NTS - 1) / NTS, TLD = (TPS + NTS - 1) / NTS;

template <typename Real>
__global__ void __launch_bounds__(NTS)
    k_scalar_zm(Geo g, const Real *__restrict__ U, const Real *__restrict__ V, const Real *__restrict__ W,
                const Real *__restrict__ T, const Real *__restrict__ S, Real *__restrict__ Snew,
                Real *__restrict__ Sadv, const uint8_t *__restrict__ code, const uint8_t *__restrict__ mx,
                const uint8_t *__restrict__ my, const uint8_t *__restrict__ mz, StageConst sc, int kchunk) {
  // double-buffered planes, level L in buffer (L - kbeg) & 1 (the level loop
  // is unrolled by 2, so the buffers are constants); a thread past the end
  // of a plane loads and stores its last element again (its owner's value),
  // so the plane loads and stores carry no branches
  __shared__ Real sS[2][SPS];
  __shared__ Real sT[2][TPS];
  __shared__ Real fX[TYS][TXS + 1];  // x faces i0-1 .. i0+TXS-1 of each row
  __shared__ Real fY[TYS + 1][TXS];  // y faces j0-1 .. j0+TYS-1 of each column
  const int tid = threadIdx.x, tx = tid % TXS, ty = tid / TXS;
  const int i0 = 1 + blockIdx.x * TXS, j0 = 1 + blockIdx.y * TYS;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = 1 + blockIdx.z * kchunk, kend = min(kbeg + kchunk - 1, g.nz);
  const long sy = g.sy, sz = g.sz;
  long soff[SLD], toff[TLD];
  int sidx[SLD], tidx[TLD];
#pragma unroll
  for (int r = 0; r < SLD; r++) {
    const int e = min(tid + r * NTS, SPS - 1);
    sidx[r] = e;
    soff[r] = gix(g, min(i0 - 2 + e % SPW, g.nx + 3), min(j0 - 2 + e / SPW, g.ny + 3), 0);
  }
#pragma unroll
  for (int r = 0; r < TLD; r++) {
    const int e = min(tid + r * NTS, TPS - 1);
    tidx[r] = e;
    toff[r] = gix(g, min(i0 - 1 + e % TPW, g.nx + 3), min(j0 - 1 + e / TPW, g.ny + 3), 0);
  }
  const long pc = gix(g, min(i, g.nx), min(j, g.ny), 0);
  const bool in_xy = i <= g.nx && j <= g.ny;
  const int ls = (ty + 2) * SPW + tx + 2, lt = (ty + 1) * TPW + tx + 1;

  // prologue: planes of level kbeg, column rings
#pragma unroll
  for (int r = 0; r < SLD; r++)
    sS[0][sidx[r]] = S[soff[r] + (long)kbeg * sz];
#pragma unroll
  for (int r = 0; r < TLD; r++)
    sT[0][tidx[r]] = T[toff[r] + (long)kbeg * sz];
  Real cs[5], ct[3], cw[2];
  uint8_t cm[2];
#pragma unroll
  for (int q = 0; q < 5; q++) cs[q] = S[pc + (long)(kbeg - 2 + q) * sz];
#pragma unroll
  for (int q = 0; q < 3; q++) ct[q] = T[pc + (long)(kbeg - 1 + q) * sz];
  cw[0] = W[pc + (long)(kbeg - 1) * sz];
  cw[1] = W[pc + (long)kbeg * sz];
  cm[0] = mz[pc + (long)(kbeg - 1) * sz];
  cm[1] = mz[pc + (long)kbeg * sz];
  __syncthreads();

  const Real Ax = 1 / (Real)g.dxmin, Ay = 1 / (Real)g.dymin, Az = 1 / (Real)g.dzmin;
  const Real Dx = 1 / (2 * ((Real)g.dxmin * (Real)g.dxmin)), Dy = 1 / (2 * ((Real)g.dymin * (Real)g.dymin)),
             Dz = 1 / (2 * ((Real)g.dzmin * (Real)g.dzmin));
  Real fz_lo = 0;
  // per-cell operands of the level being computed (face velocities and
  // masks of the lower and upper x/y faces, the code, the previous RK term),
  // fetched one level ahead like the planes
  struct Ops {
    Real um, u0, vm, v0, sadv;
    uint8_t mxm, mx0, mym, my0, cd;
  };
  auto fetch = [&](long q, Ops &o) {
    o.um = U[q - 1];
    o.u0 = U[q];
    o.vm = V[q - sy];
    o.v0 = V[q];
    o.mxm = mx[q - 1];
    o.mx0 = mx[q];
    o.mym = my[q - sy];
    o.my0 = my[q];
    o.cd = code[q];
    o.sadv = sc.stage > 0 ? Sadv[q] : (Real)0;
  };
  Ops cur, nxt;
  fetch(pc + (long)kbeg * sz, cur);
  nxt = cur;

  for (int k0 = kbeg; k0 <= kend; k0 += 2) {
#pragma unroll
  for (int b = 0; b < 2; b++) {
    const int k = k0 + b;
    if (k > kend) break;
    // next level's planes and column values (levels inside the allocation;
    // after the chunk's last level they are loaded and not used)
    Real ps[SLD], pt[TLD], ns, nt, nw;
    uint8_t nm;
#pragma unroll
    for (int r = 0; r < SLD; r++)
      ps[r] = S[soff[r] + (long)(k + 1) * sz];
#pragma unroll
    for (int r = 0; r < TLD; r++)
      pt[r] = T[toff[r] + (long)(k + 1) * sz];
    ns = S[pc + (long)(k + 3) * sz];
    nt = T[pc + (long)(k + 2) * sz];
    nw = W[pc + (long)(k + 1) * sz];
    nm = mz[pc + (long)(k + 1) * sz];
    fetch(pc + (long)(k + 1) * sz, nxt);
    const long p = gix(g, min(i, g.nx), min(j, g.ny), k);
    const Real *P = sS[b];
    auto Sx = [&](int o) { return P[ls + o]; };
    auto Sy = [&](int o) { return P[ls + o * SPW]; };
    // face fluxes of this level: upper x and y faces of the cell, the tile's
    // lower faces by its first column / row (values of out-of-domain threads
    // are never read)
    const uint8_t mxu = cur.mx0, myu = cur.my0;
    if (in_xy) {
      fX[ty][tx + 1] = mxu ? kappa_flux_v<Real>(cur.u0, Sx(-1), Sx(0), Sx(1), Sx(2)) : (Real)0;
      fY[ty + 1][tx] = myu ? kappa_flux_v<Real>(cur.v0, Sy(-1), Sy(0), Sy(1), Sy(2)) : (Real)0;
      if (tx == 0 && cur.mxm) fX[ty][0] = kappa_flux_v<Real>(cur.um, Sx(-2), Sx(-1), Sx(0), Sx(1));
      if (ty == 0 && cur.mym) fY[0][tx] = kappa_flux_v<Real>(cur.vm, Sy(-2), Sy(-1), Sy(0), Sy(1));
    }
    // z faces: the lower one is the previous level's upper one
    const bool mlo = cm[0] != 0, mhi = cm[1] != 0;
    if (k == kbeg && mlo) fz_lo = kappa_flux_v<Real>(cw[0], cs[0], cs[1], cs[2], cs[3]);
    const Real fz_hi = mhi ? kappa_flux_v<Real>(cw[1], cs[1], cs[2], cs[3], cs[4]) : (Real)0;
    __syncthreads();
    if (in_xy) {
      const uint8_t cd = cur.cd;
      {
        Real a = 0;
        if (cur.mxm) a = a + Ax * fX[ty][tx];
        if (mxu) a = a - Ax * fX[ty][tx + 1];
        if (cur.mym) a = a + Ay * fY[ty][tx];
        if (myu) a = a - Ay * fY[ty + 1][tx];
        if ((k & 1) == 0) {
          if (mhi) a = a - Az * fz_hi;
          if (mlo) a = a + Az * fz_lo;
        } else {
          if (mlo) a = a + Az * fz_lo;
          if (mhi) a = a - Az * fz_hi;
        }
        if (cd & CODE_SOLID) a = 0;
        const Real *Q = sT[b];
        const Real t0 = ct[1], s0 = cs[2];
        a = a + (Q[lt + 1] + t0) * (Sx(1) - s0) * Dx;
        a = a - (t0 + Q[lt - 1]) * (s0 - Sx(-1)) * Dx;
        a = a + (Q[lt + TPW] + t0) * (Sy(1) - s0) * Dy;
        a = a - (t0 + Q[lt - TPW]) * (s0 - Sy(-1)) * Dy;
        a = a + (ct[2] + t0) * (cs[3] - s0) * Dz;
        a = a - (t0 + ct[0]) * (s0 - cs[1]) * Dz;
        if (cd & CODE_CORRECTED) {
          Snew[p] = a;  // bulk value for the ordered-correction kernel
        } else {
          Real s2 = sc.stage > 0 ? cur.sadv * (Real)sc.rho_dt : (Real)0.0;
          s2 = s2 + a * (Real)sc.beta_dt;
          Sadv[p] = a;
          Snew[p] = s0 + s2;
        }
      }
    }
    cur = nxt;
#pragma unroll
    for (int r = 0; r < SLD; r++)
      sS[b ^ 1][sidx[r]] = ps[r];
#pragma unroll
    for (int r = 0; r < TLD; r++)
      sT[b ^ 1][tidx[r]] = pt[r];
#pragma unroll
    for (int q = 0; q < 4; q++) cs[q] = cs[q + 1];
    cs[4] = ns;
    ct[0] = ct[1];
    ct[1] = ct[2];
    ct[2] = nt;
    cw[0] = cw[1];
    cw[1] = nw;
    cm[0] = cm[1];
    cm[1] = nm;
    fz_lo = fz_hi;
    __syncthreads();
  }
  }
}

template <typename Real>
void launch_scalar_zm(const Geo &g, const Real *U, const Real *V, const Real *W, const Real *TDiff, const Real *S,
                      Real *Snew, Real *Sadv, const uint8_t *code, const uint8_t *mx, const uint8_t *my,
                      const uint8_t *mz, StageConst sc, int kchunk, hipStream_t st) {
  if (kchunk <= 0) kchunk = g.nz;
  dim3 b(NTS), gr(nblk(g.nx, TXS), nblk(g.ny, TYS), nblk(g.nz, kchunk));
  hipLaunchKernelGGL(k_scalar_zm<Real>, gr, b, 0, st, g, U, V, W, TDiff, S, Snew, Sadv, code, mx, my, mz, sc,
                     kchunk);
}

// Cells with ordered extra terms: scalar wall-face flux removal
// (src/scalar_diffusion.f90:451-505), point/volume sources (FluxKernel,
// src/volumesources.f90:757-771) and the wall-model divergence term
// (ScalarExtra, src/scalars.f90:364-379).
// term = {kind | axis<<2, a, b, 0}
//   kind 0: subtract face flux, face lower cell a (linear), axis
//   kind 1: add face flux, face lower cell a (linear), axis
//   kind 2: add src_flux[a]
//   kind 3: add wm_div[a] * S(cell)
template <typename Real>
__global__ void k_scalar_corr(Geo g, const Real *__restrict__ U, const Real *__restrict__ V,
                              const Real *__restrict__ W, const Real *__restrict__ T,
                              const Real *__restrict__ S, Real *__restrict__ Snew, Real *__restrict__ Sadv,
                              const uint8_t *__restrict__ code, const uint8_t *__restrict__ mx,
                              const uint8_t *__restrict__ my, const uint8_t *__restrict__ mz, StageConst sc,
                              TermList tl, const double *__restrict__ src_flux, const double *__restrict__ wm_div) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= tl.n) return;
  const int i = tl.cell[3 * c], j = tl.cell[3 * c + 1], k = tl.cell[3 * c + 2];
  const long p = gix(g, i, j, k);
  // the bulk kernel left scalar_cell of this cell in Snew (which receives the
  // cell's result only here)
  Real a = Snew[p];
  const Real Dd[3] = {1 / (2 * ((Real)g.dxmin * (Real)g.dxmin)), 1 / (2 * ((Real)g.dymin * (Real)g.dymin)),
                      1 / (2 * ((Real)g.dzmin * (Real)g.dzmin))};
  const long st[3] = {1, g.sy, g.sz};
  for (int t = tl.ptr[c]; t < tl.ptr[c + 1]; t++) {
    const int32_t *tt = &tl.term[4 * t];
    const int kind = tt[0] & 3, axis = (tt[0] >> 2) & 3;
    if (kind <= 1) {
      const long q = tt[1], r = q + st[axis];
      const Real fl = (T[r] + T[q]) * (S[r] - S[q]) * Dd[axis];
      a = kind == 0 ? a - fl : a + fl;
    } else if (kind == 2) {
      a = a + (Real)src_flux[tt[1]];
    } else {
      a = a + (Real)wm_div[tt[1]] * S[p];
    }
  }
  scalar_finish<Real>(sc, S, Snew, Sadv, p, a);
}

template <typename Real>
void launch_scalar_corr(const Geo &g, const Real *U, const Real *V, const Real *W, const Real *TDiff,
                        const Real *S, Real *Snew, Real *Sadv, const uint8_t *code, const uint8_t *mx,
                        const uint8_t *my, const uint8_t *mz, StageConst sc, TermList tl, const double *src_flux,
                        const double *wm_div, hipStream_t st) {
  if (!tl.n) return;
  hipLaunchKernelGGL(k_scalar_corr<Real>, dim3(nblk(tl.n, 128)), dim3(128), 0, st, g, U, V, W, TDiff, S, Snew, Sadv, code,
                     mx, my, mz, sc, tl, src_flux, wm_div);
}

// DivergenceWM (2nd order), src/wallmodels.f90:1153-1166
template <typename Real>
__global__ void k_wm_div(Geo g, WMScalarPoints wp, const int32_t *__restrict__ pt, const Real *__restrict__ U_,
                         const Real *__restrict__ V_, const Real *__restrict__ W_) {
  auto U = [&](long q) { return (double)U_[q]; };
  auto V = [&](long q) { return (double)V_[q]; };
  auto W = [&](long q) { return (double)W_[q]; };
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= wp.n) return;
  const int xi = wp.ijk[3 * n], yj = wp.ijk[3 * n + 1], zk = wp.ijk[3 * n + 2];
  const long p = gix(g, xi, yj, zk);
  const long sy = g.sy, sz = g.sz;
  double div = 0;
  if (pt[p + 1] <= 0) div = div + U(p) / g.dxmin;
  if (pt[p - 1] <= 0) div = div - U(p - 1) / g.dxmin;
  if (pt[p + sy] <= 0) div = div + V(p) / g.dymin;
  if (pt[p - sy] <= 0) div = div - V(p - sy) / g.dymin;
  if (pt[p + sz] <= 0) div = div + W(p) / g.dzPr[zk];
  if (pt[p - sz] <= 0) div = div - W(p - sz) / g.dzPr[zk];
  wp.div[n] = div;
}

template <typename Real>
void launch_wm_div(const Geo &g, WMScalarPoints p, const int32_t *prtype, const Real *U, const Real *V, const Real *W,
                   hipStream_t st) {
  if (p.n) hipLaunchKernelGGL(k_wm_div<Real>, dim3(nblk(p.n, 256)), dim3(256), 0, st, g, p, prtype, U, V, W);
}

// TDiff = (Viscosity - molecular_viscosity)/constPr_sgs + molecular_diffusivity,
// src/initial.f90:4302-4313 (constPr_sgs = 0.6 as a default-real literal)
template <typename Real>
__global__ void k_tdiff_init(Geo g, const Real *__restrict__ Visc, Real *__restrict__ TDiff, double nu,
                             double nu_scalar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long p = gix(g, i, j, k);
  TDiff[p] = (Visc[p] - (Real)nu) / (Real)(double)0.6f + (Real)nu_scalar;
}

template <typename Real>
void launch_tdiff_init(const Geo &g, const Real *Visc, Real *TDiff, double nu, double nu_scalar, hipStream_t st) {
  dim3 b(64, 4, 1), gr(nblk(g.nx, 64), nblk(g.ny, 4), g.nz);
  hipLaunchKernelGGL(k_tdiff_init<Real>, gr, b, 0, st, g, Visc, TDiff, nu, nu_scalar);
}

// The temperature in solid cells after its stage (src/scalars.f90:131-207):
// each listed solid cell (within -1..n+2, with fluid neighbours there) takes
// the mean of its fluid neighbours, summed in the order i-1, i+1, j-1, j+1,
// k-1, k+1.  Fluid cells are never written, so the cells are independent.
template <typename Real>
__global__ void k_temp_solid_fix(Geo g, const int32_t *__restrict__ ijk, int n, const int32_t *__restrict__ pt,
                                 Real *__restrict__ T) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = ijk[3 * q], j = ijk[3 * q + 1], k = ijk[3 * q + 2];
  const long p = gix(g, i, j, k), sy = g.sy, sz = g.sz;
  Real t = 0;
  int m = 0;
  if (i > -1 && pt[p - 1] <= 0) { t = t + T[p - 1]; m++; }
  if (i < g.nx + 2 && pt[p + 1] <= 0) { t = t + T[p + 1]; m++; }
  if (j > -1 && pt[p - sy] <= 0) { t = t + T[p - sy]; m++; }
  if (j < g.ny + 2 && pt[p + sy] <= 0) { t = t + T[p + sy]; m++; }
  if (k > -1 && pt[p - sz] <= 0) { t = t + T[p - sz]; m++; }
  if (k < g.nz + 2 && pt[p + sz] <= 0) { t = t + T[p + sz]; m++; }
  if (m > 0) T[p] = t / (Real)m;
}

template <typename Real>
void launch_temp_solid_fix(const Geo &g, const int32_t *ijk, int n, const int32_t *prtype, Real *T, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_temp_solid_fix<Real>, dim3(nblk(n, 256)), dim3(256), 0, st, g, ijk, n, prtype, T);
}

// The moisture in solid cells after its stage (src/scalars.f90:208):
// moisture_ref wherever Prtype > 0 in -2..n+3 (the scalar code's CODE_SOLID,
// set over that box; the row padding has code 0).
template <typename Real>
__global__ void k_moist_fill(long n, const uint8_t *__restrict__ code, Real *__restrict__ q, Real ref) {
  const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n && (code[p] & CODE_SOLID)) q[p] = ref;
}

template <typename Real>
void launch_moist_fill(const Geo &g, const uint8_t *code, Real *q, double ref, hipStream_t st) {
  hipLaunchKernelGGL(k_moist_fill<Real>, dim3(nblk(g.ntot, 256)), dim3(256), 0, st, g.ntot, code, q, (Real)ref);
}

#define ELMM_INST_SCALAR(R)                                                                                        \
  template void launch_scalar<R>(const Geo &, const R *, const R *, const R *, const R *, const R *, R *, R *,     \
                                 const uint8_t *, const uint8_t *, const uint8_t *, const uint8_t *, StageConst,   \
                                 hipStream_t);                                                                    \
  template void launch_scalar_corr<R>(const Geo &, const R *, const R *, const R *, const R *, const R *, R *, R *, \
                                      const uint8_t *, const uint8_t *, const uint8_t *, const uint8_t *,          \
                                      StageConst, TermList, const double *, const double *, hipStream_t);          \
  template void launch_wm_div<R>(const Geo &, WMScalarPoints, const int32_t *, const R *, const R *, const R *,    \
                                 hipStream_t);                                                                    \
  template void launch_tdiff_init<R>(const Geo &, const R *, R *, double, double, hipStream_t);                 \
  template void launch_temp_solid_fix<R>(const Geo &, const int32_t *, int, const int32_t *, R *, hipStream_t);  \
  template void launch_moist_fill<R>(const Geo &, const uint8_t *, R *, double, hipStream_t);                      \
  template void launch_scalar_zm<R>(const Geo &, const R *, const R *, const R *, const R *, const R *, R *, R *,  \
                                    const uint8_t *, const uint8_t *, const uint8_t *, const uint8_t *, StageConst,\
                                    int, hipStream_t);
ELMM_INST_SCALAR(double)
ELMM_INST_SCALAR(float)

}  // namespace elmm
