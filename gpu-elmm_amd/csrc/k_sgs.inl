// k_sgs.inl - Smagorinsky subgrid viscosity and the CFL partial maxima
// (the SubgridModel + TimeStepLength pass).  Included by k_sgs.hip (FP64,
// every operation rounded as in the reference) and k_sgs_f32.hip (precision
// form, ELMM_SGS_RECIP = 1: quotients by the grid spacings as products with
// one reciprocal each, contraction allowed).
#include <cmath>

#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

static inline unsigned sgs_nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------- Smagorinsky
// NuSmag/StrainIJ/Strainu, src/subgrid.f90:55-67, 170-209.  T is the field
// and arithmetic type (double: the reference's FP64; float: precision form).
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
    T v;
    if (sgs_model == 1) {
      const T dx = (T)g.dxmin, dy = (T)g.dymin;
      const T zp1 = (T)g.zPr[k + 1], zm1 = (T)g.zPr[k - 1];
      const T ddz = (T)g.dzPr[k], d2z = 2 * (zp1 - zm1), d4x = 4 * dx, d4y = 4 * dy;
#if ELMM_SGS_RECIP
      // precision form: one reciprocal per divisor, quotients as products
      const T rx = (T)g.invdxf, ry = (T)g.invdyf, rz = (T)1 / ddz, r2z = (T)1 / d2z, r4x = rx * (T)0.25,
              r4y = ry * (T)0.25;
#define ELMM_Q(a, d, r) ((a) * (r))
#else
#define ELMM_Q(a, d, r) ((a) / (d))
#endif
      T D00 = ELMM_Q(U[p] - U[p - 1], dx, rx);
      T D11 = ELMM_Q(V[p] - V[p - sy], dy, ry);
      T D22 = ELMM_Q(W[p] - W[p - sz], ddz, rz);
      T D01 = ELMM_Q(U[p + sy] + U[p - 1 + sy] - U[p - sy] - U[p - 1 - sy], d4y, r4y);
      T D02 = ELMM_Q(U[p + sz] + U[p - 1 + sz] - U[p - sz] - U[p - 1 - sz], d2z, r2z);
      T D10 = ELMM_Q(V[p + 1] + V[p + 1 - sy] - V[p - 1] - V[p - 1 - sy], d4x, r4x);
      T D12 = ELMM_Q(V[p + sz] + V[p - sy + sz] - V[p - sz] - V[p - sy - sz], d2z, r2z);
      T D20 = ELMM_Q(W[p + 1] + W[p + 1 - sz] - W[p - 1] - W[p - 1 - sz], d4x, r4x);
      T D21 = ELMM_Q(W[p + sy] + W[p + sy - sz] - W[p - sy] - W[p - sy - sz], d4y, r4y);
#undef ELMM_Q
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
      // filter width: table from the host libm (the cube root of a static
      // geometric factor per column class and level)
      T width;
      if (width_t.tab) {
        const int cls = width_t.idx ? width_t.idx[(i - 1) + (long)g.nx * (j - 1)] : 0;
        width = (T)width_t.tab[cls * width_t.ld + k];
      } else {
        width = (T)(filter_ratio * pow(g.dxPr[i] * g.dyPr[j] * g.dzPr[k], 1.0 / 3.0));
      }
      T wc = width * c_smag;
      v = Sbar * (wc * wc);
      if (nu > 0) v = v + nu;
    } else {
      v = nu;
    }
    Visc[p] = v;
    if (cfl_partial) {
      // TimeStepLength's U_dx_max, src/dynamics.f90:556-594 (NaN -> +inf)
      double a1 = fabs((double)U[p]), a0 = fabs((double)U[p - 1]);
      double b1 = fabs((double)V[p]), b0 = fabs((double)V[p - sy]);
      double c1 = fabs((double)W[p]), c0 = fabs((double)W[p - sz]);
      double dxk = g.variable_z ? g.dxPr[i] : g.dxmin;
      double dyk = g.variable_z ? g.dyPr[j] : g.dymin;
      double dzk = g.variable_z ? g.dzPr[k] : g.dzmin;
      double q = (a1 > a0 ? a1 : a0) / dxk;
      q = q + (b1 > b0 ? b1 : b0) / dyk;
      q = q + (c1 > c0 ? c1 : c0) / dzk;
      cfl = (q != q) ? INFINITY : q;
    }
  }
  if (cfl_partial) {
    // wavefront max via cross-lane shuffles, then LDS across the 4 waves
    for (int off = 32; off > 0; off >>= 1) cfl = fmax(cfl, __shfl_xor(cfl, off));
    __shared__ double red[4];
    const int lane = (threadIdx.y * blockDim.x + threadIdx.x);
    if ((lane & 63) == 0) red[lane >> 6] = cfl;
    __syncthreads();
    if (lane == 0) {
      double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      cfl_partial[(long)blockIdx.z * gridDim.y * gridDim.x + blockIdx.y * gridDim.x + blockIdx.x] = m;
    }
  }
}

template <typename T>
void launch_sgs(const Geo &g, const T *U, const T *V, const T *W, T *Visc, double nu, double c_smag,
                double filter_ratio, SgsWidth width, int sgs_model, double *cfl_partial, int, hipStream_t st) {
  dim3 b(64, 4, 1), gr(sgs_nblk(g.nx, 64), sgs_nblk(g.ny, 4), g.nz);
  hipLaunchKernelGGL(k_sgs<T>, gr, b, 0, st, g, U, V, W, Visc, nu, c_smag, filter_ratio, width, sgs_model,
                     cfl_partial);
}

}  // namespace elmm
