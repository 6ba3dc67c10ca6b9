// k_multigrid.hip - geometric multigrid solve of the pressure equation
// (poisson_solver = ELMM_POISSON_MG), the role of PoissMG
// (src/multigrid.f90:1047-1172: V-cycle MG_CGC 994-1044, smoother MG_GS
// 786-910, residual MG_res 912-977, Restrict 114-165, Prolongate 58-111).
//
// The operator is the cell-centred one of Residue_variable_z
// (src/poisson.f90:23-60) that the spectral solver inverts: periodic x and
// y, staggered Neumann walls in z, stretched z.  Levels coarsen x and y by 2
// and keep every z level, so the z coefficients are the same on all of them
// and a z-line smoother handles the strong vertical coupling of stretched
// grids.  Arrays of a level are dense, nx x ny x nz doubles, x fastest.
//
// Operation order (shared with the oracle, so the two agree bit for bit):
//   sweep (one colour, (i + j) & 1): per column, h_k = ax (P_w + P_e) +
//     ay (P_s + P_n), g_k = ((h_k - rhs_k) + az_k g_{k-1}) / den_k, then
//     P_k = g_k + e_k P_{k+1} from the top; den_k, e_k are host tables;
//   residual r = rhs - s, s = ((P_w - P) + (P_e - P)) ax + ((P_s - P) +
//     (P_n - P)) ay [+ az_k (P_b - P)] [+ cz_k (P_t - P)];
//   restriction rhs_c = ((r1 + r2) + (r3 + r4)) * 0.25 (children in x
//     pairs, lower y first), P_c = 0;
//   prolongation P += 0.75 (0.75 c + 0.25 c_x) + 0.25 (0.75 c_y + 0.25 c_xy);
//   coarsest level: exact, by discrete Fourier sums in x and y and a
//     tridiagonal elimination in z per mode (k_mg_coarse).
#include <cmath>

#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

namespace {

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ inline int wrap_lo(int i, int n) { return i > 1 ? i - 1 : n; }
__device__ inline int wrap_hi(int i, int n) { return i < n ? i + 1 : 1; }

// residual of cell (i, j, k) (1-based) of a level
__device__ inline double mg_res(const MgLevel &L, const double *__restrict__ az, const double *__restrict__ cz, int i,
                                int j, int k) {
  const int nx = L.nx, ny = L.ny;
  const long plane = (long)nx * ny, o = (long)(k - 1) * plane;
  const double *P = L.phi;
  const long c = (i - 1) + (long)(j - 1) * nx + o;
  const double p = P[c];
  double s = ((P[(wrap_lo(i, nx) - 1) + (long)(j - 1) * nx + o] - p) + (P[(wrap_hi(i, nx) - 1) + (long)(j - 1) * nx + o] - p)) *
                 L.ax +
             ((P[(i - 1) + (long)(wrap_lo(j, ny) - 1) * nx + o] - p) + (P[(i - 1) + (long)(wrap_hi(j, ny) - 1) * nx + o] - p)) *
                 L.ay;
  if (k > 1) s = s + az[k - 1] * (P[c - plane] - p);
  if (k < L.nz) s = s + cz[k - 1] * (P[c + plane] - p);
  return L.rhs[c] - s;
}

}  // namespace

// Right-hand side of the pressure equation (PrePoisson, src/pressure.f90:
// 657-717): ((U_i - U_{i-1})/dx + (V_j - V_{j-1})/dy + (W_k - W_{k-1})/dz
// - Q) / coef, into the dense level-0 array and the ghosted RHS.
template <typename T>
__global__ void __launch_bounds__(256) k_mg_rhs(Geo g, const T *__restrict__ U, const T *__restrict__ V,
                                                const T *__restrict__ W, const double *__restrict__ Q, double dt2_rec,
                                                double *__restrict__ rhs0, double *__restrict__ RHS) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  const int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long)g.nx * g.ny)) + 1;
  const long p = gix(g, i, j, k);
  const double dz = g.variable_z ? g.dzPr[k] : g.dzmin;
  double r = ((double)U[p] - (double)U[p - 1]) / g.dxmin + ((double)V[p] - (double)V[p - g.sy]) / g.dymin +
             ((double)W[p] - (double)W[p - g.sz]) / dz;
  if (Q) r = r - Q[p];
  r = r * dt2_rec;
  rhs0[t] = r;
  RHS[p] = r;
}

// Level-0 solution into the ghosted Phi interior.
__global__ void __launch_bounds__(256) k_mg_to_phi(Geo g, const double *__restrict__ phi0, double *__restrict__ Phi) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)g.nx * g.ny * g.nz) return;
  const int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long)g.nx * g.ny)) + 1;
  Phi[gix(g, i, j, k)] = phi0[t];
}

__global__ void __launch_bounds__(256) k_mg_zero(double *__restrict__ a, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.0;
}

// Line Gauss-Seidel sweep of one colour: one thread per column with (i + j)
// & 1 == color, Thomas elimination up the column (the forward values stored
// in place: the same colour's columns are not neighbours), then back down.
__global__ void __launch_bounds__(256) k_mg_sweep(MgLevel L, const double *__restrict__ az,
                                                  const double *__restrict__ den, const double *__restrict__ e,
                                                  int color) {
  const int half = L.nx / 2;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= half * L.ny) return;
  const int j = t / half + 1, ii = t - (j - 1) * half;
  const int i = 2 * ii + (((j + color) & 1) ? 1 : 2);
  const int nx = L.nx, nz = L.nz;
  const long plane = (long)nx * L.ny;
  const long c = (i - 1) + (long)(j - 1) * nx;
  const long w = (wrap_lo(i, nx) - 1) + (long)(j - 1) * nx, ea = (wrap_hi(i, nx) - 1) + (long)(j - 1) * nx;
  const long so = (i - 1) + (long)(wrap_lo(j, L.ny) - 1) * nx, no = (i - 1) + (long)(wrap_hi(j, L.ny) - 1) * nx;
  double *P = L.phi;
  const double *R = L.rhs;
  double gv = 0;
#pragma unroll 4
  for (int k = 1; k <= nz; k++) {
    const long o = (long)(k - 1) * plane;
    const double h = L.ax * (P[w + o] + P[ea + o]) + L.ay * (P[so + o] + P[no + o]);
    gv = ((h - R[c + o]) + az[k - 1] * gv) / den[k - 1];
    P[c + o] = gv;
  }
  double x = gv;
#pragma unroll 4
  for (int k = nz - 1; k >= 1; k--) {
    const long o = (long)(k - 1) * plane;
    x = P[c + o] + e[k - 1] * x;
    P[c + o] = x;
  }
}

// Maximum |residual| of a level: block maxima (a NaN counts as +inf).
__global__ void __launch_bounds__(256) k_mg_resmax(MgLevel L, const double *__restrict__ az,
                                                   const double *__restrict__ cz, double *__restrict__ partial) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0;
  if (t < (long)L.nx * L.ny * L.nz) {
    const int i = (int)(t % L.nx) + 1, j = (int)((t / L.nx) % L.ny) + 1, k = (int)(t / ((long)L.nx * L.ny)) + 1;
    v = fabs(mg_res(L, az, cz, i, j, k));
    if (v != v) v = INFINITY;
  }
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
  __shared__ double red[4];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int q = 1; q < 4; q++) m = fmax(m, red[q]);
    partial[blockIdx.x] = m;
  }
}

// Residual of the fine level restricted to the coarse one (coarse P = 0).
__global__ void __launch_bounds__(256) k_mg_restrict(MgLevel F, MgLevel Cs, const double *__restrict__ az,
                                                     const double *__restrict__ cz) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)Cs.nx * Cs.ny * Cs.nz) return;
  const int I = (int)(t % Cs.nx) + 1, J = (int)((t / Cs.nx) % Cs.ny) + 1, k = (int)(t / ((long)Cs.nx * Cs.ny)) + 1;
  const double r1 = mg_res(F, az, cz, 2 * I - 1, 2 * J - 1, k);
  const double r2 = mg_res(F, az, cz, 2 * I, 2 * J - 1, k);
  const double r3 = mg_res(F, az, cz, 2 * I - 1, 2 * J, k);
  const double r4 = mg_res(F, az, cz, 2 * I, 2 * J, k);
  Cs.rhs[t] = ((r1 + r2) + (r3 + r4)) * 0.25;
  Cs.phi[t] = 0.0;
}

// Bilinear (x, y) prolongation of the coarse correction, added to the fine P.
__global__ void __launch_bounds__(256) k_mg_prolong(MgLevel Cs, MgLevel F) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)F.nx * F.ny * F.nz) return;
  const int i = (int)(t % F.nx) + 1, j = (int)((t / F.nx) % F.ny) + 1, k = (int)(t / ((long)F.nx * F.ny)) + 1;
  const int I = (i + 1) / 2, J = (j + 1) / 2;
  const int Ix = (i & 1) ? wrap_lo(I, Cs.nx) : wrap_hi(I, Cs.nx);
  const int Jy = (j & 1) ? wrap_lo(J, Cs.ny) : wrap_hi(J, Cs.ny);
  const double *E = Cs.phi + (long)(k - 1) * Cs.nx * Cs.ny;
  const double c00 = E[(I - 1) + (long)(J - 1) * Cs.nx], c10 = E[(Ix - 1) + (long)(J - 1) * Cs.nx];
  const double c01 = E[(I - 1) + (long)(Jy - 1) * Cs.nx], c11 = E[(Ix - 1) + (long)(Jy - 1) * Cs.nx];
  const double ex0 = 0.75 * c00 + 0.25 * c10, ex1 = 0.75 * c01 + 0.25 * c11;
  F.phi[t] = F.phi[t] + (0.75 * ex0 + 0.25 * ex1);
}

// Exact solve of the coarsest level (P = 0 on entry) by one block: discrete
// Fourier sums in x, then y (twiddles w(q) = (cos 2 pi q/N, -sin 2 pi q/N)),
// a tridiagonal elimination in z per mode (the (0, 0) mode integrated as the
// singular Neumann problem with x_1 = 0), conjugate sums over y, then x.
// Work arrays Y, X (complex, nx ny nz) and cp (nx ny nz) in global memory.
__global__ void __launch_bounds__(256) k_mg_coarse(MgLevel L, MgCoarse C, const double *__restrict__ az,
                                                   const double *__restrict__ cz, const double *__restrict__ dzPr,
                                                   const double *__restrict__ dzW) {
  const int nx = L.nx, ny = L.ny, nz = L.nz;
  const long plane = (long)nx * ny, n = plane * nz;
  double2 *Y = C.Y, *X = C.X;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    const int m = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / plane);
    double ar = 0, ai = 0;
    for (int i = 0; i < nx; i++) {
      const int tq = (m * i) % nx;
      const double r = L.rhs[i + (long)j * nx + (long)k * plane];
      const double pr = r * C.cx[tq] - 0.0 * C.sx[tq], pi = r * C.sx[tq] + 0.0 * C.cx[tq];
      ar = ar + pr;
      ai = ai + pi;
    }
    Y[q] = make_double2(ar, ai);
  }
  __syncthreads();
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    const int m = (int)(q % nx), nn = (int)((q / nx) % ny), k = (int)(q / plane);
    double ar = 0, ai = 0;
    for (int j = 0; j < ny; j++) {
      const int tq = (nn * j) % ny;
      const double2 y = Y[m + (long)j * nx + (long)k * plane];
      const double pr = y.x * C.cy[tq] - y.y * C.sy[tq], pi = y.x * C.sy[tq] + y.y * C.cy[tq];
      ar = ar + pr;
      ai = ai + pi;
    }
    X[q] = make_double2(ar, ai);
  }
  __syncthreads();
  for (int mode = threadIdx.x; mode < plane; mode += blockDim.x) {
    const int m = mode % nx, nn = mode / nx;
    double2 *col = X + mode;
    if (m == 0 && nn == 0) {
      double Gr = 0, Gi = 0, xr = 0, xi = 0;
      for (int k = 1; k <= nz; k++) {
        const double2 d = col[(long)(k - 1) * plane];
        col[(long)(k - 1) * plane] = make_double2(xr, xi);
        if (k < nz) {
          Gr = Gr + d.x * dzPr[k];
          Gi = Gi + d.y * dzPr[k];
          xr = xr + Gr * dzW[k];
          xi = xi + Gi * dzW[k];
        }
      }
      continue;
    }
    const double lam = C.lam[mode];  // host table (the host's sin)
    double cprev = 0, dr = 0, di = 0;
    for (int k = 0; k < nz; k++) {
      const double b = (lam - az[k]) - cz[k];
      const double den = b - az[k] * cprev;
      const double cpk = cz[k] / den;
      C.cp[mode + (long)k * plane] = cpk;
      const double2 d = col[(long)k * plane];
      dr = (d.x - az[k] * dr) / den;
      di = (d.y - az[k] * di) / den;
      col[(long)k * plane] = make_double2(dr, di);
      cprev = cpk;
    }
    double xr = dr, xi = di;
    for (int k = nz - 2; k >= 0; k--) {
      const double2 dp = col[(long)k * plane];
      const double cpk = C.cp[mode + (long)k * plane];
      xr = dp.x - cpk * xr;
      xi = dp.y - cpk * xi;
      col[(long)k * plane] = make_double2(xr, xi);
    }
  }
  __syncthreads();
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    const int m = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / plane);
    double ar = 0, ai = 0;
    for (int nn = 0; nn < ny; nn++) {
      const int tq = (nn * j) % ny;
      const double2 x = X[m + (long)nn * nx + (long)k * plane];
      const double wr = C.cy[tq], wi = -C.sy[tq];
      const double pr = x.x * wr - x.y * wi, pi = x.x * wi + x.y * wr;
      ar = ar + pr;
      ai = ai + pi;
    }
    Y[q] = make_double2(ar, ai);
  }
  __syncthreads();
  const double scale = 1.0 / (double)plane;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    const int i = (int)(q % nx), j = (int)((q / nx) % ny), k = (int)(q / plane);
    double acc = 0;
    for (int m = 0; m < nx; m++) {
      const int tq = (m * i) % nx;
      const double2 y = Y[m + (long)j * nx + (long)k * plane];
      const double wr = C.cx[tq], wi = -C.sx[tq];
      acc = acc + (y.x * wr - y.y * wi);
    }
    L.phi[q] = acc * scale;
  }
}

template <typename T>
void launch_mg_rhs(const Geo &g, const T *U, const T *V, const T *W, const double *Q, double dt2_rec, double *rhs0,
                   double *RHS, hipStream_t st) {
  const long n = (long)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_mg_rhs<T>, dim3(nblk(n, 256)), dim3(256), 0, st, g, U, V, W, Q, dt2_rec, rhs0, RHS);
}

void launch_mg_to_phi(const Geo &g, const double *phi0, double *Phi, hipStream_t st) {
  const long n = (long)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_mg_to_phi, dim3(nblk(n, 256)), dim3(256), 0, st, g, phi0, Phi);
}

void launch_mg_zero(double *a, long n, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_zero, dim3(nblk(n, 256)), dim3(256), 0, st, a, n);
}

void launch_mg_sweep(const MgLevel &L, const double *az, const double *den, const double *e, int color,
                     hipStream_t st) {
  const long cols = (long)(L.nx / 2) * L.ny;
  hipLaunchKernelGGL(k_mg_sweep, dim3(nblk(cols, 256)), dim3(256), 0, st, L, az, den, e, color);
}

int mg_resmax_blocks(const MgLevel &L) { return (int)nblk((long)L.nx * L.ny * L.nz, 256); }

void launch_mg_resmax(const MgLevel &L, const double *az, const double *cz, double *partial, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_resmax, dim3(mg_resmax_blocks(L)), dim3(256), 0, st, L, az, cz, partial);
}

void launch_mg_restrict(const MgLevel &F, const MgLevel &Cs, const double *az, const double *cz, hipStream_t st) {
  const long n = (long)Cs.nx * Cs.ny * Cs.nz;
  hipLaunchKernelGGL(k_mg_restrict, dim3(nblk(n, 256)), dim3(256), 0, st, F, Cs, az, cz);
}

void launch_mg_prolong(const MgLevel &Cs, const MgLevel &F, hipStream_t st) {
  const long n = (long)F.nx * F.ny * F.nz;
  hipLaunchKernelGGL(k_mg_prolong, dim3(nblk(n, 256)), dim3(256), 0, st, Cs, F);
}

void launch_mg_coarse(const MgLevel &L, const MgCoarse &C, const double *az, const double *cz, const double *dzPr,
                      const double *dzW, hipStream_t st) {
  hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(256), 0, st, L, C, az, cz, dzPr, dzW);
}

template void launch_mg_rhs<double>(const Geo &, const double *, const double *, const double *, const double *,
                                    double, double *, double *, hipStream_t);
template void launch_mg_rhs<float>(const Geo &, const float *, const float *, const float *, const double *, double,
                                   double *, double *, hipStream_t);

}  // namespace elmm
