// k_fft_mr.hip - transform passes of the spectral Poisson solve for line
// lengths n = 2^a 3^b that are not powers of two (the batched-ensemble grids
// such as 96, 144 or 192 cells).
//
// The lines are transformed in LDS by a self-sorting (Stockham) mixed-radix
// FFT: radix-3 stages first, then radix-2 stages; stage s with radix r and
// span Ns (the product of the earlier radices) maps element j + q n/r
// (q < r) of the source, multiplied by w^(k q n/(Ns r)) with k = j mod Ns and
// w = exp(-2 pi i/n) (conjugate for the inverse), through an r-point DFT to
// element (j div Ns) Ns r + k + q Ns of the destination.  Every operation
// is stated so the oracle can repeat it bit for bit: complex products
// (a.x w.x - a.y w.y, a.x w.y + a.y w.x), no product when k q = 0, and the
// 3-point DFT  s = b + c, d = b - c, y0 = a + s, m = a - s/2,
// t = (s3 d.y, -s3 d.x) (forward) or (-s3 d.y, s3 d.x) (inverse) with
// s3 = sin(pi/3), y1 = m + t, y2 = m - t.
// The passes have the same inputs and outputs as the power-of-two kernels
// in k_pressure.hip (divergence -> x FFT, y FFT, inverse y, inverse x ->
// Phi), with rows mapped by division since ny need not be a power of two.
#include <cmath>

#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

namespace {

inline unsigned nblk(long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ inline double2 mk2(double x, double y) { return make_double2(x, y); }
__device__ inline float2 mk2(float x, float y) { return make_float2(x, y); }
template <typename V>
__device__ inline V cmul(V a, V b) {
  return mk2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

constexpr int kMrThreads = 256;
constexpr int kMrElems = 1024;  // line elements per block and buffer

template <typename R>
__device__ inline R sin60() {
  return (R)0.86602540378443864676;
}

// One Stockham stage over `lines` lines of length n (line l at l * n in
// src and dst).
template <typename R>
__device__ inline void mr_stage(const C2<R> *src, C2<R> *dst, int n, int lines, int radix, int ns,
                                const C2<R> *__restrict__ tw, bool inverse) {
  using V = C2<R>;
  const int nb = n / radix, tstride = n / (ns * radix);
  for (int q = threadIdx.x; q < lines * nb; q += blockDim.x) {
    const int l = q / nb, j = q - l * nb;
    const V *s = src + l * n;
    V *d = dst + l * n;
    const int k = j % ns;
    V v[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      if (r < radix) {
        v[r] = s[j + r * nb];
        if (r > 0 && k > 0) {
          V w = tw[k * r * tstride];
          if (inverse) w.y = -w.y;
          v[r] = cmul(v[r], w);
        }
      }
    }
    const int base = (j / ns) * ns * radix + k;
    if (radix == 2) {
      d[base] = mk2(v[0].x + v[1].x, v[0].y + v[1].y);
      d[base + ns] = mk2(v[0].x - v[1].x, v[0].y - v[1].y);
    } else {
      const R s3 = sin60<R>(), h = (R)0.5;
      const V sm = mk2(v[1].x + v[2].x, v[1].y + v[2].y), df = mk2(v[1].x - v[2].x, v[1].y - v[2].y);
      const V m = mk2(v[0].x - h * sm.x, v[0].y - h * sm.y);
      const V t = inverse ? mk2(-(s3 * df.y), s3 * df.x) : mk2(s3 * df.y, -(s3 * df.x));
      d[base] = mk2(v[0].x + sm.x, v[0].y + sm.y);
      d[base + ns] = mk2(m.x + t.x, m.y + t.y);
      d[base + 2 * ns] = mk2(m.x - t.x, m.y - t.y);
    }
  }
}

// Transforms `lines` lines held in a (natural order); returns the buffer
// (a or b) that holds the result.
template <typename R>
__device__ inline C2<R> *mr_fft(C2<R> *a, C2<R> *b, int n, int lines, const C2<R> *__restrict__ tw, bool inverse) {
  int rest = n, ns = 1;
  while (rest > 1) {
    const int radix = rest % 3 == 0 ? 3 : 2;
    __syncthreads();
    mr_stage<R>(a, b, n, lines, radix, ns, tw, inverse);
    C2<R> *t = a;
    a = b;
    b = t;
    ns *= radix;
    rest /= radix;
  }
  __syncthreads();
  return a;
}

template <typename R>
__device__ inline void mr_twiddles(C2<R> *dst, const C2<R> *__restrict__ src, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = src[t];
}

}  // namespace

// Pass 1 (or a refinement sweep's first pass): right-hand side of `rows`
// (j, k) rows per block, x FFT, modes 0..M-1 to the spectral array.
// mode: kMrDiv (divergence of U, V, W, minus Q where qmask says so, over
// coef; stored to rhs_out when given), kMrIn (rhs_in), kMrResid (rhs_in - A
// phi).
enum { kMrDiv = 0, kMrIn = 1, kMrResid = 2 };

template <typename T, typename R>
__global__ void __launch_bounds__(kMrThreads)
    k_fft_x_fwd_mr(Geo g, PoissonPlan<R> pp, const T *__restrict__ U, const T *__restrict__ V,
                   const T *__restrict__ W, const double *__restrict__ Q, const uint32_t *__restrict__ qmask,
                   const double *__restrict__ rhs_in, const double *__restrict__ phi, double dt2_rec,
                   double *__restrict__ rhs_out, int mode, int rows) {
  using C = C2<R>;
  extern __shared__ __align__(16) unsigned char sh_raw[];
  const int n = pp.nx;
  C *tw = reinterpret_cast<C *>(sh_raw);
  C *a = tw + n, *b = a + rows * n;
  mr_twiddles<R>(tw, pp.twx, n);
  const int nrows = pp.ny * pp.nz, row0 = blockIdx.x * rows;
  for (int e = threadIdx.x; e < rows * n; e += blockDim.x) {
    const int l = e / n, i = e - l * n;
    const int row = min(row0 + l, nrows - 1);
    const int j = row % pp.ny + 1, k = row / pp.ny + 1;
    const long p = gix(g, i + 1, j, k);
    double r;
    if (mode == kMrResid) {
      const double rx = 1 / (g.dxmin * g.dxmin), ry = 1 / (g.dymin * g.dymin);
      const double ph = phi[p];
      const double pw = phi[i + 1 > 1 ? p - 1 : p - 1 + g.nx], pe = phi[i + 1 < g.nx ? p + 1 : p + 1 - g.nx];
      const double ps = phi[j > 1 ? p - g.sy : p - g.sy + (long)g.ny * g.sy];
      const double pn = phi[j < g.ny ? p + g.sy : p + g.sy - (long)g.ny * g.sy];
      double s = ((pw - ph) + (pe - ph)) * rx + ((ps - ph) + (pn - ph)) * ry;
      if (k > 1) s = s + pp.az[k - 1] * (phi[p - g.sz] - ph);
      if (k < g.nz) s = s + pp.cz[k - 1] * (phi[p + g.sz] - ph);
      r = rhs_in[p] - s;
    } else if (mode == kMrIn) {
      r = rhs_in[p];
    } else {
      const double du = (double)U[p] - (double)U[p - 1], dv = (double)V[p] - (double)V[p - g.sy];
      const double dw = (double)W[p] - (double)W[p - g.sz];
      r = div_rn(du, g.dxmin, g.rdx) + div_rn(dv, g.dymin, g.rdy) +
          (g.variable_z ? dw / g.dzPr[k] : div_rn(dw, g.dzmin, g.rdz));
      if (Q && ((qmask[row] >> (i >> 6)) & 1u)) r = r - Q[p];
      r = r * dt2_rec;
      if (rhs_out && row0 + l < nrows) rhs_out[p] = r;
    }
    a[e] = mk2((R)r, (R)0);
  }
  const C *res = mr_fft<R>(a, b, n, rows, tw, false);
  for (int e = threadIdx.x; e < rows * pp.M; e += blockDim.x) {
    const int l = e / pp.M, m = e - l * pp.M;
    if (row0 + l < nrows) pp.X[(long)(row0 + l) * pp.Mp + m] = res[l * n + m];
  }
}

// Passes 2 and 4: y FFT of `cols` consecutive x-modes of one z-plane.
template <typename R>
__global__ void __launch_bounds__(kMrThreads) k_fft_y_mr(PoissonPlan<R> pp, int inverse, int cols) {
  using C = C2<R>;
  extern __shared__ __align__(16) unsigned char sh_raw[];
  const int n = pp.ny;
  C *tw = reinterpret_cast<C *>(sh_raw);
  C *a = tw + n, *b = a + cols * n;
  mr_twiddles<R>(tw, pp.twy, n);
  const int m0 = blockIdx.x * cols, k = blockIdx.y;
  C *X = pp.X + (long)k * n * pp.Mp;
  for (int e = threadIdx.x; e < cols * n; e += blockDim.x) {
    const int j = e / cols, c = e - j * cols;  // consecutive threads: consecutive modes of one row
    const int m = min(m0 + c, pp.M - 1);
    a[c * n + j] = m0 + c < pp.M ? X[(long)j * pp.Mp + m] : mk2((R)0, (R)0);
  }
  const C *res = mr_fft<R>(a, b, n, cols, tw, inverse != 0);
  for (int e = threadIdx.x; e < cols * n; e += blockDim.x) {
    const int j = e / cols, c = e - j * cols;
    if (m0 + c < pp.M) X[(long)j * pp.Mp + m0 + c] = res[c * n + j];
  }
}

// Pass 5: inverse x FFT of the half spectrum (Hermitian fill), real part
// scaled by 1/(nx ny) to Phi (added with accumulate).
template <typename R>
__global__ void __launch_bounds__(kMrThreads)
    k_fft_x_inv_mr(Geo g, PoissonPlan<R> pp, double *__restrict__ Phi, int accumulate, int rows) {
  using C = C2<R>;
  extern __shared__ __align__(16) unsigned char sh_raw[];
  const int n = pp.nx;
  C *tw = reinterpret_cast<C *>(sh_raw);
  C *a = tw + n, *b = a + rows * n;
  mr_twiddles<R>(tw, pp.twx, n);
  const int nrows = pp.ny * pp.nz, row0 = blockIdx.x * rows, M = pp.M;
  for (int e = threadIdx.x; e < rows * n; e += blockDim.x) {
    const int l = e / n, m = e - l * n;
    const int row = min(row0 + l, nrows - 1);
    C v = pp.X[(long)row * pp.Mp + (m < M ? m : n - m)];
    if (m >= M) v.y = -v.y;
    a[e] = v;
  }
  const C *res = mr_fft<R>(a, b, n, rows, tw, true);
  const R scale = (R)1.0 / ((R)n * pp.ny);
  for (int e = threadIdx.x; e < rows * n; e += blockDim.x) {
    const int l = e / n, i = e - l * n;
    const int row = row0 + l;
    if (row < nrows) {
      const int j = row % pp.ny + 1, k = row / pp.ny + 1;
      const long q = gix(g, i + 1, j, k);
      const double val = (double)(res[e].x * scale);
      Phi[q] = accumulate ? Phi[q] + val : val;
    }
  }
}

static int mr_rows(int n) { return n >= kMrElems ? 1 : kMrElems / n; }

template <typename R>
static size_t mr_lds(int n, int lines) {
  return ((size_t)n + 2 * (size_t)lines * n) * sizeof(C2<R>);
}

template <typename T, typename R>
void launch_fft_x_fwd_mr(const Geo &g, const PoissonPlan<R> &pp, const T *U, const T *V, const T *W, const double *Q,
                         const uint32_t *qmask, const double *rhs, const double *phi, double dt2_rec,
                         double *rhs_out, hipStream_t st) {
  const int rows = mr_rows(pp.nx), nrows = pp.ny * pp.nz;
  const int mode = phi ? kMrResid : rhs ? kMrIn : kMrDiv;
  const size_t sh = mr_lds<R>(pp.nx, rows);
  if (sh > 65536) (void)lds_limit_raise((const void *)k_fft_x_fwd_mr<T, R>, sh);
  hipLaunchKernelGGL((k_fft_x_fwd_mr<T, R>), dim3(nblk(nrows, rows)), dim3(kMrThreads), sh, st, g, pp, U, V, W, Q,
                     qmask, rhs, phi, dt2_rec, rhs_out, mode, rows);
}

template <typename R>
void launch_fft_y_mr(const PoissonPlan<R> &pp, int inverse, hipStream_t st) {
  const int cols = std::max(1, std::min(8, kMrElems / pp.ny));
  const size_t sh = mr_lds<R>(pp.ny, cols);
  if (sh > 65536) (void)lds_limit_raise((const void *)k_fft_y_mr<R>, sh);
  hipLaunchKernelGGL((k_fft_y_mr<R>), dim3(nblk(pp.M, cols), pp.nz), dim3(kMrThreads), sh, st, pp, inverse, cols);
}

template <typename R>
void launch_fft_x_inv_mr(const Geo &g, const PoissonPlan<R> &pp, double *Phi, int accumulate, hipStream_t st) {
  const int rows = mr_rows(pp.nx), nrows = pp.ny * pp.nz;
  const size_t sh = mr_lds<R>(pp.nx, rows);
  if (sh > 65536) (void)lds_limit_raise((const void *)k_fft_x_inv_mr<R>, sh);
  hipLaunchKernelGGL((k_fft_x_inv_mr<R>), dim3(nblk(nrows, rows)), dim3(kMrThreads), sh, st, g, pp, Phi, accumulate,
                     rows);
}

#define ELMM_INST_MR_R(R)                                                                                          \
  template void launch_fft_y_mr<R>(const PoissonPlan<R> &, int, hipStream_t);                                     \
  template void launch_fft_x_inv_mr<R>(const Geo &, const PoissonPlan<R> &, double *, int, hipStream_t);
ELMM_INST_MR_R(double)
ELMM_INST_MR_R(float)
#define ELMM_INST_MR_TR(T, R)                                                                                      \
  template void launch_fft_x_fwd_mr<T, R>(const Geo &, const PoissonPlan<R> &, const T *, const T *, const T *,     \
                                          const double *, const uint32_t *, const double *, const double *, double,  \
                                          double *, hipStream_t);
ELMM_INST_MR_TR(double, double)
ELMM_INST_MR_TR(float, double)
ELMM_INST_MR_TR(float, float)

}  // namespace elmm
