// k_monitor.hip - per-step diagnostics kept in device memory: the CFL
// number of TimeStepLength with the divergence test of the main loop, the
// standalone CFL scan of variable time steps, and TMarchRK3's delta.
//
// Every kernel reads and writes the handle's StepMon record only, with
// launch arguments that are constant for a configuration, so the kernels can
// sit inside captured step graphs; the host reads the record after a
// synchronisation.
#include <cmath>

#include "device_common.hpp"
#include "kernels.hpp"

namespace elmm {

namespace {

// Block reduction over 256 threads; thread 0 gets the result.
template <typename Op>
__device__ inline double block_reduce(double v, Op op) {
  for (int off = 32; off > 0; off >>= 1) v = op(v, __shfl_xor(v, off));
  __shared__ double red[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) red[wv] = v;
  __syncthreads();
  double s = red[0];
  if (threadIdx.x == 0)
    for (int w = 1; w < 4; w++) s = op(s, red[w]);
  __syncthreads();
  return s;
}

__device__ inline double nan_max(double a, double b) {
  if (a != a || b != b) return INFINITY;
  return a > b ? a : b;
}

}  // namespace

// TimeStepLength's maximum (src/dynamics.f90:556-591) from the block maxima
// the SGS kernel wrote at the first stage, then main.f90's divergence test
// (src/main.f90:161-168: time_step >= 3, CFL > CFL_max; a NaN velocity makes
// the maximum huge, src/dynamics.f90:538).
__global__ void __launch_bounds__(256)
    k_cfl_monitor(const double *__restrict__ partial, int nblocks, double dt, double cfl_max, StepMon *mon) {
  double v = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) v = nan_max(v, partial[b]);
  const double m = block_reduce(v, nan_max);
  if (threadIdx.x != 0) return;
  const double cfl = m * dt;
  const long long step = mon->step + 1;
  mon->step = step;
  mon->cfl_last = cfl;
  mon->umax_last = m;
  if (cfl_max > 0 && step >= 3 && !(cfl <= cfl_max) && mon->first_bad < 0) mon->first_bad = step;
}

void launch_cfl_monitor(const double *partial, int nblocks, double dt, double cfl_max, StepMon *mon,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_cfl_monitor, dim3(1), dim3(256), 0, st, partial, nblocks, dt, cfl_max, mon);
}

// TMarchRK3's delta (src/tsteps.f90:188-207): per stage the mean over each
// component's interior of |U - U2| (U the stage-start velocity, U2 the
// projected one), summed over the components and stages, divided by dt at
// the last stage.  Each component is summed over its own interior, thread t
// taking cell t in x-fastest order: per block the 64-lane butterfly sums of
// the four wavefronts, added in order; the block sums are then added in
// block order by one thread.
template <typename T>
__global__ void __launch_bounds__(256)
    k_delta_partial(Geo g, const T *__restrict__ U, const T *__restrict__ U2, const T *__restrict__ V,
                    const T *__restrict__ V2, const T *__restrict__ W, const T *__restrict__ W2,
                    double *__restrict__ partial, int nblocks) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const T *a[3] = {U, V, W}, *b[3] = {U2, V2, W2};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int n0 = g.n[c][0], n1 = g.n[c][1], n2 = g.n[c][2];
    double v = 0;
    if (t < (long)n0 * n1 * n2) {
      const int i = (int)(t % n0) + 1, j = (int)((t / n0) % n1) + 1, k = (int)(t / ((long)n0 * n1)) + 1;
      const long p = gix(g, i, j, k);
      v = fabs((double)a[c][p] - (double)b[c][p]);
    }
    const double s = block_reduce(v, [](double x, double y) { return x + y; });
    if (threadIdx.x == 0) partial[(long)c * nblocks + blockIdx.x] = s;
  }
}

// One thread: component sums in block order, means, accumulation into the
// step's delta (reset at the first stage, divided by dt at the last).
__global__ void k_delta_final(Geo g, const double *__restrict__ partial, int nblocks, int stage, double dt,
                              StepMon *mon) {
  if (threadIdx.x != 0) return;
  double d = stage == 0 ? 0.0 : mon->delta;
  for (int c = 0; c < 3; c++) {
    const long n = (long)g.n[c][0] * g.n[c][1] * g.n[c][2];
    if (n <= 0) continue;
    double s = 0;
    for (int b = 0; b < nblocks; b++) s += partial[(long)c * nblocks + b];
    d = d + s / (double)n;
  }
  if (stage == 2) {
    d = d / dt;
    mon->delta_last = d;
  }
  mon->delta = d;
}

int delta_blocks(const Geo &g) {
  long n = 0;
  for (int c = 0; c < 3; c++) n = std::max(n, (long)g.n[c][0] * g.n[c][1] * g.n[c][2]);
  return (int)((n + 255) / 256);
}

template <typename T>
void launch_delta(const Geo &g, const T *U, const T *U2, const T *V, const T *V2, const T *W, const T *W2,
                  double *partial, int nblocks, int stage, double dt, StepMon *mon, hipStream_t st) {
  hipLaunchKernelGGL(k_delta_partial<T>, dim3(nblocks), dim3(256), 0, st, g, U, U2, V, V2, W, W2, partial, nblocks);
  hipLaunchKernelGGL(k_delta_final, dim3(1), dim3(64), 0, st, g, partial, nblocks, stage, dt, mon);
}

template void launch_delta<double>(const Geo &, const double *, const double *, const double *, const double *,
                                   const double *, const double *, double *, int, int, double, StepMon *, hipStream_t);
template void launch_delta<float>(const Geo &, const float *, const float *, const float *, const float *,
                                  const float *, const float *, double *, int, int, double, StepMon *, hipStream_t);

}  // namespace elmm
