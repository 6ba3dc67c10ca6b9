// k_probe.hip - plain device copy of a known size: the reference point of
// the HBM traffic counters (FETCH_SIZE/WRITE_SIZE per dispatch) and of the
// achievable copy bandwidth reported beside the roofline; and the launch
// attribute helper shared by the kernels with more than 64 KiB of LDS.
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>

#include "device_common.hpp"
#include "elmm_gpu.h"
#include "errors.hpp"

namespace {

// Grid-stride copy with four independent 16-byte loads in flight per thread
// before their stores (a wavefront moves 4 KiB per round).
__global__ void __launch_bounds__(256) k_copy_probe(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                    long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t + 3 * stride < n; t += 4 * stride) {
    const float4 a = src[t], b = src[t + stride], c = src[t + 2 * stride], d = src[t + 3 * stride];
    dst[t] = a;
    dst[t + stride] = b;
    dst[t + 2 * stride] = c;
    dst[t + 3 * stride] = d;
  }
  for (; t < n; t += stride) dst[t] = src[t];
}

}  // namespace

namespace elmm {
bool lds_limit_raise(const void *fn, size_t bytes) {
  HostOpLock host;
  static std::mutex mu;
  static std::map<const void *, size_t> done;
  std::lock_guard<std::mutex> lock(mu);
  auto it = done.find(fn);
  if (it != done.end() && it->second >= bytes) return true;
  const bool ok = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) == hipSuccess;
  if (ok) done[fn] = bytes;
  return ok;
}
}  // namespace elmm

extern "C" int elmm_copy_probe(int64_t nbytes, int32_t reps, double *gbps) {
  if (nbytes < 32 || reps < 1 || !gbps) return elmm::fail(ELMM_E_ARG, "elmm_copy_probe: bad arguments");
  elmm::HostOpLock host;
  const long n = (long)(nbytes / 32) * 2;  // float4 elements (16 bytes), a multiple of 32 bytes in total
  void *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&] {
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  };
  hipError_t err;
  if ((err = hipMalloc(&a, n * 16)) != hipSuccess || (err = hipMalloc(&b, n * 16)) != hipSuccess ||
      (err = hipMemset(a, 0, n * 16)) != hipSuccess || (err = hipEventCreate(&e0)) != hipSuccess ||
      (err = hipEventCreate(&e1)) != hipSuccess) {
    cleanup();
    return elmm::fail(ELMM_E_HIP, std::string("elmm_copy_probe: ") + hipGetErrorString(err));
  }
  const int blocks = 256 * 8;  // 8 blocks of 256 threads per CU
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n);
  (void)hipEventRecord(e0, 0);
  for (int r = 0; r < reps; r++)
    hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, 0, (const float4 *)a, (float4 *)b, n);
  (void)hipEventRecord(e1, 0);
  err = hipEventSynchronize(e1);
  float ms = 0;
  if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
  cleanup();
  if (err != hipSuccess) return elmm::fail(ELMM_E_HIP, std::string("elmm_copy_probe: ") + hipGetErrorString(err));
  *gbps = 2.0 * n * 16 * reps / (ms * 1e-3) / 1e9;
  return 0;
}
