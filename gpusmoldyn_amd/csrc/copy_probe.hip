// A plain device copy of known size: the reference point that turns
// hardware byte counters into bytes moved (include/gpusmoldyn.h:
// smd_copy_probe).
#include <hip/hip_runtime.h>

#include <string>

#include "smd_setup.h"

namespace {

__global__ __launch_bounds__(256) void k_copy_probe(const float4* __restrict__ in, float4* __restrict__ out,
                                                    size_t n) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) out[i] = in[i];
}

}  // namespace

#define PROBE_CHK(x)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (x);                                                                    \
    if (e_ != hipSuccess) {                                                                 \
      smd::set_error(std::string("HIP error '") + hipGetErrorString(e_) + "' in " #x);      \
      rc = SMD_ERR_HIP;                                                                     \
      goto done;                                                                            \
    }                                                                                       \
  } while (0)

extern "C" smd_status smd_copy_probe(int device, unsigned long long bytes, int reps, double* us) {
  if (!us || reps <= 0 || bytes < 16) {
    smd::set_error("invalid copy probe arguments");
    return SMD_ERR_INVALID;
  }
  smd_status rc = SMD_OK;
  const size_t n = (size_t)(bytes / 16);
  float4 *a = nullptr, *b = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.0f;
  const unsigned blocks = 4096;
  PROBE_CHK(hipSetDevice(device));
  PROBE_CHK(hipMalloc((void**)&a, n * 16));
  PROBE_CHK(hipMalloc((void**)&b, n * 16));
  PROBE_CHK(hipMemset(a, 0, n * 16));
  PROBE_CHK(hipEventCreate(&e0));
  PROBE_CHK(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, a, b, n);  // warm-up
  PROBE_CHK(hipGetLastError());
  PROBE_CHK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_probe, dim3(blocks), dim3(256), 0, nullptr, a, b, n);
  PROBE_CHK(hipGetLastError());
  PROBE_CHK(hipEventRecord(e1, nullptr));
  PROBE_CHK(hipEventSynchronize(e1));
  PROBE_CHK(hipEventElapsedTime(&ms, e0, e1));
  *us = 1000.0 * ms / reps;
done:
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}
