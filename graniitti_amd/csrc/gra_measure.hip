// Measurement probe: the FP32 FMA rate this device sustains without packed FP32
// instructions (the library is built with packed FP32 disabled), the ceiling of
// the FP32 screening node's instruction stream.
#include <hip/hip_runtime.h>

#include "gra_hip.h"
#include "gra_internal.h"

namespace {

// Eight independent FMA chains per thread, four FMAs each per iteration, so
// every wave always has an independent FMA to issue
__global__ __launch_bounds__(256) void k_fma_rate(float* __restrict__ out, int iters, float a,
                                                  float b) {
  float acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = threadIdx.x * 1e-3f + k;
  for (int i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int k = 0; k < 8; ++k) acc[k] = fmaf(acc[k], a, b);
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < 8; ++k) s += acc[k];
  out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

extern "C" int gra_hip_fp32_fma_rate(double* tflops) {
  if (!tflops) return gra_hip::set_global_error(GRA_E_ARG, "fp32_fma_rate: tflops is NULL");
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    return gra_hip::set_global_error(GRA_E_HIP, "fp32_fma_rate: no device");
  // eight waves per SIMD (four-wave blocks, eight blocks per CU)
  const int blocks = 8 * cus, iters = 20000;
  float* d = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = GRA_OK;
  if (hipMalloc(&d, sizeof(float) * 256 * (size_t)blocks) != hipSuccess ||
      hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    rc = gra_hip::set_global_error(GRA_E_HIP, "fp32_fma_rate: set-up failed");
  } else {
    hipLaunchKernelGGL(k_fma_rate, dim3(blocks), dim3(256), 0, st, d, 100, 0.999f, 0.001f);
    (void)hipEventRecord(e0, st);
    hipLaunchKernelGGL(k_fma_rate, dim3(blocks), dim3(256), 0, st, d, iters, 0.999f, 0.001f);
    (void)hipEventRecord(e1, st);
    float ms = 0.f;
    if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&ms, e0, e1) != hipSuccess || !(ms > 0.f)) {
      rc = gra_hip::set_global_error(GRA_E_HIP, "fp32_fma_rate: launch failed");
    } else {
      const double fma = (double)blocks * 256 * iters * 32;
      *tflops = 2.0 * fma / (ms * 1e-3) / 1e12;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (d) (void)hipFree(d);
  return rc;
}
