// Issue-rate calibration of the direct kernels' roofline: how many
// v_sin_f32 per second the device sustains (the one transcendental each
// point-mode costs), alone and paired with the FP32 FMA that accumulates it.
//
// Every lane runs kChains independent dependency chains, so with 8 waves per
// SIMD the rate is set by instruction issue, not by latency; the result is
// written only if it equals an impossible value, which keeps the work live.
#include <hip/hip_runtime.h>

#include "gsc_internal.h"
#include "gstools_core_amd.h"

namespace gsc {
namespace {

constexpr int kChains = 8;
constexpr int kIters = 4096;
constexpr int kBlocks = 256 * 8;  // eight 256-thread blocks per CU of an MI355X

template <bool FMA>
__global__ void __launch_bounds__(256) k_sine_rate(float seed, float* __restrict__ sink) {
  float x[kChains], acc[kChains];
#pragma unroll
  for (int c = 0; c < kChains; ++c) {
    x[c] = seed + 0.001f * static_cast<float>(threadIdx.x + 64 * c);
    acc[c] = 0.0f;
  }
  for (int it = 0; it < kIters; ++it) {
#pragma unroll
    for (int c = 0; c < kChains; ++c) {
      const float s = __builtin_amdgcn_sinf(x[c]);
      if constexpr (FMA) {
        acc[c] = __builtin_fmaf(s, 0.75f, acc[c]);
        x[c] = s;
      } else {
        x[c] = s;
      }
    }
  }
  float t = 0.0f;
#pragma unroll
  for (int c = 0; c < kChains; ++c) t += x[c] + acc[c];
  if (t == 12345.678f) sink[threadIdx.x] = t;
}

}  // namespace
}  // namespace gsc

extern "C" int gsc_calibrate_sine(int repeats, double* sin_per_s, double* sin_fma_per_s) {
  if (repeats < 1 || repeats > 1000 || !sin_per_s || !sin_fma_per_s) return GSC_ERR_INVALID;
  float* sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  hipError_t e = hipMalloc(&sink, 256 * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
  float t_sin = 0.0f, t_fma = 0.0f;
  if (e == hipSuccess) {
    gsc::k_sine_rate<false><<<gsc::kBlocks, 256, 0, st>>>(0.1f, sink);  // warm-up
    gsc::k_sine_rate<true><<<gsc::kBlocks, 256, 0, st>>>(0.1f, sink);
    (void)hipEventRecord(ev[0], st);
    for (int r = 0; r < repeats; ++r)
      gsc::k_sine_rate<false><<<gsc::kBlocks, 256, 0, st>>>(0.1f, sink);
    (void)hipEventRecord(ev[1], st);
    for (int r = 0; r < repeats; ++r)
      gsc::k_sine_rate<true><<<gsc::kBlocks, 256, 0, st>>>(0.1f, sink);
    (void)hipEventRecord(ev[2], st);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_sin, ev[0], ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_fma, ev[1], ev[2]);
  }
  const double n = static_cast<double>(gsc::kBlocks) * 256.0 * gsc::kIters * gsc::kChains *
                   repeats;
  if (e == hipSuccess && t_sin > 0.0f && t_fma > 0.0f) {
    *sin_per_s = n / (t_sin * 1e-3);
    *sin_fma_per_s = n / (t_fma * 1e-3);
  }
  for (int i = 0; i < 3; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (st) (void)hipStreamDestroy(st);
  if (sink) (void)hipFree(sink);
  return e == hipSuccess ? GSC_OK : GSC_ERR_HIP;
}
