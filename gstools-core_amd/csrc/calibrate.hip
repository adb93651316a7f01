// Issue-rate calibration of the direct kernels' roofline: how many
// v_sin_f32 per second the device sustains (the one transcendental each
// point-mode costs), alone and paired with the FP32 FMA that accumulates it.
//
// Every lane runs kChains independent dependency chains, so with 8 waves per
// SIMD the rate is set by instruction issue, not by latency; the result is
// written only if it equals an impossible value, which keeps the work live.
// A second pair of kernels runs the cell-local kernels' whole instruction mix
// (matrix products, sines and FMAs), the ceiling of those kernels without
// their memory traffic.
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

typedef _Float16 mix_h8 __attribute__((ext_vector_type(8)));
typedef __bf16 mix_bf8 __attribute__((ext_vector_type(8)));
typedef float mix_f4 __attribute__((ext_vector_type(4)));
typedef float mix_f16 __attribute__((ext_vector_type(16)));

constexpr int kMixIters = 2048;
constexpr int kMixBlocks = 256 * 16;  // sixteen 256-thread blocks per CU, four waves per SIMD at once

// The instruction mix of the cell-local kernels per 1024 point-modes of a
// wave, without their memory traffic: one v_mfma_f32_32x32x16_f16, 16 sines
// and 16 FMAs (scalar, k_direct_local32), or four v_mfma_f32_16x16x32_bf16,
// 16 sines and 48 FMAs (vector, k_direct_local_vec with three components).
// As in the kernels, the products of the next tile go out before the sines
// of the current one, four waves per SIMD, and the FMA operands are VGPRs.
template <bool VEC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_local_mix(float seed, float* __restrict__ sink) {
  constexpr int NF = VEC ? 3 : 1;
  const int lane = threadIdx.x & 63;
  float am[NF], acc[NF][16];
#pragma unroll
  for (int c = 0; c < NF; ++c) {
    am[c] = seed * static_cast<float>(c + 1) + 1e-3f * static_cast<float>(lane);
#pragma unroll
    for (int v = 0; v < 16; ++v) acc[c][v] = 0.0f;
  }
  if constexpr (!VEC) {
    mix_h8 a, b;
    mix_f16 zero;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      a[i] = static_cast<_Float16>(1e-3f * static_cast<float>(lane + i));
      b[i] = static_cast<_Float16>(2e-3f * static_cast<float>(lane - i));
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) zero[v] = 0.0f;
    mix_f16 t = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, zero, 0, 0, 0);
    for (int it = 0; it < kMixIters; ++it) {
      b[0] = static_cast<_Float16>(static_cast<float>(it));  // new products every tile
      const mix_f16 tn = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, zero, 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int v = 0; v < 16; ++v)
        acc[0][v] = __builtin_fmaf(am[0], __builtin_amdgcn_sinf(t[v]), acc[0][v]);
      t = tn;
    }
  } else {
    mix_bf8 a[4], b;  // four distinct products per tile
    const mix_f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
#pragma unroll
      for (int k = 0; k < 4; ++k)
        a[k][i] = static_cast<__bf16>(1e-3f * static_cast<float>(k + 1) *
                                      static_cast<float>(lane + i));
      b[i] = static_cast<__bf16>(2e-3f * static_cast<float>(lane - i));
    }
    mix_f4 t[4];
#pragma unroll
    for (int k = 0; k < 4; ++k)
      t[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[k], b, zero, 0, 0, 0);
    for (int it = 0; it < kMixIters; ++it) {
      b[0] = static_cast<__bf16>(static_cast<float>(it));
      mix_f4 tn[4];
#pragma unroll
      for (int k = 0; k < 4; ++k)
        tn[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[k], b, zero, 0, 0, 0);
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int k = 0; k < 4; ++k)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float sv = __builtin_amdgcn_sinf(t[k][v]);
#pragma unroll
          for (int c = 0; c < 3; ++c)
            acc[c][4 * k + v] = __builtin_fmaf(am[c], sv, acc[c][4 * k + v]);
        }
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = tn[k];
    }
  }
  float r = 0.0f;
#pragma unroll
  for (int c = 0; c < NF; ++c)
#pragma unroll
    for (int v = 0; v < 16; ++v) r += acc[c][v];
  if (r == 12345.678f) sink[threadIdx.x] = r;
}

}  // namespace
}  // namespace gsc

extern "C" int gsc_calibrate_local_mix(int repeats, double* scalar_pm_per_s,
                                       double* vector_pm_per_s) {
  if (repeats < 1 || repeats > 1000 || !scalar_pm_per_s || !vector_pm_per_s)
    return GSC_ERR_INVALID;
  float* sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  hipError_t e = hipMalloc(&sink, 256 * sizeof(float));
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  for (int i = 0; i < 3 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
  float t_s = 0.0f, t_v = 0.0f;
  if (e == hipSuccess) {
    gsc::k_local_mix<false><<<gsc::kMixBlocks, 256, 0, st>>>(0.1f, sink);  // warm-up
    gsc::k_local_mix<true><<<gsc::kMixBlocks, 256, 0, st>>>(0.1f, sink);
    (void)hipEventRecord(ev[0], st);
    for (int r = 0; r < repeats; ++r)
      gsc::k_local_mix<false><<<gsc::kMixBlocks, 256, 0, st>>>(0.1f, sink);
    (void)hipEventRecord(ev[1], st);
    for (int r = 0; r < repeats; ++r)
      gsc::k_local_mix<true><<<gsc::kMixBlocks, 256, 0, st>>>(0.1f, sink);
    (void)hipEventRecord(ev[2], st);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipEventSynchronize(ev[2]);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_s, ev[0], ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&t_v, ev[1], ev[2]);
  }
  // 16 point-modes per lane and tile
  const double n = static_cast<double>(gsc::kMixBlocks) * 256.0 * gsc::kMixIters * 16.0 * repeats;
  if (e == hipSuccess && t_s > 0.0f && t_v > 0.0f) {
    *scalar_pm_per_s = n / (t_s * 1e-3);
    *vector_pm_per_s = n / (t_v * 1e-3);
  }
  for (int i = 0; i < 3; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (st) (void)hipStreamDestroy(st);
  if (sink) (void)hipFree(sink);
  return e == hipSuccess ? GSC_OK : GSC_ERR_HIP;
}

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
