// Device code shared by the direct kernels (direct.hip) and the ensemble
// kernels (factored.hip): the per-mode FP64 preparation, the revolution sine
// and the VALU mode loop of the scalar summator.
#pragma once

#include <hip/hip_runtime.h>

#include "gsc_internal.h"

namespace gsc {

constexpr int kChunk = 32;  // modes per FP32 partial sum of the VALU mode loop

constexpr double kInv2Pi = 0.15915494309189533576888376337251436;  // 1 / (2 pi)

__device__ __forceinline__ float sin_rev(double t) {
  // t - floor(t) is exact in binary64; the cast rounds r in [0, 1) to FP32
  // (error <= 2^-25 revolutions); v_sin_f32 computes sin(2 pi r).
  return __builtin_amdgcn_sinf(static_cast<float>(__builtin_amdgcn_fract(t)));
}

template <typename OutT>
__device__ __forceinline__ void store(OutT* p, double v) { *p = static_cast<OutT>(v); }

// Prepares mode m (m < 16 ntiles) of one mode set: its ModeRec (m < M) and
// its entries of the matrix-core tile tables.
__device__ __forceinline__ void prep_mode(int kind, int dim, int64_t M, int64_t m,
                                          const double* __restrict__ cov,
                                          const double* __restrict__ z1,
                                          const double* __restrict__ z2,
                                          const double* __restrict__ sf,
                                          ModeRec* __restrict__ table, double* __restrict__ a_tab,
                                          float* __restrict__ amp_tab) {
  ModeRec r;
  r.kap[0] = r.kap[1] = r.kap[2] = 0.0;
  r.th = 0.0;
  r.amp[0] = r.amp[1] = r.amp[2] = r.amp[3] = 0.0f;
  if (m < M) {  // modes past M pad the last MFMA tile with zeros
    double k[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < dim; ++i) k[i] = cov[i * M + m];
    const double a = z1[m], b = z2[m];
    const double A = hypot(a, b);
    for (int i = 0; i < 3; ++i) r.kap[i] = k[i] * kInv2Pi;
    r.th = atan2(a, b) * kInv2Pi;
    if (kind == KIND_SCALAR) {
      r.amp[0] = static_cast<float>(A);
    } else if (kind == KIND_FOURIER) {
      r.amp[0] = static_cast<float>(sf[m] * A);
    } else {
      // projector p_c = delta_c0 - k_c k_0 / |k|^2 (field.rs:138-151); k = 0
      // gives 0/0 = NaN exactly as the reference does.
      double kk = 0.0;
      for (int i = 0; i < dim; ++i) kk += k[i] * k[i];
      for (int c = 0; c < dim; ++c) {
        const double pc = (c == 0 ? 1.0 : 0.0) - (k[c] * k[0]) / kk;
        r.amp[c] = static_cast<float>(pc * A);
      }
    }
    table[m] = r;
  }
  // MFMA tile tables (tile t = m / 16, row i = m % 16):
  //   a_tab[t][kk][i]: A operand of v_mfma_f64_16x16x4, lane l = 16 kk + i holds
  //                    parameter kk of mode i (kap_0..kap_{dim-1}, then th, then 0)
  //   amp_tab[t][c][4 g + it]: amplitude of component c of mode 4 it + g, the
  //                    mode in D row 4 it + g seen by lanes 16 g .. 16 g + 15
  const int t = static_cast<int>(m >> 4), i = static_cast<int>(m & 15);
  for (int kk = 0; kk < 4; ++kk) {
    double v = 0.0;
    if (kk < dim) v = r.kap[kk];
    else if (kk == dim) v = r.th;
    a_tab[static_cast<int64_t>(t) * 64 + kk * 16 + i] = v;
  }
  const int it = i >> 2, g = i & 3;
  for (int c = 0; c < 3; ++c) amp_tab[(static_cast<int64_t>(t) * 3 + c) * 16 + 4 * g + it] = r.amp[c];
}


// Mode loop shared by the point-list and grid forms: tot[p] = sum over all
// modes of amp0 sin(2 pi (kap . x_p + th)).
// Global-memory view of a pointer (for pointers read from device memory,
// which the compiler would otherwise access with flat instructions).
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* global_ptr(T* p) {
  return (__attribute__((address_space(1))) T*)(p);  // address-space conversion
}

// Mode record m read field by field (works for either pointer view).
template <typename MP>
__device__ __forceinline__ ModeRec load_mode(MP modes, int m) {
  ModeRec r;
#pragma unroll
  for (int k = 0; k < 3; ++k) r.kap[k] = modes[m].kap[k];
  r.th = modes[m].th;
#pragma unroll
  for (int k = 0; k < 4; ++k) r.amp[k] = modes[m].amp[k];
  return r;
}

// MP: const ModeRec* or its global_ptr view.
template <int D, int P, typename MP>
__device__ __forceinline__ void sum_modes_scalar(MP __restrict__ modes, int M,
                                                 const double (&x)[P][D], double (&tot)[P]) {
#pragma unroll
  for (int p = 0; p < P; ++p) tot[p] = 0.0;
  // Software-pipelined mode loads: the record for m + 1 is requested while
  // mode m is being summed, so the scalar-load latency hides behind VALU work.
  ModeRec nxt;
  if (M > 0) nxt = load_mode(modes, 0);
  for (int m0 = 0; m0 < M; m0 += kChunk) {
    const int m1 = min(m0 + kChunk, M);
    float acc[P];
#pragma unroll
    for (int p = 0; p < P; ++p) acc[p] = 0.0f;
    for (int m = m0; m < m1; ++m) {
      const ModeRec md = nxt;
      nxt = load_mode(modes, min(m + 1, M - 1));
#pragma unroll
      for (int p = 0; p < P; ++p) {
        double t = __builtin_fma(md.kap[0], x[p][0], md.th);
        if constexpr (D > 1) t = __builtin_fma(md.kap[1], x[p][1], t);
        if constexpr (D > 2) t = __builtin_fma(md.kap[2], x[p][2], t);
        acc[p] = __builtin_fmaf(md.amp[0], sin_rev(t), acc[p]);
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) tot[p] += static_cast<double>(acc[p]);
  }
}

// index i with start[i] <= t < start[i + 1] (start has n + 1 entries)
__device__ __forceinline__ int find_item(const int64_t* __restrict__ start, int n, int64_t t) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace gsc
