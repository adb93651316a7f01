// Direct (per-point) randomization-method summators for gfx950.
//
// Reference: src/field.rs:53-62 (summator), 130-174 (summator_incompr) and
// 236-246 (summator_fourier) evaluate, for every point x_j and mode m,
//   z12 = z1_m cos(k_m . x_j) + z2_m sin(k_m . x_j)
// in FP64 with libm cos/sin.  Here each point-mode costs:
//   t  = kap . x + th          FP64 FMA chain (phase in revolutions)
//   r  = fract(t)              exact in FP64: the range reduction that FP32
//                              cannot do once |k . x| reaches ~100 rad
//   s  = v_sin_f32((float) r)  hardware sine of 2 pi r (input in revolutions)
//   acc = fma(amp, s, acc)     FP32, one accumulator per 32-mode chunk
// and every chunk is added into an FP64 total, so FP32 rounding of the running
// sum never grows with M.  The mode table (ModeRec) is wave-uniform: the
// compiler reads it with scalar loads into SGPRs, which VALU instructions take
// as operands for free.  Positions are (dim, N) row-major: coalesced FP64
// loads, once per point, kept in VGPRs for the whole mode loop.
#include "gsc_internal.h"
#include "gsc_factored.h"

#include <hip/hip_runtime.h>

#include <cmath>

namespace gsc {
namespace {

constexpr int kBlock = 256;     // 4 waves of 64
constexpr int kChunk = 32;      // modes per FP32 partial sum
constexpr double kInv2Pi = 0.15915494309189533576888376337251436;  // 1 / (2 pi)

__device__ __forceinline__ float sin_rev(double t) {
  // t - floor(t) is exact in binary64; the cast rounds r in [0, 1) to FP32
  // (error <= 2^-25 revolutions); v_sin_f32 computes sin(2 pi r).
  return __builtin_amdgcn_sinf(static_cast<float>(__builtin_amdgcn_fract(t)));
}

template <typename OutT>
__device__ __forceinline__ void store(OutT* p, double v) { *p = static_cast<OutT>(v); }

// ---------------------------------------------------------------------------
// Mode preparation (FP64).  One thread per mode.
// ---------------------------------------------------------------------------
__global__ void k_prep_small(int kind, int dim, int64_t M, const double* __restrict__ cov,
                             const double* __restrict__ z1, const double* __restrict__ z2,
                             const double* __restrict__ sf, ModeRec* __restrict__ table) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double k[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < dim; ++i) k[i] = cov[i * M + m];
  const double a = z1[m], b = z2[m];
  const double A = hypot(a, b);
  ModeRec r;
  for (int i = 0; i < 3; ++i) r.kap[i] = k[i] * kInv2Pi;
  r.th = atan2(a, b) * kInv2Pi;
  r.amp[0] = r.amp[1] = r.amp[2] = r.amp[3] = 0.0f;
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
      const double p = (c == 0 ? 1.0 : 0.0) - (k[c] * k[0]) / kk;
      r.amp[c] = static_cast<float>(p * A);
    }
  }
  table[m] = r;
}

// dim > 3 (scalar and Fourier summators only): [th, amp, kap_0..kap_{d-1}].
__global__ void k_prep_generic(int kind, int dim, int64_t M, const double* __restrict__ cov,
                               const double* __restrict__ z1, const double* __restrict__ z2,
                               const double* __restrict__ sf, double* __restrict__ table) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const double a = z1[m], b = z2[m];
  double A = hypot(a, b);
  if (kind == KIND_FOURIER) A *= sf[m];
  double* r = table + m * (dim + 2);
  r[0] = atan2(a, b) * kInv2Pi;
  r[1] = static_cast<double>(static_cast<float>(A));
  for (int i = 0; i < dim; ++i) r[2 + i] = cov[i * M + m] * kInv2Pi;
}

// ---------------------------------------------------------------------------
// K1: scalar summator.  P points per thread (strided by the block size so
// every load and store is coalesced); grid = ceil(N / (256 P)) blocks.
// ---------------------------------------------------------------------------
// Mode loop shared by the point-list and grid forms: tot[p] = sum over all
// modes of amp0 sin(2 pi (kap . x_p + th)).
template <int D, int P>
__device__ __forceinline__ void sum_modes_scalar(const ModeRec* __restrict__ modes, int M,
                                                 const double (&x)[P][D], double (&tot)[P]) {
#pragma unroll
  for (int p = 0; p < P; ++p) tot[p] = 0.0;
  // Software-pipelined mode loads: the record for m + 1 is requested while
  // mode m is being summed, so the scalar-load latency hides behind VALU work.
  ModeRec nxt;
  if (M > 0) nxt = modes[0];
  for (int m0 = 0; m0 < M; m0 += kChunk) {
    const int m1 = min(m0 + kChunk, M);
    float acc[P];
#pragma unroll
    for (int p = 0; p < P; ++p) acc[p] = 0.0f;
    for (int m = m0; m < m1; ++m) {
      const ModeRec md = nxt;
      nxt = modes[min(m + 1, M - 1)];
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

template <int D, int P, typename OutT>
__global__ void __launch_bounds__(kBlock)
k_direct_scalar(const ModeRec* __restrict__ modes, int M, const double* __restrict__ pos,
                int64_t pos_ld, int64_t N, OutT* __restrict__ out, double scale) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * (kBlock * P) + threadIdx.x;
  double x[P][D];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    int64_t j = base + p * kBlock;
    j = j < N ? j : N - 1;  // tail lanes recompute the last point, never store it
#pragma unroll
    for (int i = 0; i < D; ++i) x[p][i] = pos[i * pos_ld + j];
  }
  double tot[P];
  sum_modes_scalar<D, P>(modes, M, x, tot);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int64_t j = base + p * kBlock;
    if (j < N) store(out + j, tot[p] * scale);
  }
}

// Grid form: point j of the C-order grid (s0, s1, s2) has coordinates
// (ax0[j / (s1 s2)], ax1[(j / s2) % s1], ax2[j % s2]) (fewer axes for D < 3).
template <int D, int P, typename OutT>
__global__ void __launch_bounds__(kBlock)
k_direct_grid(const ModeRec* __restrict__ modes, int M, const double* __restrict__ ax0,
              const double* __restrict__ ax1, const double* __restrict__ ax2, int64_t s1,
              int64_t s2, int64_t N, OutT* __restrict__ out, double scale) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * (kBlock * P) + threadIdx.x;
  double x[P][D];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    int64_t j = base + p * kBlock;
    j = j < N ? j : N - 1;
    if constexpr (D == 1) {
      x[p][0] = ax0[j];
    } else if constexpr (D == 2) {
      x[p][0] = ax0[j / s1];
      x[p][1] = ax1[j % s1];
    } else {
      const int64_t i0 = j / (s1 * s2), r = j - i0 * (s1 * s2);
      x[p][0] = ax0[i0];
      x[p][1] = ax1[r / s2];
      x[p][2] = ax2[r % s2];
    }
  }
  double tot[P];
  sum_modes_scalar<D, P>(modes, M, x, tot);
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int64_t j = base + p * kBlock;
    if (j < N) store(out + j, tot[p] * scale);
  }
}

// Generic dimension (4..16), one point per thread.
template <typename OutT>
__global__ void __launch_bounds__(kBlock)
k_direct_generic(const double* __restrict__ modes, int dim, int M, const double* __restrict__ pos,
                 int64_t pos_ld, int64_t N, OutT* __restrict__ out, double scale) {
  const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const int64_t j = j0 < N ? j0 : N - 1;
  double x[kMaxDim];
#pragma unroll
  for (int i = 0; i < kMaxDim; ++i) x[i] = i < dim ? pos[i * pos_ld + j] : 0.0;
  double tot = 0.0;
  const int rec = dim + 2;
  for (int m0 = 0; m0 < M; m0 += kChunk) {
    const int m1 = min(m0 + kChunk, M);
    float acc = 0.0f;
    for (int m = m0; m < m1; ++m) {
      const double* md = modes + static_cast<int64_t>(m) * rec;
      double t = md[0];
#pragma unroll
      for (int i = 0; i < kMaxDim; ++i)
        if (i < dim) t = __builtin_fma(md[2 + i], x[i], t);
      acc = __builtin_fmaf(static_cast<float>(md[1]), sin_rev(t), acc);
    }
    tot += static_cast<double>(acc);
  }
  if (j0 < N) store(out + j0, tot * scale);
}

// ---------------------------------------------------------------------------
// K2: incompressible vector summator, D in {2, 3}; D outputs per point.
// ---------------------------------------------------------------------------
template <int D, int P, typename OutT>
__global__ void __launch_bounds__(kBlock)
k_direct_incompr(const ModeRec* __restrict__ modes, int M, const double* __restrict__ pos,
                 int64_t pos_ld, int64_t N, OutT* __restrict__ out, int64_t out_ld,
                 double scale) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * (kBlock * P) + threadIdx.x;
  double x[P][D];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    int64_t j = base + p * kBlock;
    j = j < N ? j : N - 1;
#pragma unroll
    for (int i = 0; i < D; ++i) x[p][i] = pos[i * pos_ld + j];
  }
  double tot[P][D];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int c = 0; c < D; ++c) tot[p][c] = 0.0;

  ModeRec nxt;
  if (M > 0) nxt = modes[0];
  for (int m0 = 0; m0 < M; m0 += kChunk) {
    const int m1 = min(m0 + kChunk, M);
    float acc[P][D];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int c = 0; c < D; ++c) acc[p][c] = 0.0f;
    for (int m = m0; m < m1; ++m) {
      const ModeRec md = nxt;
      nxt = modes[min(m + 1, M - 1)];
#pragma unroll
      for (int p = 0; p < P; ++p) {
        double t = __builtin_fma(md.kap[0], x[p][0], md.th);
        t = __builtin_fma(md.kap[1], x[p][1], t);
        if constexpr (D > 2) t = __builtin_fma(md.kap[2], x[p][2], t);
        const float s = sin_rev(t);
#pragma unroll
        for (int c = 0; c < D; ++c) acc[p][c] = __builtin_fmaf(md.amp[c], s, acc[p][c]);
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int c = 0; c < D; ++c) tot[p][c] += static_cast<double>(acc[p][c]);
  }
#pragma unroll
  for (int p = 0; p < P; ++p) {
    const int64_t j = base + p * kBlock;
    if (j < N) {
#pragma unroll
      for (int c = 0; c < D; ++c) store(out + c * out_ld + j, tot[p][c] * scale);
    }
  }
}

inline unsigned grid_for(int64_t n, int per_block) {
  return static_cast<unsigned>((n + per_block - 1) / per_block);
}

template <int D, typename OutT>
hipError_t scalar_t(int64_t M, int64_t N, const void* table, const double* pos, int64_t pos_ld,
                    void* out, double scale, hipStream_t s) {
  constexpr int P = D == 1 ? 8 : (D == 2 ? 8 : 8);
  k_direct_scalar<D, P, OutT><<<grid_for(N, kBlock * P), kBlock, 0, s>>>(
      static_cast<const ModeRec*>(table), static_cast<int>(M), pos, pos_ld, N,
      static_cast<OutT*>(out), scale);
  return hipGetLastError();
}

template <int D, typename OutT>
hipError_t incompr_t(int64_t M, int64_t N, const void* table, const double* pos, int64_t pos_ld,
                     void* out, int64_t out_ld, double scale, hipStream_t s) {
  constexpr int P = 4;
  k_direct_incompr<D, P, OutT><<<grid_for(N, kBlock * P), kBlock, 0, s>>>(
      static_cast<const ModeRec*>(table), static_cast<int>(M), pos, pos_ld, N,
      static_cast<OutT*>(out), out_ld, scale);
  return hipGetLastError();
}

}  // namespace

size_t mode_table_bytes(int dim, int64_t M) {
  const size_t per = dim <= 3 ? sizeof(ModeRec) : static_cast<size_t>(dim + 2) * sizeof(double);
  const size_t b = per * static_cast<size_t>(M > 0 ? M : 1);
  return (b + 255) & ~static_cast<size_t>(255);
}

hipError_t launch_prep(int kind, int dim, int64_t M, const double* cov, const double* z1,
                       const double* z2, const double* sf, void* table, hipStream_t stream) {
  if (M <= 0) return hipSuccess;
  const unsigned g = grid_for(M, 256);
  if (dim <= 3)
    k_prep_small<<<g, 256, 0, stream>>>(kind, dim, M, cov, z1, z2, sf,
                                        static_cast<ModeRec*>(table));
  else
    k_prep_generic<<<g, 256, 0, stream>>>(kind, dim, M, cov, z1, z2, sf,
                                          static_cast<double*>(table));
  return hipGetLastError();
}

hipError_t launch_direct_scalar(int dim, int64_t M, int64_t N, const void* table,
                                const double* pos, int64_t pos_ld, void* out, int out_dtype,
                                double scale, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  const bool f32 = out_dtype == 1;
  switch (dim) {
    case 1: return f32 ? scalar_t<1, float>(M, N, table, pos, pos_ld, out, scale, s)
                       : scalar_t<1, double>(M, N, table, pos, pos_ld, out, scale, s);
    case 2: return f32 ? scalar_t<2, float>(M, N, table, pos, pos_ld, out, scale, s)
                       : scalar_t<2, double>(M, N, table, pos, pos_ld, out, scale, s);
    case 3: return f32 ? scalar_t<3, float>(M, N, table, pos, pos_ld, out, scale, s)
                       : scalar_t<3, double>(M, N, table, pos, pos_ld, out, scale, s);
    default: {
      const unsigned g = grid_for(N, kBlock);
      if (f32)
        k_direct_generic<float><<<g, kBlock, 0, s>>>(static_cast<const double*>(table), dim,
                                                     static_cast<int>(M), pos, pos_ld, N,
                                                     static_cast<float*>(out), scale);
      else
        k_direct_generic<double><<<g, kBlock, 0, s>>>(static_cast<const double*>(table), dim,
                                                      static_cast<int>(M), pos, pos_ld, N,
                                                      static_cast<double*>(out), scale);
      return hipGetLastError();
    }
  }
}

hipError_t launch_direct_incompr(int dim, int64_t M, int64_t N, const void* table,
                                 const double* pos, int64_t pos_ld, void* out, int64_t out_ld,
                                 int out_dtype, double scale, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  const bool f32 = out_dtype == 1;
  if (dim == 2)
    return f32 ? incompr_t<2, float>(M, N, table, pos, pos_ld, out, out_ld, scale, s)
               : incompr_t<2, double>(M, N, table, pos, pos_ld, out, out_ld, scale, s);
  return f32 ? incompr_t<3, float>(M, N, table, pos, pos_ld, out, out_ld, scale, s)
             : incompr_t<3, double>(M, N, table, pos, pos_ld, out, out_ld, scale, s);
}

hipError_t launch_direct_grid(int dim, int64_t M, const int64_t* shape, const double* const* axes,
                              const void* table, void* out, int out_dtype, double scale,
                              hipStream_t s) {
  int64_t N = 1;
  for (int i = 0; i < dim; ++i) N *= shape[i];
  if (N <= 0) return hipSuccess;
  constexpr int P = 8;
  const unsigned g = grid_for(N, kBlock * P);
  const ModeRec* md = static_cast<const ModeRec*>(table);
  const int Mi = static_cast<int>(M);
  const int64_t s1 = dim > 1 ? shape[1] : 1, s2 = dim > 2 ? shape[2] : 1;
  const double* a0 = axes[0];
  const double* a1 = dim > 1 ? axes[1] : nullptr;
  const double* a2 = dim > 2 ? axes[2] : nullptr;
#define GSC_GRID(D, T) \
  k_direct_grid<D, P, T><<<g, kBlock, 0, s>>>(md, Mi, a0, a1, a2, s1, s2, N, static_cast<T*>(out), scale)
  const bool f32 = out_dtype == 1;
  if (dim == 1) { if (f32) GSC_GRID(1, float); else GSC_GRID(1, double); }
  else if (dim == 2) { if (f32) GSC_GRID(2, float); else GSC_GRID(2, double); }
  else { if (f32) GSC_GRID(3, float); else GSC_GRID(3, double); }
#undef GSC_GRID
  return hipGetLastError();
}

}  // namespace gsc
