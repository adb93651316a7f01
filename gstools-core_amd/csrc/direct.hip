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
#include "gsc_device.h"
#include "gsc_factored.h"
#include "gsc_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>

namespace gsc {
namespace {

constexpr int kBlock = 256;     // 4 waves of 64
// ---------------------------------------------------------------------------
// Mode preparation (FP64).  One thread per mode.
// ---------------------------------------------------------------------------
__global__ void k_prep_small(int kind, int dim, int64_t M, int ntiles,
                             const double* __restrict__ cov, const double* __restrict__ z1,
                             const double* __restrict__ z2, const double* __restrict__ sf,
                             ModeRec* __restrict__ table, double* __restrict__ a_tab,
                             float* __restrict__ amp_tab) {
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= static_cast<int64_t>(ntiles) * 16) return;
  prep_mode(kind, dim, M, m, cov, z1, z2, sf, table, a_tab, amp_tab);
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

// ---------------------------------------------------------------------------
// K1m / K2m: direct summators with the FP64 phase on the matrix cores.
//
// One v_mfma_f64_16x16x4f64 per 16 modes x 16 points computes every phase
//   t(i, j) = sum_kk A[i][kk] B[kk][j] = kap_i . x_j + th_i   (exact FP64 FMA
// chains, K = 4 = three coordinates + the constant 1 for th).  Operand and
// result layouts (verified on gfx950):
//   A: lane l holds A[l % 16][l / 16]   (mode l % 16, parameter l / 16)
//   B: lane l holds B[l / 16][l % 16]   (coordinate l / 16 of point l % 16)
//   D: lane l, item it holds D[4 it + l / 16][l % 16]
// so a lane sees 4 modes (4 it + g, g = l / 16) of one point (j = l % 16).
// The VALU keeps only fract + cvt + v_sin_f32 + FMA per point-mode, and the
// matrix core runs the phase work concurrently.  A wave owns P groups of 16
// points; partial sums of the four lane groups are combined at the end.
// ---------------------------------------------------------------------------
typedef double f64x4_t __attribute__((ext_vector_type(4)));

template <int D, int NC, int P, bool GRID, typename OutT>
__global__ void __launch_bounds__(kBlock)
k_direct_mfma(const double* __restrict__ a_tab, const float* __restrict__ amp_tab, int ntiles,
              const double* __restrict__ pos, int64_t pos_ld, const double* __restrict__ ax1,
              const double* __restrict__ ax2, int64_t s1, int64_t s2, int64_t N,
              OutT* __restrict__ out, int64_t out_ld, double scale) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4, j = lane & 15;
  const int64_t base =
      (static_cast<int64_t>(blockIdx.x) * (kBlock / 64) + (threadIdx.x >> 6)) * (16 * P);
  // B operands: coordinate g of point base + 16 p + j (1 for g == D, 0 above)
  double b[P];
#pragma unroll
  for (int p = 0; p < P; ++p) {
    int64_t pt = base + 16 * p + j;
    pt = pt < N ? pt : N - 1;
    double v = g == D ? 1.0 : 0.0;
    if (g < D) {
      if constexpr (!GRID) {
        v = pos[g * pos_ld + pt];
      } else if constexpr (D == 1) {
        v = pos[pt];
      } else if constexpr (D == 2) {
        v = g == 0 ? pos[pt / s1] : ax1[pt % s1];
      } else {
        const int64_t i0 = pt / (s1 * s2), r = pt - i0 * (s1 * s2);
        v = g == 0 ? pos[i0] : (g == 1 ? ax1[r / s2] : ax2[r % s2]);
      }
    }
    b[p] = v;
  }
  double tot[P][NC];
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int c = 0; c < NC; ++c) tot[p][c] = 0.0;

  const f64x4_t zero = {0.0, 0.0, 0.0, 0.0};
  // software-pipelined tile loads
  double a_nxt = 0.0;
  float4 am_nxt[NC];
  auto load_tile = [&](int t) {
    a_nxt = a_tab[static_cast<int64_t>(t) * 64 + lane];
#pragma unroll
    for (int c = 0; c < NC; ++c)
      am_nxt[c] = *reinterpret_cast<const float4*>(amp_tab + (static_cast<int64_t>(t) * 3 + c) * 16 + 4 * g);
  };
  if (ntiles > 0) load_tile(0);
  constexpr int kTileChunk = 8;  // 8 tiles = 32 modes per lane per FP32 partial sum
  for (int t0 = 0; t0 < ntiles; t0 += kTileChunk) {
    const int t1 = min(t0 + kTileChunk, ntiles);
    float acc[P][NC];
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[p][c] = 0.0f;
    for (int t = t0; t < t1; ++t) {
      const double a = a_nxt;
      float4 am[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) am[c] = am_nxt[c];
      load_tile(min(t + 1, ntiles - 1));
      f64x4_t ph[P];
#pragma unroll
      for (int p = 0; p < P; ++p) ph[p] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b[p], zero, 0, 0, 0);
#pragma unroll
      for (int p = 0; p < P; ++p) {
        const float s0 = sin_rev(ph[p][0]), s1v = sin_rev(ph[p][1]);
        const float s2v = sin_rev(ph[p][2]), s3 = sin_rev(ph[p][3]);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          float v = acc[p][c];
          v = __builtin_fmaf(am[c].x, s0, v);
          v = __builtin_fmaf(am[c].y, s1v, v);
          v = __builtin_fmaf(am[c].z, s2v, v);
          v = __builtin_fmaf(am[c].w, s3, v);
          acc[p][c] = v;
        }
      }
    }
#pragma unroll
    for (int p = 0; p < P; ++p)
#pragma unroll
      for (int c = 0; c < NC; ++c) tot[p][c] += static_cast<double>(acc[p][c]);
  }
  // lanes j, j + 16, j + 32, j + 48 hold partial sums of the same point
#pragma unroll
  for (int p = 0; p < P; ++p)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      double v = tot[p][c];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      tot[p][c] = v;
    }
  if (g == 0) {
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int64_t pt = base + 16 * p + j;
      if (pt < N) {
#pragma unroll
        for (int c = 0; c < NC; ++c) store(out + c * out_ld + pt, tot[p][c] * scale);
      }
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

namespace {

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
int mfma_tiles(int64_t M) { return static_cast<int>((M + 15) / 16); }

// Table layout for dim <= 3: [ModeRec x M][a_tab: tiles x 64 f64][amp_tab: tiles x 3 x 16 f32]
size_t rec_bytes(int64_t M) { return align256(sizeof(ModeRec) * static_cast<size_t>(M > 0 ? M : 1)); }
size_t atab_bytes(int64_t M) { return align256(sizeof(double) * 64 * static_cast<size_t>(mfma_tiles(M))); }
size_t amptab_bytes(int64_t M) { return align256(sizeof(float) * 48 * static_cast<size_t>(mfma_tiles(M))); }
const double* atab_of(const void* t, int64_t M) {
  return reinterpret_cast<const double*>(static_cast<const char*>(t) + rec_bytes(M));
}
const float* amptab_of(const void* t, int64_t M) {
  return reinterpret_cast<const float*>(static_cast<const char*>(t) + rec_bytes(M) + atab_bytes(M));
}

template <int D, int NC, bool GRID, typename OutT>
hipError_t mfma_t(int64_t M, int64_t N, const void* table, const double* pos, int64_t pos_ld,
                  const double* ax1, const double* ax2, int64_t s1, int64_t s2, void* out,
                  int64_t out_ld, double scale, hipStream_t s) {
  constexpr int P = NC == 1 ? 8 : 4;  // points per lane group: 16 P per wave
  const unsigned g = grid_for(N, (kBlock / 64) * 16 * P);
  k_direct_mfma<D, NC, P, GRID, OutT><<<g, kBlock, 0, s>>>(
      atab_of(table, M), amptab_of(table, M), mfma_tiles(M), pos, pos_ld, ax1, ax2, s1, s2, N,
      static_cast<OutT*>(out), out_ld, scale);
  return hipGetLastError();
}

}  // namespace

size_t mode_table_bytes(int dim, int64_t M) {
  if (dim > 3) return align256(static_cast<size_t>(dim + 2) * sizeof(double) *
                               static_cast<size_t>(M > 0 ? M : 1));
  return rec_bytes(M) + atab_bytes(M) + amptab_bytes(M);
}

PrepArgs prep_args(int kind, int dim, int64_t M, const double* cov, const double* z1,
                   const double* z2, const double* sf, void* table) {
  PrepArgs a;
  a.kind = kind;
  a.dim = dim;
  a.M = M;
  a.ntiles = M > 0 ? mfma_tiles(M) : 0;
  a.cov = cov;
  a.z1 = z1;
  a.z2 = z2;
  a.sf = sf;
  a.table = static_cast<ModeRec*>(table);
  a.a_tab = const_cast<double*>(atab_of(table, M));
  a.amp_tab = const_cast<float*>(amptab_of(table, M));
  return a;
}

hipError_t launch_prep(int kind, int dim, int64_t M, const double* cov, const double* z1,
                       const double* z2, const double* sf, void* table, hipStream_t stream) {
  if (M <= 0) return hipSuccess;
  if (dim <= 3) {
    const int nt = mfma_tiles(M);
    k_prep_small<<<grid_for(static_cast<int64_t>(nt) * 16, 256), 256, 0, stream>>>(
        kind, dim, M, nt, cov, z1, z2, sf, static_cast<ModeRec*>(table),
        const_cast<double*>(atab_of(table, M)), const_cast<float*>(amptab_of(table, M)));
  } else {
    k_prep_generic<<<grid_for(M, 256), 256, 0, stream>>>(kind, dim, M, cov, z1, z2, sf,
                                                          static_cast<double*>(table));
  }
  return hipGetLastError();
}

hipError_t launch_direct_scalar(int dim, int64_t M, int64_t N, const void* table,
                                const double* pos, int64_t pos_ld, void* out, int out_dtype,
                                double scale, int variant, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  const bool f32 = out_dtype == 1;
  if (dim <= 3 && variant == 0) {
#define GSC_M(D) (f32 ? mfma_t<D, 1, false, float>(M, N, table, pos, pos_ld, nullptr, nullptr, 1, 1, out, 0, scale, s) \
                      : mfma_t<D, 1, false, double>(M, N, table, pos, pos_ld, nullptr, nullptr, 1, 1, out, 0, scale, s))
    if (dim == 1) return GSC_M(1);
    if (dim == 2) return GSC_M(2);
    return GSC_M(3);
#undef GSC_M
  }
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
                                 int out_dtype, double scale, int variant, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  const bool f32 = out_dtype == 1;
  if (variant == 0) {
    if (dim == 2)
      return f32 ? mfma_t<2, 2, false, float>(M, N, table, pos, pos_ld, nullptr, nullptr, 1, 1, out, out_ld, scale, s)
                 : mfma_t<2, 2, false, double>(M, N, table, pos, pos_ld, nullptr, nullptr, 1, 1, out, out_ld, scale, s);
    return f32 ? mfma_t<3, 3, false, float>(M, N, table, pos, pos_ld, nullptr, nullptr, 1, 1, out, out_ld, scale, s)
               : mfma_t<3, 3, false, double>(M, N, table, pos, pos_ld, nullptr, nullptr, 1, 1, out, out_ld, scale, s);
  }
  if (dim == 2)
    return f32 ? incompr_t<2, float>(M, N, table, pos, pos_ld, out, out_ld, scale, s)
               : incompr_t<2, double>(M, N, table, pos, pos_ld, out, out_ld, scale, s);
  return f32 ? incompr_t<3, float>(M, N, table, pos, pos_ld, out, out_ld, scale, s)
             : incompr_t<3, double>(M, N, table, pos, pos_ld, out, out_ld, scale, s);
}

hipError_t launch_direct_grid(int dim, int64_t M, const int64_t* shape, const double* const* axes,
                              const void* table, void* out, int out_dtype, double scale,
                              int variant, hipStream_t s) {
  int64_t N = 1;
  for (int i = 0; i < dim; ++i) N *= shape[i];
  if (N <= 0) return hipSuccess;
  const ModeRec* md = static_cast<const ModeRec*>(table);
  const int Mi = static_cast<int>(M);
  const int64_t s1 = dim > 1 ? shape[1] : 1, s2 = dim > 2 ? shape[2] : 1;
  const double* a0 = axes[0];
  const double* a1 = dim > 1 ? axes[1] : nullptr;
  const double* a2 = dim > 2 ? axes[2] : nullptr;
  const bool f32 = out_dtype == 1;
  if (variant == 0) {
#define GSC_MG(D) (f32 ? mfma_t<D, 1, true, float>(M, N, table, a0, 0, a1, a2, s1, s2, out, 0, scale, s) \
                       : mfma_t<D, 1, true, double>(M, N, table, a0, 0, a1, a2, s1, s2, out, 0, scale, s))
    if (dim == 1) return GSC_MG(1);
    if (dim == 2) return GSC_MG(2);
    return GSC_MG(3);
#undef GSC_MG
  }
  constexpr int P = 8;
  const unsigned g = grid_for(N, kBlock * P);
#define GSC_GRID(D, T) \
  k_direct_grid<D, P, T><<<g, kBlock, 0, s>>>(md, Mi, a0, a1, a2, s1, s2, N, static_cast<T*>(out), scale)
  if (dim == 1) { if (f32) GSC_GRID(1, float); else GSC_GRID(1, double); }
  else if (dim == 2) { if (f32) GSC_GRID(2, float); else GSC_GRID(2, double); }
  else { if (f32) GSC_GRID(3, float); else GSC_GRID(3, double); }
#undef GSC_GRID
  return hipGetLastError();
}

}  // namespace gsc
