// Factored summation on structured grids, contracted on the MFMA cores.
//
// On a grid x = (x_a, y_b, z_c) the per-mode phase separates, and with the
// amplitude-phase form used throughout this library
//   z1 cos(phi) + z2 sin(phi) = A sin(phi + theta),   A = hypot(z1, z2),
//   theta = atan2(z1, z2),
// the field is the imaginary part of a product of per-axis tables:
//   u(a, b, c) = Im sum_m P(a, m) Q(b, m) R(c, m)
//   P(a, m) = A_m e^{i(theta_m + k_xm x_a)},  Q(b, m) = e^{i k_ym y_b},
//   R(c, m) = e^{i k_zm z_c}.
// With W = P Q (complex, per row (a, b)) and K = 2M interleaved as
// k = 2m -> (Re W, Im R), k = 2m + 1 -> (Im W, Re R):
//   u(row, c) = sum_k A(row, k) B(k, c),  A = [Re W | Im W],  B = [Im R ; Re R]
// a real GEMM (rows = Nx Ny, cols = Nz).  In 2-D, W = P and B is built from Q.
//
// Tables are evaluated in FP64 (phase in revolutions, exact fract range
// reduction, sincospi) and rounded to FP32.  The MFMA operands are split
// into NS parts of T (fp16 or bf16): x = sum_s x_s, x_0 = T(x),
// x_{s+1} = T(x - sum_{t<=s} x_t) (each difference exact in FP32), and the
// products with i + j < NS are accumulated in FP32 on the matrix cores:
// NS = 1 -> 1 product, NS = 2 -> 3 products (x3), NS = 3 -> 6 products (x6).
//
// Kernel shape: 256 threads (4 waves, 2 x 2), block tile 128 x 128, wave tile
// 64 x 64 = 2 x 2 v_mfma_f32_32x32x16 blocks, K chunks of 32 (16 modes)
// staged in LDS.  The A tile is generated from P and Q in the chunk prologue
// (the Khatri-Rao operand is never written to HBM).
#include "gsc_internal.h"
#include "gsc_factored.h"

#include <hip/hip_runtime.h>

namespace gsc {
namespace {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int kTM = 128;   // block rows
constexpr int kTN = 128;   // block cols
constexpr int kKC = 32;    // k per chunk (16 modes)
constexpr int kPad = 8;    // LDS row padding (elements)
constexpr int kLd = kKC + kPad;
// Accumulation (ACC template argument).  The MFMA accumulates in FP32; over
// K = 2M terms its rounding grows with the partial sum (~8e-6 sigma at
// M = 4000).  Every kFlush chunks (64 modes) the FP32 accumulators are added
// into a second set and cleared, so FP32 rounding only ever sees a 64-mode
// partial sum:
//   ACC_F64   second set in FP64  (most accurate, +128 VGPRs)
//   ACC_F32X2 second set in FP32  (~1e-6 sigma at M = 4000, +64 VGPRs)
//   ACC_F32   no second set       (fastest, error grows with M)
constexpr int kFlush = 4;
enum { ACC_F64 = 0, ACC_F32X2 = 1, ACC_F32 = 2 };
template <int ACC> struct AccT { using type = double; };
template <> struct AccT<ACC_F32X2> { using type = float; };
template <> struct AccT<ACC_F32> { using type = float; };

template <typename T> struct Vec8;
template <> struct Vec8<_Float16> { using type = half8_t; };
template <> struct Vec8<__bf16> { using type = bf16x8_t; };

__device__ __forceinline__ f32x16_t mfma(half8_t a, half8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16_t mfma(bf16x8_t a, bf16x8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

template <typename T, int NS>
__device__ __forceinline__ void split(float x, T (&o)[NS]) {
  float r = x;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const T h = static_cast<T>(r);
    o[s] = h;
    r = r - static_cast<float>(h);
  }
}

// ---------------------------------------------------------------------------
// Per-axis complex table: tab[a * mp + m] = amp_m e^{2 pi i (kap_{axis,m} x_a + th_m)}
// (amp = A, th = theta/2pi when with_amp, else amp = 1, th = 0), FP64 -> FP32.
// Modes m >= M are zero.  One thread per (a, m), 1-D grid over n * mp.
// ---------------------------------------------------------------------------
__global__ void k_axis_table(const ModeRec* __restrict__ modes, int M, int mp, int axis,
                             const double* __restrict__ x, int64_t n, int with_amp,
                             float2* __restrict__ tab) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= n * mp) return;
  const int64_t a = idx / mp;
  const int m = static_cast<int>(idx - a * mp);
  float2 v = make_float2(0.0f, 0.0f);
  if (m < M) {
    const ModeRec md = modes[m];
    double t = md.kap[axis] * x[a];
    double amp = 1.0;
    if (with_amp) {
      t += md.th;
      amp = static_cast<double>(md.amp[0]);
    }
    const double r = t - rint(t);  // exact: |r| <= 1/2 revolution
    double s, c;
    sincospi(2.0 * r, &s, &c);
    v = make_float2(static_cast<float>(amp * c), static_cast<float>(amp * s));
  }
  tab[a * mp + m] = v;
}

// B operand: bt[s][col][k] for col < ncp (padded), k < kp = 2 mp:
// k = 2m -> Im T(col, m), k = 2m + 1 -> Re T(col, m); zero for col >= ncols.
template <typename T, int NS>
__global__ void k_split_b(const float2* __restrict__ tab, int64_t ncols, int64_t ncp, int mp,
                          T* __restrict__ bt) {
  const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (idx >= ncp * mp) return;
  const int64_t col = idx / mp;
  const int m = static_cast<int>(idx - col * mp);
  float2 v = make_float2(0.0f, 0.0f);
  if (col < ncols) v = tab[col * mp + m];
  T im[NS], re[NS];
  split<T, NS>(v.y, im);
  split<T, NS>(v.x, re);
  const int64_t kp = 2 * static_cast<int64_t>(mp);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    T* row = bt + (s * ncp + col) * kp;
    row[2 * m] = im[s];
    row[2 * m + 1] = re[s];
  }
}

// ---------------------------------------------------------------------------
// The GEMM.  rows = n_rows (KR: n_rows = na * nb, row = a * nb + b), cols.
// out[row * ldo + col] = scale * sum_k A(row, k) B(k, col).
// ---------------------------------------------------------------------------
template <typename T, int NS, bool KR, typename OutT, int ACC>
__global__ void __launch_bounds__(256)
k_factored_gemm(const float2* __restrict__ P, const float2* __restrict__ Q, int64_t nb,
                int64_t n_rows, int mp, const T* __restrict__ bt, int64_t ncp, int64_t ncols,
                OutT* __restrict__ out, int64_t ldo, double scale) {
  using V8 = typename Vec8<T>::type;
  __shared__ __attribute__((aligned(16))) T As[NS][kTM][kLd];
  __shared__ __attribute__((aligned(16))) T Bs[NS][kTN][kLd];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t row0 = static_cast<int64_t>(blockIdx.x) * kTM;
  const int64_t col0 = static_cast<int64_t>(blockIdx.y) * kTN;
  const int64_t kp = 2 * static_cast<int64_t>(mp);

  // generator assignment: thread -> (row gr, 8 modes at half gh of the chunk)
  const int gr = tid >> 1, gh = tid & 1;
  const int64_t grow = row0 + gr;
  const bool grow_ok = grow < n_rows;
  const int64_t ga = KR ? (grow_ok ? grow / nb : 0) : (grow_ok ? grow : 0);
  const int64_t gb = KR ? (grow_ok ? grow - ga * nb : 0) : 0;
  const float2* prow = P + ga * mp;
  const float2* qrow = KR ? Q + gb * mp : nullptr;
  // B loader assignment: thread -> (col bc, k half bh)
  const int bc = tid >> 1, bh = tid & 1;

  using A2 = typename AccT<ACC>::type;
  f32x16_t acc[2][2];
  A2 acc2[2][2][ACC == ACC_F32 ? 1 : 16];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
#pragma unroll
      for (int v = 0; v < (ACC == ACC_F32 ? 1 : 16); ++v) acc2[i][j][v] = 0;
    }
  auto flush = [&]() {
    if constexpr (ACC != ACC_F32) {
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
          for (int v = 0; v < 16; ++v) {
            acc2[i][j][v] += static_cast<A2>(acc[i][j][v]);
            acc[i][j][v] = 0.0f;
          }
    }
  };
  int since_flush = 0;

  for (int m0 = 0; m0 < mp; m0 += kKC / 2) {
    // ---- A tile: W(row, m0 + 8 gh + 0..7) -> k = 16 gh + 0..15
    {
      T ws[NS][16];
      if (grow_ok) {
        const float4* p4 = reinterpret_cast<const float4*>(prow + m0 + 8 * gh);
        const float4* q4 = KR ? reinterpret_cast<const float4*>(qrow + m0 + 8 * gh) : nullptr;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float4 p = p4[i];  // two modes: (re, im, re, im)
          float wr0 = p.x, wi0 = p.y, wr1 = p.z, wi1 = p.w;
          if constexpr (KR) {
            const float4 q = q4[i];
            const float r0 = wr0 * q.x - wi0 * q.y, i0 = wr0 * q.y + wi0 * q.x;
            const float r1 = wr1 * q.z - wi1 * q.w, i1 = wr1 * q.w + wi1 * q.z;
            wr0 = r0; wi0 = i0; wr1 = r1; wi1 = i1;
          }
          T t[NS];
          split<T, NS>(wr0, t);
#pragma unroll
          for (int s = 0; s < NS; ++s) ws[s][4 * i + 0] = t[s];
          split<T, NS>(wi0, t);
#pragma unroll
          for (int s = 0; s < NS; ++s) ws[s][4 * i + 1] = t[s];
          split<T, NS>(wr1, t);
#pragma unroll
          for (int s = 0; s < NS; ++s) ws[s][4 * i + 2] = t[s];
          split<T, NS>(wi1, t);
#pragma unroll
          for (int s = 0; s < NS; ++s) ws[s][4 * i + 3] = t[s];
        }
      } else {
#pragma unroll
        for (int s = 0; s < NS; ++s)
#pragma unroll
          for (int i = 0; i < 16; ++i) ws[s][i] = static_cast<T>(0.0f);
      }
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        V8* dst = reinterpret_cast<V8*>(&As[s][gr][16 * gh]);
        V8 lo, hi;
#pragma unroll
        for (int i = 0; i < 8; ++i) { lo[i] = ws[s][i]; hi[i] = ws[s][8 + i]; }
        dst[0] = lo;
        dst[1] = hi;
      }
    }
    // ---- B tile: bt[s][col0 + bc][2 m0 + 16 bh + 0..15]
    {
      const int64_t k0 = 2 * static_cast<int64_t>(m0) + 16 * bh;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        const V8* src = reinterpret_cast<const V8*>(bt + (s * ncp + col0 + bc) * kp + k0);
        V8* dst = reinterpret_cast<V8*>(&Bs[s][bc][16 * bh]);
        dst[0] = src[0];
        dst[1] = src[1];
      }
    }
    __syncthreads();
    // ---- MFMAs: two k-steps of 16
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      V8 af[NS][2], bf[NS][2];
      const int kk = ks * 16 + 8 * (lane >> 5);
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          af[s][i] = *reinterpret_cast<const V8*>(&As[s][wr * 64 + i * 32 + (lane & 31)][kk]);
          bf[s][i] = *reinterpret_cast<const V8*>(&Bs[s][wc * 64 + i * 32 + (lane & 31)][kk]);
        }
#pragma unroll
      for (int sa = 0; sa < NS; ++sa)
#pragma unroll
        for (int sb = 0; sb < NS - sa; ++sb)
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma(af[sa][i], bf[sb][j], acc[i][j]);
    }
    __syncthreads();
    if (++since_flush == kFlush) {
      flush();
      since_flush = 0;
    }
  }
  flush();

  // ---- epilogue: D[8 (v/4) + 4 (lane/32) + v%4][lane%32] of each 32 x 32 block
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int64_t col = col0 + wc * 64 + j * 32 + (lane & 31);
      if (col >= ncols) continue;
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int64_t row = row0 + wr * 64 + i * 32 + 8 * (v >> 2) + 4 * (lane >> 5) + (v & 3);
        double val;
        if constexpr (ACC == ACC_F32) val = static_cast<double>(acc[i][j][v]);
        else val = static_cast<double>(acc2[i][j][v]);
        if (row < n_rows) out[row * ldo + col] = static_cast<OutT>(val * scale);
      }
    }
}

template <typename T, int NS, typename OutT, int ACC>
void gemm_acc(bool kr, dim3 grid, const float2* P, const float2* Q, int64_t nb, int64_t rows,
              int mp, const void* bt, int64_t ncp, int64_t ncols, void* out, int64_t ldo,
              double scale, hipStream_t s) {
  if (kr)
    k_factored_gemm<T, NS, true, OutT, ACC><<<grid, 256, 0, s>>>(
        P, Q, nb, rows, mp, static_cast<const T*>(bt), ncp, ncols, static_cast<OutT*>(out), ldo,
        scale);
  else
    k_factored_gemm<T, NS, false, OutT, ACC><<<grid, 256, 0, s>>>(
        P, Q, nb, rows, mp, static_cast<const T*>(bt), ncp, ncols, static_cast<OutT*>(out), ldo,
        scale);
}

template <typename T, int NS, typename OutT>
hipError_t gemm_t(bool kr, int accum, const float2* P, const float2* Q, int64_t nb, int64_t rows,
                  int mp, const void* bt, int64_t ncp, int64_t ncols, void* out, int64_t ldo,
                  double scale, hipStream_t s) {
  dim3 grid(static_cast<unsigned>((rows + kTM - 1) / kTM), static_cast<unsigned>(ncp / kTN));
  if (accum == ACC_F32X2)
    gemm_acc<T, NS, OutT, ACC_F32X2>(kr, grid, P, Q, nb, rows, mp, bt, ncp, ncols, out, ldo, scale, s);
  else if (accum == ACC_F32)
    gemm_acc<T, NS, OutT, ACC_F32>(kr, grid, P, Q, nb, rows, mp, bt, ncp, ncols, out, ldo, scale, s);
  else
    gemm_acc<T, NS, OutT, ACC_F64>(kr, grid, P, Q, nb, rows, mp, bt, ncp, ncols, out, ldo, scale, s);
  return hipGetLastError();
}

template <typename T, int NS>
hipError_t split_t(const float2* tab, int64_t ncols, int64_t ncp, int mp, void* bt,
                   hipStream_t s) {
  const unsigned grid = static_cast<unsigned>((ncp * mp + 255) / 256);
  k_split_b<T, NS><<<grid, 256, 0, s>>>(tab, ncols, ncp, mp, static_cast<T*>(bt));
  return hipGetLastError();
}

}  // namespace

FactoredPlan plan_factored(int dim, int64_t M, const int64_t* shape, int precision) {
  FactoredPlan p{};
  p.dim = dim;
  p.ns = precision_splits(precision);
  p.bf16 = precision_is_bf16(precision);
  p.mp = static_cast<int>(((M > 0 ? M : 1) + 15) / 16 * 16);
  p.kp = 2 * static_cast<int64_t>(p.mp);
  if (dim == 2) {
    p.na = shape[0];
    p.nb = 1;
    p.rows = shape[0];
    p.ncols = shape[1];
  } else {
    p.na = shape[0];
    p.nb = shape[1];
    p.rows = shape[0] * shape[1];
    p.ncols = shape[2];
  }
  p.ncp = (p.ncols + kTN - 1) / kTN * kTN;
  auto al = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
  size_t off = al(mode_table_bytes(3, M));
  p.off_p = off;
  off += al(sizeof(float2) * static_cast<size_t>(p.na) * p.mp);
  p.off_q = off;
  if (dim == 3) off += al(sizeof(float2) * static_cast<size_t>(p.nb) * p.mp);
  p.off_r = off;
  off += al(sizeof(float2) * static_cast<size_t>(p.ncols) * p.mp);
  p.off_bt = off;
  off += al(2 * static_cast<size_t>(p.ns) * p.ncp * p.kp);
  p.bytes = off;
  return p;
}

hipError_t launch_factored(const FactoredPlan& p, int64_t M, const double* const* axes,
                           void* work, void* out, int out_dtype, double scale, int accum,
                           hipStream_t s) {
  char* w = static_cast<char*>(work);
  const ModeRec* modes = reinterpret_cast<const ModeRec*>(w);
  float2* P = reinterpret_cast<float2*>(w + p.off_p);
  float2* Q = reinterpret_cast<float2*>(w + p.off_q);
  float2* R = reinterpret_cast<float2*>(w + p.off_r);
  void* bt = w + p.off_bt;
  const int Mi = static_cast<int>(M);
  // axis tables (the mode table was written by launch_prep)
  auto tgrid = [&](int64_t n) { return static_cast<unsigned>((n * p.mp + 255) / 256); };
  k_axis_table<<<tgrid(p.na), 256, 0, s>>>(modes, Mi, p.mp, 0, axes[0], p.na, 1, P);
  if (p.dim == 3)
    k_axis_table<<<tgrid(p.nb), 256, 0, s>>>(modes, Mi, p.mp, 1, axes[1], p.nb, 0, Q);
  const int last = p.dim - 1;
  k_axis_table<<<tgrid(p.ncols), 256, 0, s>>>(modes, Mi, p.mp, last, axes[last], p.ncols, 0, R);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // B operand from the last axis
  if (p.bf16) {
    e = p.ns == 1 ? split_t<__bf16, 1>(R, p.ncols, p.ncp, p.mp, bt, s)
      : p.ns == 2 ? split_t<__bf16, 2>(R, p.ncols, p.ncp, p.mp, bt, s)
                  : split_t<__bf16, 3>(R, p.ncols, p.ncp, p.mp, bt, s);
  } else {
    e = p.ns == 1 ? split_t<_Float16, 1>(R, p.ncols, p.ncp, p.mp, bt, s)
      : p.ns == 2 ? split_t<_Float16, 2>(R, p.ncols, p.ncp, p.mp, bt, s)
                  : split_t<_Float16, 3>(R, p.ncols, p.ncp, p.mp, bt, s);
  }
  if (e != hipSuccess) return e;
  const bool kr = p.dim == 3;
  const double sc = scale;
  const bool f32 = out_dtype == 1;
#define GSC_GEMM(T, NS)                                                                        \
  (f32 ? gemm_t<T, NS, float>(kr, accum, P, Q, p.nb, p.rows, p.mp, bt, p.ncp, p.ncols, out,    \
                              p.ncols, sc, s)                                                  \
       : gemm_t<T, NS, double>(kr, accum, P, Q, p.nb, p.rows, p.mp, bt, p.ncp, p.ncols, out,   \
                               p.ncols, sc, s))
  if (p.bf16) {
    if (p.ns == 1) return GSC_GEMM(__bf16, 1);
    if (p.ns == 2) return GSC_GEMM(__bf16, 2);
    return GSC_GEMM(__bf16, 3);
  }
  if (p.ns == 1) return GSC_GEMM(_Float16, 1);
  if (p.ns == 2) return GSC_GEMM(_Float16, 2);
  return GSC_GEMM(_Float16, 3);
#undef GSC_GEMM
}

}  // namespace gsc
