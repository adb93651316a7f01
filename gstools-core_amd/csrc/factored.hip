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
// staged in LDS.  The A tile is generated from P and Q while the previous
// chunk's MFMAs run (the Khatri-Rao operand is never written to HBM).
//
// Table layouts are chunk-major so that one chunk of one block is contiguous:
//   P/Q/R: tab[(c * n + a) * 16 + mm]            mode m = 16 c + mm, a < n
//   B:     bt[((s * nch + c) * ncp + col) * 32 + kk], kk = 2 mm (+1 for Re)
// with the mode count padded to mp = 32 * ceil(M / 32) (nch = mp / 16 chunks,
// an even number; padded modes are zero).
#include "gsc_internal.h"
#include "gsc_factored.h"

#include <hip/hip_runtime.h>

#include <algorithm>

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
// M = 4000).  Every kFlush chunks (256 modes) the FP32 accumulators are added
// into a second set and cleared, so FP32 rounding only ever sees a 256-mode
// partial sum (256 modes):
//   ACC_F64   second set in FP64  (most accurate, +128 VGPRs)
//   ACC_F32X2 second set in FP32  (~1e-6 sigma at M = 4000, +64 VGPRs)
//   ACC_F32   no second set       (fastest, error grows with M)
constexpr int kFlush = 16;  // even: the main loop is unrolled by two
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
// Per-axis complex tables, tab[m / 16][a][m % 16] = amp_m e^{2 pi i (kap_{axis,m} x_a + th_m)}
// (amp = A, th = theta/2pi for axis 0; amp = 1, th = 0 otherwise), FP64 -> FP32.
// Modes m >= M are zero.
// ---------------------------------------------------------------------------
// All tables of one call in one launch.  blockIdx.z selects the table:
//   0: P (axis 0, amplitude and theta folded in), chunk-major float2
//   1: Q (axis 1, 3-D only), chunk-major float2
//   2: the B operand, split straight from the FP64 value of R (last axis),
//      columns ncols <= col < ncp zero-filled.
// block = 16 rows x 16 modes of one chunk: grid (ceil(max(na, nb, ncp) / 16), mp / 16, 3).
template <typename T, int NS>
__global__ void __launch_bounds__(256)
k_tables(const ModeRec* __restrict__ modes, int M, int dim, const double* __restrict__ x0,
         const double* __restrict__ x1, const double* __restrict__ xl, int64_t na, int64_t nb,
         int64_t ncols, int64_t ncp, float2* __restrict__ P, float2* __restrict__ Q,
         T* __restrict__ bt) {
  const int which = blockIdx.z;
  if (which == 1 && dim != 3) return;
  const int64_t a = static_cast<int64_t>(blockIdx.x) * 16 + (threadIdx.x >> 4);
  const int c = blockIdx.y;
  const int nch = gridDim.y;
  const int mm = threadIdx.x & 15;
  const int m = c * 16 + mm;
  const int64_t n = which == 0 ? na : (which == 1 ? nb : ncp);
  if (a >= n) return;
  const double* x = which == 0 ? x0 : (which == 1 ? x1 : xl);
  const int axis = which == 2 ? dim - 1 : which;
  double re = 0.0, im = 0.0;
  if (m < M && (which != 2 || a < ncols)) {
    const ModeRec md = modes[m];
    double t = md.kap[axis] * x[a];
    double amp = 1.0;
    if (which == 0) {
      t += md.th;
      amp = static_cast<double>(md.amp[0]);
    }
    const double r = t - rint(t);  // exact: |r| <= 1/2 revolution
    double sn, cs;
    sincospi(2.0 * r, &sn, &cs);
    re = amp * cs;
    im = amp * sn;
  }
  if (which < 2) {
    float2* tab = which == 0 ? P : Q;
    tab[(static_cast<int64_t>(c) * n + a) * 16 + mm] =
        make_float2(static_cast<float>(re), static_cast<float>(im));
    return;
  }
  double ri = im, rr = re;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const T hi = static_cast<T>(static_cast<float>(ri));
    const T hr = static_cast<T>(static_cast<float>(rr));
    ri -= static_cast<double>(static_cast<float>(hi));
    rr -= static_cast<double>(static_cast<float>(hr));
    T* row = bt + ((static_cast<int64_t>(s) * nch + c) * ncp + a) * 32;
    row[2 * mm] = hi;
    row[2 * mm + 1] = hr;
  }
}

// ---------------------------------------------------------------------------
// The GEMM.  rows = n_rows (KR: n_rows = na * nb, row = a * nb + b), cols.
// out[row * ldo + col] = scale * sum_k A(row, k) B(k, col).
//
// Software pipeline, one barrier per K chunk, two register sets for the
// global loads and two LDS buffers: in the step for chunk c the wave
//   1. issues the global loads of chunk c + 2 (P, Q rows, B tile) into the
//      register set freed by chunk c,
//   2. reads chunk c's A/B fragments from LDS buffer c % 2 and issues its
//      MFMAs,
//   3. generates chunk c + 1's A tile from the register set loaded one step
//      earlier (complex products, splits) and writes it with the B tile to
//      the other LDS buffer, interleaved with the MFMAs of step 2.
// The loop is unrolled by two so buffer and register-set indices are
// compile-time constants.
// ---------------------------------------------------------------------------
template <typename T, int NS, bool KR, typename OutT, int ACC>
__global__ void __launch_bounds__(256)
k_factored_gemm(const float2* __restrict__ P, const float2* __restrict__ Q, int64_t na,
                int64_t nb, int64_t n_rows, int nch, const T* __restrict__ bt, int64_t ncp,
                int64_t ncols, OutT* __restrict__ out, int64_t ldo, double scale) {
  using V8 = typename Vec8<T>::type;
  __shared__ __attribute__((aligned(16))) T As[2][NS][kTM][kLd];
  __shared__ __attribute__((aligned(16))) T Bs[2][NS][kTN][kLd];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t row0 = static_cast<int64_t>(blockIdx.x) * kTM;
  const int64_t col0 = static_cast<int64_t>(blockIdx.y) * kTN;

  // generator assignment: thread -> (row gr, 8 modes at half gh of the chunk)
  const int gr = tid >> 1, gh = tid & 1;
  const int64_t grow = row0 + gr;
  const bool grow_ok = grow < n_rows;
  const int64_t ga = KR ? (grow_ok ? grow / nb : 0) : (grow_ok ? grow : 0);
  const int64_t gb = KR ? (grow_ok ? grow - ga * nb : 0) : 0;
  // float4 = two modes; chunk c of row a starts at float2 index (c n + a) 16
  const float4* prow = reinterpret_cast<const float4*>(P + ga * 16 + 8 * gh);
  const float4* qrow = KR ? reinterpret_cast<const float4*>(Q + gb * 16 + 8 * gh) : nullptr;
  const int64_t pstep = na * 8, qstep = nb * 8;  // float4 per chunk
  // B loader assignment: thread -> (col bc, k half bh); chunk stride ncp * 32
  const int bc = tid >> 1, bh = tid & 1;
  const T* brow = bt + (col0 + bc) * 32 + 16 * bh;
  const int64_t bstep = ncp * 32, bsplit = static_cast<int64_t>(nch) * ncp * 32;

  float4 pr[2][4], qr[2][4];
  V8 br[2][NS][2];
  auto gload = [&](const int set, int c) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      pr[set][i] = prow[c * pstep + i];
      if constexpr (KR) qr[set][i] = qrow[c * qstep + i];
    }
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const V8* src = reinterpret_cast<const V8*>(brow + s * bsplit + c * bstep);
      br[set][s][0] = src[0];
      br[set][s][1] = src[1];
    }
  };
  auto gstore = [&](const int set, const int buf) {
    T ws[NS][16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      float wr0 = pr[set][i].x, wi0 = pr[set][i].y, wr1 = pr[set][i].z, wi1 = pr[set][i].w;
      if constexpr (KR) {
        const float4 q = qr[set][i];
        const float r0 = wr0 * q.x - wi0 * q.y, i0 = wr0 * q.y + wi0 * q.x;
        const float r1 = wr1 * q.z - wi1 * q.w, i1 = wr1 * q.w + wi1 * q.z;
        wr0 = r0; wi0 = i0; wr1 = r1; wi1 = i1;
      }
      if (!grow_ok) wr0 = wi0 = wr1 = wi1 = 0.0f;
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
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      V8* dst = reinterpret_cast<V8*>(&As[buf][s][gr][16 * gh]);
      V8 lo, hi;
#pragma unroll
      for (int i = 0; i < 8; ++i) { lo[i] = ws[s][i]; hi[i] = ws[s][8 + i]; }
      dst[0] = lo;
      dst[1] = hi;
      V8* bd = reinterpret_cast<V8*>(&Bs[buf][s][bc][16 * bh]);
      bd[0] = br[set][s][0];
      bd[1] = br[set][s][1];
    }
  };

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

  auto mfma_chunk = [&](const int buf) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      V8 af[NS][2], bf[NS][2];
      const int kk = ks * 16 + 8 * (lane >> 5);
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          af[s][i] = *reinterpret_cast<const V8*>(&As[buf][s][wr * 64 + i * 32 + (lane & 31)][kk]);
          bf[s][i] = *reinterpret_cast<const V8*>(&Bs[buf][s][wc * 64 + i * 32 + (lane & 31)][kk]);
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
  };
  // Interleave hints for one step: the loads first, the first k-step's LDS
  // reads, then every MFMA paired with LDS reads, generator VALU work and
  // LDS writes of the next chunk.
  constexpr int kMfma = 2 * 4 * (NS * (NS + 1) / 2);
  auto sched = [&]() {
    __builtin_amdgcn_sched_group_barrier(0x020, 4 * (KR ? 2 : 1) + 2 * NS, 0);  // VMEM read
    __builtin_amdgcn_sched_group_barrier(0x100, 4 * NS, 0);                    // DS read
#pragma unroll
    for (int i = 0; i < kMfma; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      if (i < 4 * NS) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);  // VALU
      if (i >= kMfma - 4 * NS) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // DS write
    }
  };

  gload(0, 0);
  gload(1, nch > 1 ? 1 : 0);
  gstore(0, 0);
  __syncthreads();
  for (int c0 = 0; c0 < nch; c0 += kFlush) {
    const int c1 = min(c0 + kFlush, nch);
    for (int c = c0; c < c1; c += 2) {
      gload(0, min(c + 2, nch - 1));
      mfma_chunk(0);
      gstore(1, 1);
      sched();
      __syncthreads();
      gload(1, min(c + 3, nch - 1));
      mfma_chunk(1);
      gstore(0, 0);
      sched();
      __syncthreads();
    }
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
  }

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
void gemm_acc(bool kr, dim3 grid, const float2* P, const float2* Q, int64_t na, int64_t nb,
              int64_t rows, int nch, const void* bt, int64_t ncp, int64_t ncols, void* out,
              int64_t ldo, double scale, hipStream_t s) {
  if (kr)
    k_factored_gemm<T, NS, true, OutT, ACC><<<grid, 256, 0, s>>>(
        P, Q, na, nb, rows, nch, static_cast<const T*>(bt), ncp, ncols, static_cast<OutT*>(out),
        ldo, scale);
  else
    k_factored_gemm<T, NS, false, OutT, ACC><<<grid, 256, 0, s>>>(
        P, Q, na, nb, rows, nch, static_cast<const T*>(bt), ncp, ncols, static_cast<OutT*>(out),
        ldo, scale);
}

template <typename T, int NS, typename OutT>
hipError_t gemm_t(bool kr, int accum, const float2* P, const float2* Q, int64_t na, int64_t nb,
                  int64_t rows, int nch, const void* bt, int64_t ncp, int64_t ncols, void* out,
                  int64_t ldo, double scale, hipStream_t s) {
  dim3 grid(static_cast<unsigned>((rows + kTM - 1) / kTM), static_cast<unsigned>(ncp / kTN));
  if (accum == ACC_F32X2)
    gemm_acc<T, NS, OutT, ACC_F32X2>(kr, grid, P, Q, na, nb, rows, nch, bt, ncp, ncols, out, ldo,
                                     scale, s);
  else if (accum == ACC_F32)
    gemm_acc<T, NS, OutT, ACC_F32>(kr, grid, P, Q, na, nb, rows, nch, bt, ncp, ncols, out, ldo,
                                   scale, s);
  else
    gemm_acc<T, NS, OutT, ACC_F64>(kr, grid, P, Q, na, nb, rows, nch, bt, ncp, ncols, out, ldo,
                                   scale, s);
  return hipGetLastError();
}

}  // namespace

FactoredPlan plan_factored(int dim, int64_t M, const int64_t* shape, int precision) {
  FactoredPlan p{};
  p.dim = dim;
  p.ns = precision_splits(precision);
  p.bf16 = precision_is_bf16(precision);
  p.mp = static_cast<int>(((M > 0 ? M : 1) + 31) / 32 * 32);
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
  void* bt = w + p.off_bt;
  const int Mi = static_cast<int>(M);
  // tables (the mode table was written by launch_prep): one launch
  const int64_t nmax = std::max(std::max(p.na, p.dim == 3 ? p.nb : 0), p.ncp);
  const dim3 tg(static_cast<unsigned>((nmax + 15) / 16), static_cast<unsigned>(p.mp / 16), 3);
  const double* xl = axes[p.dim - 1];
  const double* x1 = p.dim == 3 ? axes[1] : nullptr;
#define GSC_TABLES(T, NS)                                                                     \
  k_tables<T, NS><<<tg, 256, 0, s>>>(modes, Mi, p.dim, axes[0], x1, xl, p.na, p.nb, p.ncols, \
                                     p.ncp, P, Q, static_cast<T*>(bt))
  if (p.bf16) {
    if (p.ns == 1) GSC_TABLES(__bf16, 1);
    else if (p.ns == 2) GSC_TABLES(__bf16, 2);
    else GSC_TABLES(__bf16, 3);
  } else {
    if (p.ns == 1) GSC_TABLES(_Float16, 1);
    else if (p.ns == 2) GSC_TABLES(_Float16, 2);
    else GSC_TABLES(_Float16, 3);
  }
#undef GSC_TABLES
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const bool kr = p.dim == 3;
  const double sc = scale;
  const bool f32 = out_dtype == 1;
#define GSC_GEMM(T, NS)                                                                        \
  (f32 ? gemm_t<T, NS, float>(kr, accum, P, Q, p.na, p.nb, p.rows, p.mp / 16, bt, p.ncp,       \
                              p.ncols, out, p.ncols, sc, s)                                    \
       : gemm_t<T, NS, double>(kr, accum, P, Q, p.na, p.nb, p.rows, p.mp / 16, bt, p.ncp,      \
                               p.ncols, out, p.ncols, sc, s))
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
