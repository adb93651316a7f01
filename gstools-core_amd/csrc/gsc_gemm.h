// Factored summation on structured grids, contracted on the MFMA cores:
// device templates, instantiated per operand type in gemm_*.hip.
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
#pragma once

#include "gsc_device.h"
#include "gsc_factored.h"
#include "gsc_internal.h"

#include <hip/hip_runtime.h>

namespace gsc {
namespace gemm {

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
// sin and cos of 2 pi r for |r| <= 1/2, r exact in FP64: r = rh + rl with
// rh = fl32(r), and sin(2 pi (rh + rl)) = sin(2 pi rh) + 2 pi rl cos(2 pi rh)
// + O(rl^2) with |rl| <= 2^-26, so the FP32 result keeps the ~1 ulp of
// sincospif instead of inheriting the rounding of r.
__device__ __forceinline__ void sincos_2pi(double r, float& sn, float& cs) {
  const float rh = static_cast<float>(r);
  const float rl = static_cast<float>(r - static_cast<double>(rh));
  float sh, ch;
  sincospif(2.0f * rh, &sh, &ch);
  const float dl = 6.283185307179586f * rl;
  sn = __builtin_fmaf(dl, ch, sh);
  cs = __builtin_fmaf(-dl, sh, ch);
}

template <typename T>
struct alignas(4) Pair {
  T im, re;
};

// One entry (row a, mode 16 c + mm) of a table: which = 0 -> P (axis 0,
// amplitude and theta folded in), 1 -> Q (axis 1, 3-D), 2 -> B operand (last
// axis, split into NS parts; columns ncols <= a < ncp are zero).
template <typename T, int NS>
__device__ __forceinline__ void table_entry(const ModeRec* __restrict__ modes, int M, int dim,
                                            const double* __restrict__ x0,
                                            const double* __restrict__ x1,
                                            const double* __restrict__ xl, int64_t na, int64_t nb,
                                            int64_t ncols, int64_t ncp, float2* __restrict__ P,
                                            float2* __restrict__ Q, T* __restrict__ bt, int which,
                                            int64_t a, int c, int nch, int mm) {
  if (which == 1 && dim != 3) return;
  const int m = c * 16 + mm;
  const int64_t n = which == 0 ? na : (which == 1 ? nb : ncp);
  if (a >= n || c >= nch) return;
  const double* x = which == 0 ? x0 : (which == 1 ? x1 : xl);
  const int axis = which == 2 ? dim - 1 : which;
  float re = 0.0f, im = 0.0f;
  if (m < M && (which != 2 || a < ncols)) {
    const ModeRec md = modes[m];
    double t = md.kap[axis] * x[a];
    if (which == 0) t += md.th;
    float sn, cs;
    sincos_2pi(t - rint(t), sn, cs);  // t - rint(t) is exact: |r| <= 1/2 revolution
    const float amp = which == 0 ? md.amp[0] : 1.0f;
    re = amp * cs;
    im = amp * sn;
  }
  if (which < 2) {
    float2* tab = which == 0 ? P : Q;
    tab[(static_cast<int64_t>(c) * n + a) * 16 + mm] = make_float2(re, im);
    return;
  }
  T hi[NS], hr[NS];
  split<T, NS>(im, hi);
  split<T, NS>(re, hr);
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    Pair<T>* row = reinterpret_cast<Pair<T>*>(bt + ((static_cast<int64_t>(s) * nch + c) * ncp + a) * 32);
    row[mm] = Pair<T>{hi[s], hr[s]};
  }
}

template <typename T, int NS>
__global__ void __launch_bounds__(256)
k_tables(const ModeRec* __restrict__ modes, int M, int dim, const double* __restrict__ x0,
         const double* __restrict__ x1, const double* __restrict__ xl, int64_t na, int64_t nb,
         int64_t ncols, int64_t ncp, float2* __restrict__ P, float2* __restrict__ Q,
         T* __restrict__ bt) {
  table_entry<T, NS>(modes, M, dim, x0, x1, xl, na, nb, ncols, ncp, P, Q, bt, blockIdx.z,
                     static_cast<int64_t>(blockIdx.x) * 16 + (threadIdx.x >> 4), blockIdx.y,
                     gridDim.y, threadIdx.x & 15);
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
template <typename T, int NS>
struct GemmSmem {
  T As[2][NS][kTM][kLd];
  T Bs[2][NS][kTN][kLd];
};

// One 128 x 128 output tile (rows row0.., cols col0..) of one GEMM; used by
// the grid kernel below and by the persistent ensemble kernel.  Ends with a
// barrier after the last LDS read, so the caller may start another tile.
template <typename T, int NS, bool KR, int ACC>
__device__ __forceinline__ void gemm_tile(GemmSmem<T, NS>& sm, const float2* __restrict__ P,
                                          const float2* __restrict__ Q, int64_t na, int64_t nb,
                                          int64_t n_rows, int nch, const T* __restrict__ bt,
                                          int64_t ncp, int64_t ncols, void* __restrict__ out,
                                          bool out_f64, int64_t ldo, double scale, int64_t row0,
                                          int64_t col0) {
  using V8 = typename Vec8<T>::type;
  auto& As = sm.As;
  auto& Bs = sm.Bs;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;

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
        if (row < n_rows) {
          if (out_f64) static_cast<double*>(out)[row * ldo + col] = val * scale;
          else static_cast<float*>(out)[row * ldo + col] = static_cast<float>(val * scale);
        }
      }
    }
}

template <typename T, int NS, bool KR, int ACC>
__global__ void __launch_bounds__(256)
k_factored_gemm(const float2* __restrict__ P, const float2* __restrict__ Q, int64_t na,
                int64_t nb, int64_t n_rows, int nch, const T* __restrict__ bt, int64_t ncp,
                int64_t ncols, void* __restrict__ out, bool out_f64, int64_t ldo, double scale) {
  __shared__ __attribute__((aligned(16))) GemmSmem<T, NS> sm;
  gemm_tile<T, NS, KR, ACC>(sm, P, Q, na, nb, n_rows, nch, bt, ncp, ncols, out, out_f64, ldo,
                            scale, static_cast<int64_t>(blockIdx.x) * kTM,
                            static_cast<int64_t>(blockIdx.y) * kTN);
}

// Tables of all factored realisations: one block per task (realisation,
// table, group of 16 rows); the block walks every 16-mode chunk of its rows.
// task_start has 3 nfact + 1 entries: tasks of fact[i] / table w start at
// task_start[3 i + w].
template <typename T, int NS>
__global__ void __launch_bounds__(256)
k_ens_tables(const EnsReal* __restrict__ d, const int* __restrict__ fact,
             const int64_t* __restrict__ task_start, int nfact) {
  const int64_t task = blockIdx.x;
  const int k = find_item(task_start, 3 * nfact, task);
  const EnsReal& r = d[fact[k / 3]];
  const int which = k % 3;
  const int64_t a = (task - task_start[k]) * 16 + (threadIdx.x >> 4);
  const int nch = r.mp / 16;
  for (int c = 0; c < nch; ++c)
    table_entry<T, NS>(r.table, static_cast<int>(r.M), r.dim, r.ax0, r.ax1,
                       r.dim == 3 ? r.ax2 : r.ax1, r.na, r.nb, r.ncols, r.ncp, r.P, r.Q,
                       static_cast<T*>(r.bt), which, a, c, nch, threadIdx.x & 15);
}

// Persistent GEMM over the tiles of all factored realisations.  The tiles
// are split into kXcdGroups groups of whole realisations (balanced on the
// host); block b starts on group b % kXcdGroups -- workgroups are dispatched
// to the XCDs round-robin, so a realisation's table bands stay in one XCD's
// L2 -- and, once its group is exhausted, helps the others in turn.  Every
// group counter only grows, so every block leaves every group and exits.
template <typename T, int NS, bool KR, int ACC>
__global__ void __launch_bounds__(256)
k_ens_gemm(const EnsReal* __restrict__ d, const int* __restrict__ fact,
           const int64_t* __restrict__ tile_start, int nfact,
           const int64_t* __restrict__ group_start, int* __restrict__ counters, bool out_f64) {
  __shared__ __attribute__((aligned(16))) GemmSmem<T, NS> sm;
  __shared__ int64_t tile_s;
  const int home = blockIdx.x % kXcdGroups;
  for (int k = 0; k < kXcdGroups; ++k) {
    const int g = (home + k) % kXcdGroups;
    const int64_t g0 = group_start[g], g1 = group_start[g + 1];
    for (;;) {
      if (threadIdx.x == 0) tile_s = g0 + atomicAdd(&counters[g], 1);
      __syncthreads();
      const int64_t t = tile_s;
      __syncthreads();  // every wave has read tile_s before thread 0 fetches the next
      if (t >= g1) break;
      const int i = find_item(tile_start, nfact, t);
      const EnsReal& r = d[fact[i]];
      const int64_t local = t - tile_start[i];
      const int64_t rt = local / r.col_tiles, ct = local - rt * r.col_tiles;
      gemm_tile<T, NS, KR, ACC>(sm, r.P, r.Q, r.na, r.nb, r.rows, r.mp / 16,
                                static_cast<const T*>(r.bt), r.ncp, r.ncols, r.out, out_f64,
                                r.ncols, r.scale, rt * kTM, ct * kTN);
    }
  }
}

}  // namespace gemm

template <typename T, int NS>
hipError_t tables_launch(const TableArgs& a, dim3 grid, hipStream_t s) {
  gemm::k_tables<T, NS><<<grid, 256, 0, s>>>(a.modes, a.M, a.dim, a.x0, a.x1, a.xl, a.na, a.nb,
                                              a.ncols, a.ncp, a.P, a.Q, static_cast<T*>(a.bt));
  return hipGetLastError();
}

template <typename T, int NS>
hipError_t factored_gemm_launch(const GemmArgs& a, dim3 grid, hipStream_t s) {
  using namespace gemm;
  const T* bt = static_cast<const T*>(a.bt);
#define GSC_G(KR, ACC)                                                                          \
  k_factored_gemm<T, NS, KR, ACC><<<grid, 256, 0, s>>>(a.P, a.Q, a.na, a.nb, a.rows, a.nch, bt, \
                                                       a.ncp, a.ncols, a.out, a.out_f64, a.ldo,  \
                                                       a.scale)
  if (a.kr) {
    if (a.accum == ACC_F32X2) GSC_G(true, ACC_F32X2);
    else if (a.accum == ACC_F32) GSC_G(true, ACC_F32);
    else GSC_G(true, ACC_F64);
  } else {
    if (a.accum == ACC_F32X2) GSC_G(false, ACC_F32X2);
    else if (a.accum == ACC_F32) GSC_G(false, ACC_F32);
    else GSC_G(false, ACC_F64);
  }
#undef GSC_G
  return hipGetLastError();
}

template <typename T, int NS>
hipError_t ens_launch(const EnsLaunch& e, int accum, bool out_f64, hipStream_t s) {
  using namespace gemm;
  k_ens_tables<T, NS><<<static_cast<unsigned>(e.n_table_tasks), 256, 0, s>>>(
      e.d, e.fact, e.task_start, e.nfact);
#define GSC_E(KR, ACC)                                                                      \
  k_ens_gemm<T, NS, KR, ACC><<<e.gemm_blocks, 256, 0, s>>>(e.d, e.fact, e.tile_start, e.nfact, \
                                                           e.group_start, e.counters, out_f64)
  if (e.dim == 3) {
    if (accum == ACC_F32X2) GSC_E(true, ACC_F32X2);
    else if (accum == ACC_F32) GSC_E(true, ACC_F32);
    else GSC_E(true, ACC_F64);
  } else {
    if (accum == ACC_F32X2) GSC_E(false, ACC_F32X2);
    else if (accum == ACC_F32) GSC_E(false, ACC_F32);
    else GSC_E(false, ACC_F64);
  }
#undef GSC_E
  return hipGetLastError();
}

}  // namespace gsc
