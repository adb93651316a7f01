// Factored summation on structured grids, contracted on the MFMA cores:
// device templates, instantiated per operand format in gemm_*.hip.
//
// On a grid x = (x_a, y_b, z_c) the per-mode phase separates, and with the
// amplitude-phase form used throughout this library
//   z1 cos(phi) + z2 sin(phi) = A sin(phi + theta),   A = hypot(z1, z2),
//   theta = atan2(z1, z2),
// the field is the imaginary part of a product of per-axis tables:
//   u(a, b, c) = Im sum_m S(a, m) Q(b, m) C(c, m)
//   S(a, m) = A_m e^{i(theta_m + k_xm x_a)},  Q(b, m) = e^{i k_ym y_b},
//   C(c, m) = e^{i k_zm z_c}.
// For a fixed a this is a real GEMM over K = 2M, rows b, columns c:
//   u(a, b, c) = sum_k A(b, k) B_a(k, c),  A = Q, B_a = S(a) C,
// with K interleaved as k = 2m -> (Re A, Im B), k = 2m + 1 -> (Im A, Re B),
// so that Re A Im B + Im A Re B = Im(A B).  In 2-D the A operand is the
// axis-0 table (amplitude and theta folded in) and B is the axis-1 table.
//
// Tables are evaluated with the phase in FP64 revolutions, reduced exactly
// (t - rint(t)), rounded to FP32 and passed to the FP32 sine and cosine
// instructions (argument in revolutions).  MFMA operands are split into NS
// parts of T (fp16 or bf16): x = sum_s x_s, x_0 = T(x),
// x_{s+1} = T(x - sum_{t<=s} x_t) (each difference exact in FP32), and the
// products with i + j < NS are accumulated in FP32 on the matrix cores:
// NS = 1 -> 1 product, NS = 2 -> 3 products (x3), NS = 3 -> 6 products (x6).
//
// Kernel shape: 256 threads = 4 waves stacked in rows, block tile
// 256 rows x 32 JN columns; a wave owns 64 rows x 32 JN columns
// (2 x JN v_mfma_f32_32x32x16 blocks).  The A operand is split once, when the
// table is built, and every wave loads its fragments straight from the
// table (global memory / L2) into registers in the MFMA layout: no LDS and
// no arithmetic for A.  The B operand (S(a) C in 3-D, C in 2-D) is built per
// block and 16-mode chunk by the four waves (complex product, split) into a
// double-buffered LDS tile, one barrier per chunk.  Per chunk a wave issues
// 2 JN NS(NS+1)/2 MFMAs per 16 k and 4 x (JN > 1 ? 32 JN / 64 : 1) complex
// products for B.
//
// Table layouts, chosen so that every load instruction of the GEMM reads
// whole cache lines (mode m = 16 c + mm, pair = (Re, Im)):
//   S (float2):        S[(c * na + a) * 16 + mm]
//   A (packed pairs):  A[((((s * nch + c) * 2 + ks) * 2 + h) * arows + row) * 4 + e],
//                      mm = 8 ks + 4 h + e: the 4 pairs one MFMA lane holds
//   C (float2):        C[((m / 2) * ncp + col) * 2 + m % 2]: mode pairs of one
//                      column contiguous, columns of one mode pair contiguous
// with the mode count padded to mp = 32 ceil(M / 32) (nch = mp / 16 chunks),
// padded modes, rows and columns zero.
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
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int kLdB = 40;    // LDS row of the B tile: 32 k + 8 pad (T elements, 80 bytes)
constexpr int kFlush = 16;  // chunks (256 modes) per FP32 accumulation interval
// Accumulation (ACC template argument).  The MFMA accumulates in FP32; over
// K = 2M terms its rounding grows with the partial sum (~8e-6 sigma at
// M = 4000).  Every kFlush chunks the FP32 accumulators are added into a
// second set and restarted, so FP32 rounding only ever sees 256 modes:
//   ACC_F64   second set in FP64  (1 column block per wave)
//   ACC_F32X2 second set in FP32  (~1e-6 sigma at M = 4000)
//   ACC_F32   no second set       (error grows with M)
enum { ACC_F64 = 0, ACC_F32X2 = 1, ACC_F32 = 2 };
template <int ACC> struct AccT { using type = double; };
template <> struct AccT<ACC_F32X2> { using type = float; };
template <> struct AccT<ACC_F32> { using type = float; };

// Column blocks (of 32) per wave: the accumulators, the second set and the
// operand fragments must fit one wave's register file.
template <int NS, int ACC>
constexpr int cols_blocks() {
  return ACC == ACC_F64 ? 1 : (NS == 3 ? 2 : 4);
}

// Waves per SIMD the GEMM kernels are built for.  RB = 2 (256-row blocks,
// 64 rows per wave): the wide tiles need more than 256 registers, one wave
// per SIMD; the one-column-block FP64 flush variant fits 256, so a second
// block hides the first one's barrier and load waits.  RB = 1 (128-row
// blocks, 32 rows per wave): half the accumulators, two blocks per CU, so one
// block's barrier and fragment waits are covered by the other's MFMAs.
template <int ACC, int RB>
constexpr int gemm_waves() {
  return RB == 1 ? (ACC == ACC_F64 ? 4 : 2) : (ACC == ACC_F64 ? 2 : 1);
}

static_assert(32 * cols_blocks<2, ACC_F32X2>() == gemm_tile_cols(2, ACC_F32X2) &&
                  32 * cols_blocks<3, ACC_F32>() == gemm_tile_cols(3, ACC_F32) &&
                  32 * cols_blocks<1, ACC_F64>() == gemm_tile_cols(1, ACC_F64),
              "host and device tile widths agree");

template <typename T> struct Fmt;
template <> struct Fmt<_Float16> {
  using v8 = half8_t;
  using v2 = _Float16 __attribute__((ext_vector_type(2)));
};
template <> struct Fmt<__bf16> {
  using v8 = bf16x8_t;
  using v2 = __bf16 __attribute__((ext_vector_type(2)));
};

__device__ __forceinline__ f32x16_t mfma(half8_t a, half8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x16_t mfma(bf16x8_t a, bf16x8_t b, f32x16_t c) {
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// (x, y) -> NS packed pairs of T with (x, y) = sum_s pair_s, round to nearest
// even; each residual x - sum_{t<=s} x_t is exact in FP32.
template <typename T, int NS>
__device__ __forceinline__ void split_pair(f2v v, uint32_t (&o)[NS]) {
  using V2 = typename Fmt<T>::v2;
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const V2 h = __builtin_convertvector(v, V2);
    o[s] = __builtin_bit_cast(uint32_t, h);
    if (s + 1 < NS) v = v - __builtin_convertvector(h, f2v);
  }
}

// Global-memory view of a pointer.  The ensemble passes its tables through
// descriptors in device memory; without the address space the compiler
// would use flat accesses, whose waits also cover the LDS counter.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* gmem(T* p) {
  return (__attribute__((address_space(1))) T*)(p);  // address-space conversion
}

// Four table entries from their modes' data: v[e] = amp e^{2 pi i (kx x +
// th)} of mode m0 + e (amp = 1, th = 0 unless with_amp), zero for modes past
// M and for padding rows.  FP64 phase in revolutions, exact reduction to
// [-1/2, 1/2], then the FP32 cosine and sine instructions.
__device__ __forceinline__ void table_eval(const double (&kx)[4], const double (&th)[4],
                                           const float (&am)[4], double x, bool with_amp,
                                           bool valid, int m0, int M, f2v (&v)[4]) {
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    double t = kx[e] * x;
    if (with_amp) t += th[e];
    const float r = static_cast<float>(t - rint(t));
    const float a = with_amp ? am[e] : 1.0f;
    const f2v w = f2v{a * __builtin_amdgcn_cosf(r), a * __builtin_amdgcn_sinf(r)};
    v[e] = (valid && m0 + e < M) ? w : f2v{0.0f, 0.0f};
  }
}

// Stores of four entries (modes 16 c + 4 g .. + 3) of row `row` of the
// role's table: role 0 -> S (3-D axis 0), 1 -> A (split), 2 -> C.  Each
// role's store is one 16-byte vector per part, consecutive rows contiguous
// for A and C.
template <typename T, int NS>
__device__ __forceinline__ void table_store(const f2v (&v)[4], int role, int64_t row, int64_t cc,
                                            int g, int nch, int64_t na, int64_t arows,
                                            int64_t ncpt, float2* S, uint32_t* A, float2* C) {
  if (role == 0) {
    auto* d = gmem(reinterpret_cast<f4v*>(S + (cc * na + row) * 16 + 4 * g));
    d[0] = f4v{v[0].x, v[0].y, v[1].x, v[1].y};
    d[1] = f4v{v[2].x, v[2].y, v[3].x, v[3].y};
  } else if (role == 1) {
    uint32_t o[4][NS];
#pragma unroll
    for (int e = 0; e < 4; ++e) split_pair<T, NS>(v[e], o[e]);
#pragma unroll
    for (int s = 0; s < NS; ++s)
      *gmem(reinterpret_cast<u32x4*>(A + ((((s * nch + cc) * 4) + g) * arows + row) * 4)) =
          u32x4{o[0][s], o[1][s], o[2][s], o[3][s]};
  } else {
    auto* d = gmem(reinterpret_cast<f4v*>(C));
    d[(8 * cc + 2 * g) * ncpt + row] = f4v{v[0].x, v[0].y, v[1].x, v[1].y};
    d[(8 * cc + 2 * g + 1) * ncpt + row] = f4v{v[2].x, v[2].y, v[3].x, v[3].y};
  }
}

// Four table entries: role 0 -> S (3-D axis 0), 1 -> A (split), 2 -> C; row
// `row` of that table, modes 16 c + 4 g .. + 3 (g < 4).  The four modes'
// data are loaded together (index clamped to M - 1), then evaluated and
// stored.  C has ncomp blocks of ncp rows; with ncomp > 1 block i holds the
// entries times the projector p_i(k) = delta_i0 - k_i k_0 / |k|^2 (FP64 from
// the mode's kap = k / 2 pi; k = 0 gives NaN, as field.rs:138).
template <typename T, int NS>
__device__ __forceinline__ void table_quad(const ModeRec* __restrict__ modes, int M, int dim,
                                           int nch, const double* __restrict__ xs,
                                           const double* __restrict__ xa,
                                           const double* __restrict__ xc, int64_t na,
                                           int64_t nrow, int64_t arows, int64_t ncols,
                                           int64_t ncp, float2* __restrict__ S,
                                           uint32_t* __restrict__ A, float2* __restrict__ C,
                                           int role, int64_t row, int c, int g, int trans,
                                           int ncomp, const ProdGeom& pg) {
  if (c >= nch || M <= 0) return;
  const int m0 = c * 16 + 4 * g;
  int axis, comp = 0;
  bool amp, valid;
  double x = 0.0, x2 = 0.0;
  const bool prod = pg.prod != 0;
  if (role == 0) {  // 3-D axis 0, or the product form's tile corners
    if ((dim != 3 && !prod) || row >= na) return;
    axis = prod ? (trans ? 1 : 0) : 0;
    amp = true;
    valid = true;
    if (prod) {
      const int64_t rh = row / pg.col_tiles, ch = row - rh * pg.col_tiles;
      x = pg.org0 + static_cast<double>(rh * pg.brows) * pg.step0;
      x2 = pg.org1 + static_cast<double>(ch * pg.tcols) * pg.step1;
    } else {
      x = gmem(xs)[row];
    }
  } else if (role == 1) {  // a transposed 2-D plan takes its rows from axis 1
    if (row >= arows) return;
    axis = trans ? 1 : dim - 2;
    amp = dim == 2 && !prod;
    valid = prod || row < nrow;
    if (prod) x = static_cast<double>(row) * pg.step0;
    else if (valid) x = gmem(xa)[row];
  } else {
    if (row >= ncp * ncomp) return;
    comp = static_cast<int>(row / ncp);
    const int64_t col = row - comp * ncp;
    axis = trans ? 0 : dim - 1;
    amp = false;
    valid = prod || col < ncols;
    if (prod) x = static_cast<double>(col) * pg.step1;
    else if (valid) x = gmem(xc)[col];
  }
  const int axis2 = trans ? 0 : 1;  // product-form S: the column axis
  const __attribute__((address_space(1))) ModeRec* gm = gmem(modes);
  double kx[4], th[4];
  float am[4], pr[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const int m = min(m0 + e, M - 1);
    kx[e] = gm[m].kap[axis];
    // the product form's S: the column axis' part joins the offset
    th[e] = prod && role == 0 ? __builtin_fma(gm[m].kap[axis2], x2, gm[m].th) : gm[m].th;
    am[e] = gm[m].amp[0];
    pr[e] = 1.0f;
    if (ncomp > 1 && role == 2) {
      const double k0 = gm[m].kap[0], k1 = gm[m].kap[1], k2 = gm[m].kap[2];
      const double kc = comp == 0 ? k0 : (comp == 1 ? k1 : k2);
      pr[e] = static_cast<float>((comp == 0 ? 1.0 : 0.0) - (kc * k0) / (k0 * k0 + k1 * k1 + k2 * k2));
    }
  }
  f2v v[4];
  table_eval(kx, th, am, x, amp, valid, m0, M, v);
  if (ncomp > 1 && role == 2) {
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = v[e] * pr[e];
  }
  table_store<T, NS>(v, role, row, c, g, nch, na, arows, ncp * ncomp, S, A, C);
}

// All tables of one call in one launch.  block = 64 rows x 16 modes of one
// chunk (thread: row t % 64, modes 4 (t / 64) .. + 3):
// grid (ceil(max rows / 64), nch, 3), blockIdx.z = role.
template <typename T, int NS>
__global__ void __launch_bounds__(256)
k_tables(const ModeRec* __restrict__ modes, int M, int dim, int nch,
         const double* __restrict__ xs, const double* __restrict__ xa,
         const double* __restrict__ xc, int64_t na, int64_t nrow, int64_t arows, int64_t ncols,
         int64_t ncp, float2* __restrict__ S, uint32_t* __restrict__ A, float2* __restrict__ C,
         int trans, int ncomp, ProdGeom pg) {
  table_quad<T, NS>(modes, M, dim, nch, xs, xa, xc, na, nrow, arows, ncols, ncp, S, A, C,
                    blockIdx.z, static_cast<int64_t>(blockIdx.x) * 64 + (threadIdx.x & 63),
                    blockIdx.y, threadIdx.x >> 6, trans, ncomp, pg);
}

// ---------------------------------------------------------------------------
// The GEMM.  One block tile: rows arow0 .. arow0 + 255 of the A table
// (output rows out_row0 + r for r < nvalid), columns col0 .. col0 + 32 JN - 1
// (stored for col < ncols):  out[(out_row0 + r) * ldo + col] = scale * u.
//
// Software pipeline (128-row blocks), one barrier per chunk: in the step for
// chunk c a wave
//   1. requests chunk c + 1's A fragments into the other register set and
//      reads chunk c's ks = 1 B fragments from LDS buffer c % 2,
//   2. issues chunk c's ks = 0 MFMAs (fragments read in the previous step)
//      and, between them, builds chunk c + 1's B tile into the other buffer
//      from C (and S) values loaded one step earlier,
//   3. waits at the barrier, requests the C (and S) values of chunk c + 2
//      and reads chunk c + 1's ks = 0 B fragments,
//   4. issues chunk c's ks = 1 MFMAs.
// 256-row blocks (RB = 2) build the B tile two chunks ahead into four slots
// and synchronise once per two chunks; their C (and S) values are loaded a
// step and a half ahead, into two register sets.
// ---------------------------------------------------------------------------
template <int RB>
constexpr int b_slots() {
  return RB == 2 ? 4 : 2;
}
template <typename T, int NS, int JN, int SLOTS>
struct GemmSmem {
  T B[SLOTS][NS][32 * JN][kLdB];
};

struct TileJob {
  const uint32_t* A;
  int64_t arows;
  int nch;
  const float2* S;  // 3-D: S table and the block's index a
  int64_t na, a;
  const float2* C;
  int64_t ncp;      // row length of the C table (all components)
  int64_t arow0;
  int64_t ccol0;    // first C-table column of the tile
  int64_t col0;     // its output column (within the tile's field component)
  void* out;        // the component's output
  bool out_f64;
  int64_t ldo, out_row0, nvalid, ncols;
  double scale;
  int64_t ldc;  // column stride of the output (1, or the row count of a transposed store)
};

// Output of one wave's 32 RB x 32 JN tile (rows past nvalid and columns past
// ncols are not stored).
template <int ACC, int JN, int RB, bool TR, typename OutT, typename A2>
__device__ __forceinline__ void gemm_store(OutT* __restrict__ out, const TileJob& jb,
                                           const f32x16_t (&acc)[RB][JN], const A2& acc2,
                                           int wave, int hl, int lr) {
  const int64_t ldo = jb.ldo;
  constexpr int kRows = 128 * RB;
  const int nvalid = static_cast<int>(jb.nvalid < kRows ? jb.nvalid : kRows);
  const float scale32 = static_cast<float>(jb.scale);
#pragma unroll
  for (int i = 0; i < RB; ++i) {
    const int r0 = 32 * RB * wave + 32 * i + 4 * hl;
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      const int64_t col = jb.col0 + 32 * j + lr;
      if (col >= jb.ncols) continue;
      __attribute__((address_space(1))) OutT* const p =
          gmem(out) + (jb.out_row0 + r0) * ldo + (TR ? col * jb.ldc : col);
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        const int rr = 8 * (v >> 2) + (v & 3);
        if (r0 + rr >= nvalid) continue;
        OutT val;
        if constexpr (ACC == ACC_F64) {
          val = static_cast<OutT>(acc2[i][j][v] * jb.scale);
        } else {
          const float a = ACC == ACC_F32 ? acc[i][j][v] : static_cast<float>(acc2[i][j][v]);
          if constexpr (sizeof(OutT) == 8) val = static_cast<double>(a) * jb.scale;
          else val = a * scale32;
        }
        // a transposed store writes 4-byte pieces of many lines: ordinary
        // stores, so that L2 assembles each line before it is written
        if constexpr (TR) p[rr * ldo] = val;
        else __builtin_nontemporal_store(val, p + rr * ldo);
      }
    }
  }
}

template <typename T, int NS, bool KR, int ACC, int JN, int RB, bool TR = false>
__device__ __forceinline__ void gemm_tile(GemmSmem<T, NS, JN, b_slots<RB>()>& sm,
                                          const TileJob& jb) {
  using V8 = typename Fmt<T>::v8;
  using A2 = typename AccT<ACC>::type;
  constexpr int GC = JN == 4 ? 2 : 1;  // B columns built per lane
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int hl = lane >> 5;  // k half of an MFMA fragment
  const int lr = lane & 31;  // row / column within a 32 x 32 block
  const int nch = jb.nch;

  // A fragments: row arow0 + 32 RB wave + 32 i + lr, packed pairs 8 ks + 4 hl .. + 3
  const __attribute__((address_space(1))) uint32_t* abase =
      gmem(jb.A) + (hl * jb.arows + jb.arow0 + 32 * RB * wave + lr) * 4;
  const int64_t a_ks = 2 * jb.arows * 4;
  const int64_t a_chunk = 2 * a_ks;
  const int64_t a_part = a_chunk * nch;
  // B builder: this wave builds modes 4 wave .. + 3 (k = 8 wave .. + 7) of
  // each chunk for columns lane + 64 g (JN = 4), lane (JN = 2) or lr (JN = 1)
  const bool gen_on = JN > 1 || lane < 32;
  const int gcol = JN == 1 ? lr : lane;
  // mode pair 8 c + 2 wave + h of column col0 + gcol + 64 g: float4 index
  // (8 c + 2 wave + h) ncp + col0 + gcol + 64 g
  const __attribute__((address_space(1))) f4v* cbase =
      gmem(reinterpret_cast<const f4v*>(jb.C)) + 2 * wave * jb.ncp + jb.ccol0 + gcol;
  const int64_t c_pair = jb.ncp;
  const int64_t c_chunk = 8 * jb.ncp;
  // S(a, m) of one chunk: lane l < 16 loads mode l with a vector load (a
  // scalar load would make every LDS wait in the loop wait for it too); the
  // wave takes its four modes with readlane
  const __attribute__((address_space(1))) f2v* sbase =
      KR ? gmem(reinterpret_cast<const f2v*>(jb.S)) + jb.a * 16 + (lane & 15) : nullptr;
  const int64_t s_chunk = jb.na * 16;

  // C values (and, in 3-D, the wave's four S values) of one chunk: with
  // RB = 1 loaded a step before they are used, with RB = 2 into two register
  // sets (chunk parity) a step and a half before (an L2 hit outlasts half a
  // step of MFMAs)
  constexpr int CB = RB == 2 ? 2 : 1;
  f4v cr[CB][GC][2];
  float2 sl[CB];
  auto gload = [&](int c, const int b) {
#pragma unroll
    for (int g = 0; g < GC; ++g) {
      const __attribute__((address_space(1))) f4v* p = cbase + c * c_chunk + 64 * g;
      cr[b][g][0] = p[0];
      cr[b][g][1] = p[c_pair];
    }
    if constexpr (KR) {
      // one 8-byte load, then plain float members for the readlanes below
      // (bit-casting elements of a vector-typed sl lost its .y half)
      const f2v t = sbase[c * s_chunk];
      sl[b].x = t[0];
      sl[b].y = t[1];
    }
  };
  auto gstore = [&](const int buf, const int b) {
    float2 sv[4];
    if constexpr (KR) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        sv[q].x = __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                __builtin_bit_cast(int, sl[b].x), 4 * wave + q));
        sv[q].y = __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                __builtin_bit_cast(int, sl[b].y), 4 * wave + q));
      }
    }
#pragma unroll
    for (int g = 0; g < GC; ++g) {
      uint32_t w[NS][4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const f4v v = cr[b][g][q >> 1];
        float re = (q & 1) ? v.z : v.x, im = (q & 1) ? v.w : v.y;
        if constexpr (KR) {
          const float r2 = sv[q].x * re - sv[q].y * im;
          im = sv[q].x * im + sv[q].y * re;
          re = r2;
        }
        uint32_t o[NS];
        split_pair<T, NS>(f2v{im, re}, o);  // B pair (Im, Re)
#pragma unroll
        for (int s = 0; s < NS; ++s) w[s][q] = o[s];
      }
      if (gen_on) {
#pragma unroll
        for (int s = 0; s < NS; ++s)
          *reinterpret_cast<u32x4*>(&sm.B[buf][s][gcol + 64 * g][8 * wave]) =
              u32x4{w[s][0], w[s][1], w[s][2], w[s][3]};
      }
    }
  };

  V8 af[2][2][NS][RB];  // [chunk parity][ks][part][row block]
  auto aload = [&](const int set, int c) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int s = 0; s < NS; ++s)
#pragma unroll
        for (int i = 0; i < RB; ++i)
          af[set][ks][s][i] = *reinterpret_cast<const __attribute__((address_space(1))) V8*>(
              abase + s * a_part + c * a_chunk + ks * a_ks + 128 * i);
  };
  // B fragments: both k halves of a chunk held at once with RB = 2 (read a
  // half-step ahead); with RB = 1 one half at a time, read just before its
  // MFMAs (the registers go to the second wave on the SIMD, which covers the
  // LDS latency)
  constexpr int KB = RB == 2 ? 2 : 1;
  V8 bf[KB][NS][JN];  // [ks][part][column block]
  auto bread = [&](const int ks, const int buf) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
#pragma unroll
      for (int j = 0; j < JN; ++j)
        bf[ks % KB][s][j] =
            *reinterpret_cast<const V8*>(&sm.B[buf][s][32 * j + lr][16 * ks + 8 * hl]);
  };

  f32x16_t acc[RB][JN];
  A2 acc2[RB][JN][ACC == ACC_F32 ? 1 : 16];
#pragma unroll
  for (int i = 0; i < RB; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) {
#pragma unroll
      for (int v = 0; v < 16; ++v) acc[i][j][v] = 0.0f;
#pragma unroll
      for (int v = 0; v < (ACC == ACC_F32 ? 1 : 16); ++v) acc2[i][j][v] = 0;
    }

  auto mfma_ks = [&](const int set, const int ks) {
#pragma unroll
    for (int sa = 0; sa < NS; ++sa)
#pragma unroll
      for (int sb = 0; sb < NS - sa; ++sb)
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < JN; ++j)
            acc[i][j] = mfma(af[set][ks][sa][i], bf[ks % KB][sb][j], acc[i][j]);
  };

  // One chunk c (register set and LDS buffer c % 2).  The barrier sits
  // between building chunk c + 1's B tile and reading its first fragments;
  // chunk c's fragments were all read before it, so the next build into
  // this buffer (after the following barrier) cannot overtake a read.
  // A wave issues in order, so the memory and ALU work of each half is
  // spread between its MFMAs (scheduling groups): a burst of loads would
  // stall the MFMAs behind it.
  constexpr int kM = RB * JN * NS * (NS + 1) / 2;  // MFMAs per half
  constexpr int kA = 2 * RB * NS;                   // A fragment loads per chunk
  constexpr int kB = NS * JN;                      // B fragment reads per half
  constexpr int kW = NS * GC;                      // B tile writes per chunk
  constexpr int kG = 2 * GC + (KR ? 1 : 0);        // C (and S) loads per chunk
  // RB = 1: one barrier per chunk after both halves, all of chunk c's
  // fragments read and chunk c + 1's tile written by then
  auto step1 = [&](int c, const int set) {
    const int cn = c + 1 < nch ? c + 1 : c;
    aload(set ^ 1, cn);
    bread(0, set);
    mfma_ks(set, 0);
    gstore(set ^ 1, 0);
    bread(1, set);
    mfma_ks(set, 1);
    __syncthreads();
    gload(c + 2 < nch ? c + 2 : cn, 0);
  };
  // RB = 2: chunk c's fragments come from slot c % 4; the step builds chunk
  // c + 2 into slot (c + 2) % 4, and odd steps end with the barrier that
  // publishes the two chunks built since the last one.  Slot (c + 2) % 4 held
  // chunk c - 2, whose last reads precede that barrier (see the order below:
  // ks = 1 fragments at the start of a step, the next chunk's ks = 0
  // fragments at its end).
  auto step = [&](int c, const int set) {
    if constexpr (RB == 1) {
      step1(c, set);
      return;
    }
    const int cn = c + 1 < nch ? c + 1 : c;
    aload(set ^ 1, cn);
    bread(1, c & 3);
    mfma_ks(set, 0);
    // past the last chunk this rebuilds the last one into a slot that no
    // wave reads again
    gstore((c + 2) & 3, set);
#pragma unroll
    for (int k = 0; k < kM; ++k) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      if (k < kA) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);            // VMEM read
      else if (k < kA + kB) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
      __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);                        // VALU
      if (k >= kM - kW) __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);      // DS write
    }
    __builtin_amdgcn_sched_barrier(0);
    if (set == 1) __syncthreads();  // set = c % 2, fixed by the unrolled pair below
    gload(c + 4 < nch ? c + 4 : nch - 1, set);
    bread(0, (c + 1) & 3);
    mfma_ks(set, 1);
#pragma unroll
    for (int k = 0; k < kM; ++k) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      if (k < kG) __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);            // VMEM read
      else if (k < kG + kB) __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);  // DS read
      __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);                        // VALU
    }
    __builtin_amdgcn_sched_barrier(0);
  };

  gload(0, 0);
  gstore(0, 0);
  if constexpr (RB == 2) {
    gload(1, 1);  // nch is even, at least 2
    gstore(1, 1);
    gload(nch > 2 ? 2 : nch - 1, 0);
    gload(nch > 3 ? 3 : nch - 1, 1);
  } else {
    gload(nch > 1 ? 1 : 0, 0);
  }
  aload(0, 0);
  __syncthreads();
  if constexpr (RB == 2) bread(0, 0);
  // nch is even (modes padded to 32): the unrolled pair fixes set and buffer
  for (int c = 0; c < nch; c += 2) {
    step(c, 0);
    step(c + 1, 1);
    if constexpr (ACC != ACC_F32) {
      if ((c + 2) % kFlush == 0 || c + 2 == nch) {
#pragma unroll
        for (int i = 0; i < RB; ++i)
#pragma unroll
          for (int j = 0; j < JN; ++j)
#pragma unroll
            for (int v = 0; v < 16; ++v) {
              acc2[i][j][v] += static_cast<A2>(acc[i][j][v]);
              acc[i][j][v] = 0.0f;
            }
      }
    }
  }
  __syncthreads();  // the caller may reuse the LDS tile for its next tile

  // ---- epilogue: D[8 (v/4) + 4 (lane/32) + v%4][lane%32] of each 32 x 32 block.
  // One pointer per column block; the row steps 8 (v/4) + v%4 are uniform
  // multiples of ldo.  FP32 output is scaled in FP32 (the sums are FP32
  // already unless ACC_F64).
  if (jb.out_f64)
    gemm_store<ACC, JN, RB, TR>(static_cast<double*>(jb.out), jb, acc, acc2, wave, hl, lr);
  else
    gemm_store<ACC, JN, RB, TR>(static_cast<float*>(jb.out), jb, acc, acc2, wave, hl, lr);
}

// Tile job of row tile rt, column tile ct of one contraction.  Column tiles
// run component-major: tiles_per_comp tiles of 32 JN output columns per field
// component, component i's C-table columns starting at i ncp and its output
// at out + i comp_stride.
template <bool KR, int JN, int RB>
__device__ __forceinline__ TileJob make_job(const uint32_t* A, int64_t arows, int nch,
                                            const float2* S, int64_t na, const float2* C,
                                            int64_t ncpt, int64_t nrow, int64_t nb_out,
                                            void* out, bool out_f64, int64_t ldo,
                                            int64_t ncols, double scale, int64_t rt,
                                            int64_t ct, int64_t ldc, int64_t tiles_per_comp,
                                            int64_t ncp, int64_t comp_stride) {
  const int64_t nbt = arows / (128 * RB);
  const int64_t a = KR ? rt / nbt : 0;
  const int64_t arow0 = (rt - a * nbt) * (128 * RB);
  const int64_t comp = ct / tiles_per_comp;
  TileJob jb;
  jb.A = A;
  jb.arows = arows;
  jb.nch = nch;
  jb.S = S;
  jb.na = na;
  jb.a = a;
  jb.C = C;
  jb.ncp = ncpt;
  jb.arow0 = arow0;
  jb.col0 = (ct - comp * tiles_per_comp) * 32 * JN;
  jb.ccol0 = comp * ncp + jb.col0;
  jb.out = static_cast<char*>(out) + comp * comp_stride * (out_f64 ? 8 : 4);
  jb.out_f64 = out_f64;
  jb.ldo = ldo;
  jb.out_row0 = KR ? a * nb_out + arow0 : arow0;
  jb.nvalid = nrow - arow0;
  jb.ncols = ncols;
  jb.scale = scale;
  jb.ldc = ldc;
  return jb;
}

template <typename T, int NS, bool KR, int ACC, bool TR, int RB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(gemm_waves<ACC, RB>(), gemm_waves<ACC, RB>())))
k_factored_gemm(GemmArgs g) {
  constexpr int JN = cols_blocks<NS, ACC>();
  __shared__ __attribute__((aligned(16))) GemmSmem<T, NS, JN, b_slots<RB>()> sm;
  const int64_t tpc = (g.ncols + 32 * JN - 1) / (32 * JN);
  TileJob jb = make_job<KR, JN, RB>(g.A, g.arows, g.nch, g.S, g.na, g.C, g.ncpt, g.nrow,
                                    g.nb_out, g.out, g.out_f64, g.ldo, g.ncols, g.scale,
                                    blockIdx.x, blockIdx.y, g.ldc, tpc, g.ncp, g.comp_stride);
  if (KR && g.prod) {
    // product form: S row = tile (blockIdx.x, blockIdx.y), A and C tables one tile each
    jb.a = static_cast<int64_t>(blockIdx.x) * g.col_tiles + blockIdx.y;
    jb.arow0 = 0;
    jb.ccol0 = 0;
    jb.col0 = static_cast<int64_t>(blockIdx.y) * (32 * JN);
    jb.out_row0 = static_cast<int64_t>(blockIdx.x) * g.arows;
    jb.nvalid = g.nrow - jb.out_row0;
  }
  gemm_tile<T, NS, KR, ACC, JN, RB, TR>(sm, jb);
}

// Tables of all factored realisations: one block per task (realisation,
// role, group of 64 rows); the block walks every chunk of its rows.
// task_start has 3 nfact + 1 entries: tasks of fact[i] / role w start at
// task_start[3 i + w].  The descriptor and the row coordinate are read once
// and the modes' data for the role's axis are staged in LDS, kTabStage modes
// at a time, so the chunk loop's only global accesses are its stores (the
// vector memory counter counts stores too: a global load in the loop would
// wait for the stores before it).  Product-form realisations (EnsReal.prod)
// take their coordinates from origin and spacing: S rows are tiles (phase
// th + k_r x_r + k_c x_c at the tile's corner), A rows and C columns the
// offsets within a tile.
constexpr int kTabStage = 512;
template <typename T, int NS>
__global__ void __launch_bounds__(256)
k_ens_tables(const EnsReal* __restrict__ d, const int* __restrict__ fact,
             const int64_t* __restrict__ task_start, int nfact) {
  __shared__ double2 s_kt[kTabStage];  // (slope along the role's axis, phase offset)
  __shared__ double s_k2[kTabStage];   // product-form S: slope along the column axis
  __shared__ float s_am[kTabStage];    // amplitude
  const int64_t task = blockIdx.x;
  const int k = find_item(task_start, 3 * nfact, task);
  const EnsReal r = d[fact[k / 3]];
  const int M = static_cast<int>(r.M);
  if (M <= 0) return;  // the whole block: M is the realisation's
  const int role = k % 3;
  const int nch = r.mp / 16;
  const int g = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
  const int64_t row = (task - task_start[k]) * 64 + (threadIdx.x & 63);
  const bool prod = r.prod != 0;
  int axis;
  bool amp;
  const double* xp;
  int64_t nvalid, nrows;
  if (role == 0) {  // 3-D axis 0, or the product form's tile corners
    axis = prod ? (r.trans ? 1 : 0) : 0;
    amp = true;
    xp = r.ax0;
    nvalid = r.na;
    nrows = r.na;
  } else if (role == 1) {  // a transposed 2-D realisation takes its rows from y
    axis = r.trans ? 1 : r.dim - 2;
    amp = r.dim == 2 && !prod;
    xp = r.dim == 3 ? r.ax1 : (r.trans ? r.ax1 : r.ax0);
    nvalid = prod ? r.arows : r.nrow;
    nrows = r.arows;
  } else {
    axis = r.trans ? 0 : r.dim - 1;
    amp = false;
    xp = r.dim == 3 ? r.ax2 : (r.trans ? r.ax0 : r.ax1);
    nvalid = prod ? r.ncp : r.ncols;
    nrows = r.ncp;
  }
  const int axis2 = r.trans ? 0 : 1;  // product-form S: the column axis
  const bool active = row < nrows;  // rows of the last S group past na store nothing
  const bool valid = row < nvalid;  // padding rows of A and C hold zeros
  double x = 0.0, x2 = 0.0;
  if (valid) {
    if (!prod) {
      x = gmem(xp)[row];
    } else if (role == 0) {
      const int64_t rh = row / r.col_tiles, ch = row - rh * r.col_tiles;
      x = r.org[0] + static_cast<double>(rh * r.arows) * r.step[0];
      x2 = r.org[1] + static_cast<double>(ch * r.tcols) * r.step[1];
    } else {
      x = static_cast<double>(row) * r.step[role - 1];
    }
  }
  const __attribute__((address_space(1))) ModeRec* gm = gmem(r.table);
  float2* S = r.S;
  uint32_t* A = r.A;
  float2* C = r.C;
  for (int b0 = 0; b0 < 16 * nch; b0 += kTabStage) {
    __syncthreads();  // the previous stage is read
    for (int i = threadIdx.x; i < kTabStage; i += 256) {
      const int m = min(b0 + i, M - 1);
      s_kt[i] = double2{gm[m].kap[axis], gm[m].th};
      s_k2[i] = gm[m].kap[axis2];
      s_am[i] = gm[m].amp[0];
    }
    __syncthreads();
    const int c1 = min(nch, (b0 + kTabStage) / 16);
    for (int c = b0 / 16; c < c1; ++c) {
      const int m0 = 16 * c + 4 * g;
      double kx[4], th[4];
      float am[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double2 kt = s_kt[m0 - b0 + e];
        kx[e] = kt.x;
        // the product form's S: the column axis' part joins the offset
        th[e] = prod && role == 0 ? __builtin_fma(s_k2[m0 - b0 + e], x2, kt.y) : kt.y;
        am[e] = s_am[m0 - b0 + e];
      }
      f2v v[4];
      table_eval(kx, th, am, x, amp, valid, m0, M, v);
      if (active) table_store<T, NS>(v, role, row, c, g, nch, r.na, r.arows, r.ncp, S, A, C);
    }
  }
}

// Persistent GEMM over the tiles of all factored realisations.  The tiles
// are split into kXcdGroups groups of whole realisations (balanced on the
// host); block b starts on group b % kXcdGroups -- workgroups are dispatched
// to the XCDs round-robin, so a realisation's tables stay in one XCD's L2 --
// and, once its group is exhausted, helps the others in turn.  Every group
// counter only grows, so every block leaves every group and exits.
template <typename T, int NS, bool KR, int ACC, int RB>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(gemm_waves<ACC, RB>(), gemm_waves<ACC, RB>())))
k_ens_gemm(const EnsReal* __restrict__ d, const int* __restrict__ fact,
           const int64_t* __restrict__ tile_start, int nfact,
           const int64_t* __restrict__ group_start, int* __restrict__ counters, bool out_f64) {
  constexpr int JN = cols_blocks<NS, ACC>();
  __shared__ __attribute__((aligned(16))) GemmSmem<T, NS, JN, b_slots<RB>()> sm;
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
      // a transposed realisation (2-D): rows are y, columns x of the C-order
      // (x, y) output, so the row stride is 1 and the column stride nrow
      TileJob jb = make_job<KR, JN, RB>(r.A, r.arows, r.mp / 16, r.S, r.na, r.C, r.ncp, r.nrow,
                                        r.nb_out, r.out, out_f64, r.trans ? 1 : r.ncols,
                                        r.ncols, r.scale, rt, ct, r.trans ? r.nrow : 1,
                                        r.col_tiles, r.ncp, 0);
      if (KR && r.prod) {
        // product form: S row = tile, A and C tables one tile each
        jb.a = local;
        jb.arow0 = 0;
        jb.ccol0 = 0;
        jb.col0 = ct * (32 * JN);
        jb.out_row0 = rt * r.arows;
        jb.nvalid = r.nrow - rt * r.arows;
      }
      if (r.trans) gemm_tile<T, NS, KR, ACC, JN, RB, true>(sm, jb);
      else gemm_tile<T, NS, KR, ACC, JN, RB, false>(sm, jb);
    }
  }
}

}  // namespace gemm

template <typename T, int NS>
hipError_t tables_launch(const TableArgs& a, dim3 grid, hipStream_t s) {
  gemm::k_tables<T, NS><<<grid, 256, 0, s>>>(a.modes, a.M, a.dim, a.nch, a.xs, a.xa, a.xc, a.na,
                                              a.nrow, a.arows, a.ncols, a.ncp, a.S, a.A, a.C,
                                              a.trans, a.ncomp, a.pg);
  return hipGetLastError();
}

template <typename T, int NS>
hipError_t factored_gemm_launch(const GemmArgs& a, dim3 grid, hipStream_t s) {
  using namespace gemm;
#define GSC_GR(KR, ACC, RB)                                                                \
  do {                                                                                     \
    if (a.ldc != 1) k_factored_gemm<T, NS, KR, ACC, true, RB><<<grid, 256, 0, s>>>(a);         \
    else k_factored_gemm<T, NS, KR, ACC, false, RB><<<grid, 256, 0, s>>>(a);                  \
  } while (0)
#define GSC_G(KR, ACC)                        \
  do {                                        \
    if (a.brows == 128) GSC_GR(KR, ACC, 1);   \
    else GSC_GR(KR, ACC, 2);                  \
  } while (0)
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
#undef GSC_GR
  return hipGetLastError();
}

template <typename T, int NS>
hipError_t ens_launch(const EnsLaunch& e, int part, int accum, bool out_f64, hipStream_t s) {
  using namespace gemm;
  if (part == 0) {
    k_ens_tables<T, NS><<<static_cast<unsigned>(e.n_table_tasks), 256, 0, s>>>(
        e.d, e.fact, e.task_start, e.nfact);
    return hipGetLastError();
  }
  for (int q = 0; q < e.ngemm; ++q) {
    const EnsGemmGroup& g = e.gemm[q];
    if (g.nfact == 0) continue;
#define GSC_ER(KR, ACC, RB)                                                                \
  k_ens_gemm<T, NS, KR, ACC, RB><<<g.gemm_blocks, 256, 0, s>>>(e.d, g.fact, g.tile_start, \
                                                              g.nfact, g.group_start,    \
                                                              g.counters, out_f64)
#define GSC_E(KR, ACC)                          \
  do {                                          \
    if (g.brows == 128) GSC_ER(KR, ACC, 1);     \
    else GSC_ER(KR, ACC, 2);                    \
  } while (0)
    if (g.kr) {
      if (accum == ACC_F32X2) GSC_E(true, ACC_F32X2);
      else if (accum == ACC_F32) GSC_E(true, ACC_F32);
      else GSC_E(true, ACC_F64);
    } else {
      if (accum == ACC_F32X2) GSC_E(false, ACC_F32X2);
      else if (accum == ACC_F32) GSC_E(false, ACC_F32);
      else GSC_E(false, ACC_F64);
    }
#undef GSC_E
#undef GSC_ER
  }
  return hipGetLastError();
}

}  // namespace gsc
