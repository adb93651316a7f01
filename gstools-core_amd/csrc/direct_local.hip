// Direct summation with local phases on the matrix cores (point lists, dim 2
// and 3): the "binned" direct variant.
//
// The phase of point x and mode m splits at the centre X_c of a spatial cell:
//   t = kap . x + th = (kap . X_c + th) + kap . xi,   xi = x - X_c.
// r0(c, m) = (kap . X_c + th) - rint(...) is evaluated in FP64 once per
// (wave, mode) and rounded to FP32; kap . xi is small (|xi| <= half a cell)
// and is computed, together with + r0, on the matrix cores as a product
//   D[i][j] = sum_k A[i][k] B[k][j]
// of split factors whose parts are exact in FP32 products, in two forms:
// - scalar fields (k_direct_local32): 32 points x 32 modes per
//   v_mfma_f32_32x32x16_f16, with power-of-two scaled xi and kap split into
//   two fp16 parts each (22 significant bits; products hi hi, hi lo, lo hi
//   per dimension) and r0 split into two fp16 parts as two more K slots;
// - incompressible fields (k_direct_local_vec): 16 points x 16 modes per
//   v_mfma_f32_16x16x32_bf16, with xi and kap split into three bf16 parts
//   each (24 significant bits) and K holding, per dimension d, the six
//   products of parts whose orders sum to at most 2 (h h, h m, m h, h l,
//   m m, l h) in slots 8 d .. 8 d + 5; r0 enters through three more K slots
//   (from the per-cell table) or as the FP32 accumulator C (evaluated per
//   wave pair when the table would be too large).
// The dropped products are below 2^-22 (fp16) or 2^-24 (bf16) of the phase.
// The mode-only parts of the B fragments are read from tables shared by all
// waves; r0 comes from per-cell rows (or per-wave LDS tables) and the
// amplitudes from LDS.  The VALU is left with one v_sin_f32 (argument in
// revolutions) and one FMA per point-mode (D for the vector form); on gfx950
// the matrix instructions hardly overlap that VALU stream, which is why the
// scalar form uses the instruction with the fewest matrix cycles per phase.
// FP32 partial sums over a chunk of modes go into FP32 (scalar) or FP64
// (vector) totals; the mode columns of a point are summed in FP64 (the scalar
// form first adds groups of four columns in FP32).
//
// Modes whose local phase could exceed kRoughRev revolutions in some cell
// (sum_k |kap_k| h_k / 2 > kRoughRev, e.g. the far tail of a heavy-tailed
// spectrum) keep the FP64 phase: after the matrix-core pass every lane
// evaluates them for its own point with the VALU mode loop of the other
// direct kernels.
//
// Cells: the bounding box of the finite points is cut into a grid of about
// N / kCellPts cells (at most kMaxCells), numbered in blocks of f0 x f1 x f2
// cells (at most 8 blocks per dimension) so that neighbouring cells have
// nearby numbers.  A counting sort lists the points of every cell, padded to
// whole waves of kWavePts slots, so every wave has one origin X_c: point
// block b (a contiguous range) counts its points per cell in LDS, a column
// scan over the blocks gives every (block, cell) pair its base slot, and the
// scatter ranks the points of one block and cell with LDS atomics and writes
// their records (xi, index) -- no device-scope atomics, which are slow across
// the XCDs.  A point's result depends only on its own cell, never on the slot
// it lands in, so the output is bitwise reproducible.
#include "gsc_device.h"
#include "gsc_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

namespace gsc {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

// K slots of the split phase product, per dimension d: slots 8 d + s (s < 6)
// hold (x part kXPart[s]) * (kap part kKPart[s]); the rest are zero.
constexpr int kXPart[6] = {0, 0, 1, 0, 1, 2};
constexpr int kKPart[6] = {0, 1, 0, 2, 1, 0};

__device__ __forceinline__ void split3(float x, uint16_t (&p)[3]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const __bf16 h = static_cast<__bf16>(r);
    p[i] = __builtin_bit_cast(uint16_t, h);
    r = r - static_cast<float>(h);
  }
}

// A fragment (16 points x 32 K slots, this lane's 8 slots) of the point
// whose record starts at rec: lane group kg = lane / 16 holds slots
// 8 kg .. 8 kg + 7, the split parts of coordinate kg scaled by sc = 2^-e_kg
// (the power-of-two cell scaling of the fp16 form below; zero for kg >= D).
template <int D>
__device__ __forceinline__ bf8 a_fragment(const float* rec, int kg, float sc) {
  const float x = rec[kg];  // kg = 3: the index field, not used
  uint16_t p[3];
  split3(kg < D ? x * sc : 0.0f, p);
  uint16_t v[8];
#pragma unroll
  for (int t = 0; t < 6; ++t) v[t] = p[kXPart[t]];
  v[6] = v[7] = 0;
  return __builtin_bit_cast(bf8, v);
}

// ---- fp16 form of the phase product (32 x 32 x 16 matrix instructions) ----
// Coordinates are scaled per dimension by 2^-e with 2^e > h / 2 >= 2^(e-1)
// (h the cell width), so |u| = |xi| 2^-e <= 1, and kap by 2^e, so
// sum_k |kap'_k| <= 2 kRoughRev; both scalings are exact.  Each factor is
// split into two fp16 parts (x = hi + lo + O(2^-22 |x|)) and, per dimension,
// the three products hi hi, hi lo, lo hi go into K slots 3 d .. 3 d + 2
// (the dropped lo lo is below 2^-22 of the term).  Slots 10 and 11 (one
// 32-bit word of the fragment) hold r0 = r0_hi + r0_lo against a constant 1
// in A; slots 9 and 12..15 are zero.
// A lane holds the eight slots 8 hl .. 8 hl + 7 (hl = lane / 32), so the
// lower half-wave's B fragment (slots 0..7) depends on the mode only, and the
// upper half's (slots 8..15) on the mode's r0 in the wave's cell as well.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16x __attribute__((ext_vector_type(16)));

// e with 2^e > h / 2 >= 2^(e-1), clamped to [-126, 126] so that 2^-e stays
// a finite normal FP32 number (kap' = kap 2^e <= kap h <= 2 for smooth modes
// needs no bound).  The clamp is only reached by cells narrower than 2^-126,
// whose offsets xi the FP32 point records cannot hold anyway.
__device__ __forceinline__ int half_cell_exp(double h) {
  int e;
  (void)frexp(0.5 * h, &e);
  return min(max(e, -126), 126);
}

__device__ __forceinline__ void split2h(float x, uint16_t& hi, uint16_t& lo) {
  const _Float16 h = static_cast<_Float16>(x);
  hi = __builtin_bit_cast(uint16_t, h);
  lo = __builtin_bit_cast(uint16_t, static_cast<_Float16>(x - static_cast<float>(h)));
}

// A fragment (32 points x 16 K slots, this lane's slots 8 hl .. 8 hl + 7) of
// the point whose record starts at rec; usc[k] = 2^-e_k.
template <int D>
__device__ __forceinline__ h8 a_fragment32(const float* rec, const float (&usc)[3], int hl) {
  uint16_t ph[3], pl[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) split2h(k < D ? rec[k] * usc[k] : 0.0f, ph[k], pl[k]);
  constexpr uint16_t kOne = 0x3C00;  // fp16 1.0
  uint16_t v[8];
  if (hl == 0) {
    v[0] = ph[0]; v[1] = ph[0]; v[2] = pl[0];
    v[3] = ph[1]; v[4] = ph[1]; v[5] = pl[1];
    v[6] = ph[2]; v[7] = ph[2];
  } else {
    v[0] = pl[2]; v[1] = 0; v[2] = kOne; v[3] = kOne;
    v[4] = v[5] = v[6] = v[7] = 0;
  }
  return __builtin_bit_cast(h8, v);
}

constexpr int kCellPts = 4096;      // target points per cell
constexpr int kMinCellPts = 64;     // the cell count never exceeds N / kMinCellPts
constexpr double kCellRev = 0.75;   // cells keep sum_k |kap_k| h_k / 2 below this (revolutions)
                                    // for all but the widest kBandSkip of the modes
constexpr double kBandSkip = 0.005;
constexpr int kQBins = 512;         // quarter-octave bins of sum_k |kap_k|
constexpr int kMaxCells = 16384;    // cell counters of one block fill 64 KB of LDS
constexpr int kMaxSortBlocks = 256;
constexpr int kSortThreads = 1024;  // threads of the histogram and scatter blocks
constexpr int kSortUnroll = 4;      // points in flight per thread there
constexpr int kColBlocks = 16;      // point blocks per group of the column scan
constexpr int kColGroups = kMaxSortBlocks / kColBlocks;
constexpr int kWavePts = 64;        // point slots per record group (one wave)
constexpr int kModeChunk = 1024;    // modes per (r0, amplitude) table of a wave (32 tiles)
constexpr double kRoughRev = 1.0;   // bound on |kap . xi| (revolutions) of the FP32 phase
constexpr int kBoxBlocks = 1024;    // partial bounding boxes
constexpr int kBoxUnroll = 4;       // points in flight per thread there
constexpr size_t kR0TableCap = size_t(256) << 20;  // largest (cell, mode) r0 table
constexpr int kVecR0Pad = 64;       // zero r0 entries past the smooth modes in a vector-kernel row

struct alignas(16) BinParams {
  double lo[3];
  double h[3];
  double inv_h[3];
  double inv_f[3];  // 1 / f
  int n[3];       // cells per dimension
  int f[3];       // cells per coarse cell and dimension
  int C[3];       // coarse cells per dimension
  int F;          // f0 f1 f2
  int ncoarse;    // C0 C1 C2
  int ncells;     // ncoarse F: cell numbers (coarse-major, some empty past n)
  int nwaves;
  int n_smooth;
  int n_rough;
};

struct alignas(32) SmoothRec {  // FP64 data of a smooth mode for r0
  double kap[3];
  double th;
};

constexpr size_t al256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

struct BinLayout {
  int64_t max_cells, max_waves, nblk, blk_pts;
  size_t params, box, centre, bcount, counts, offsets, wave_cell, recs, btpl, btpl32, amp, sm,
      rough, r0w, bytes;
  int64_t r0ld;   // row length of the (cell, mode) r0 words of the scalar kernel; 0: no table
  int64_t r0vld;  // row length of the (cell, mode) r0 parts of the vector kernel; 0: no table
};

BinLayout bin_layout(int64_t M, int64_t N) {
  BinLayout L{};
  L.max_cells = std::min<int64_t>(kMaxCells, N / kMinCellPts + 1);
  L.max_waves = (N + kWavePts - 1) / kWavePts + L.max_cells;
  // point blocks of at least one round (kSortThreads x kSortUnroll points)
  L.nblk = std::max<int64_t>(1, std::min<int64_t>(kMaxSortBlocks, (N + 4095) / 4096));
  L.blk_pts = (N + L.nblk - 1) / L.nblk;
  const size_t m = static_cast<size_t>(M > 0 ? M : 1);
  size_t off = 0;
  auto take = [&](size_t b) { const size_t o = off; off += al256(b); return o; };
  L.params = take(sizeof(BinParams));
  L.box = take(sizeof(double) * 6 * kBoxBlocks);
  L.centre = take(sizeof(double) * 4 * static_cast<size_t>(L.max_cells));
  L.bcount = take(sizeof(int) * static_cast<size_t>(L.nblk * L.max_cells));
  L.counts = take(sizeof(int) * static_cast<size_t>(L.max_cells));
  L.offsets = take(sizeof(int) * static_cast<size_t>(L.max_cells + 1));

  L.wave_cell = take(sizeof(int) * static_cast<size_t>(L.max_waves));
  L.recs = take(sizeof(float4) * static_cast<size_t>(L.max_waves) * kWavePts);
  L.btpl = take(sizeof(uint4) * 4 * (m + kModeChunk));
  L.btpl32 = take(sizeof(uint4) * 2 * (m + kModeChunk));
  L.amp = take(sizeof(float4) * (m + kModeChunk));
  L.sm = take(sizeof(SmoothRec) * (m + kModeChunk));
  L.rough = take(sizeof(ModeRec) * m);
  // r0 of every (cell, smooth mode), when the table stays below kR0TableCap
  // bytes; otherwise every wave evaluates its cell's r0 itself.  The scalar
  // kernel's fp16 words (4 bytes, rows padded by a chunk) and the vector
  // kernel's bf16 parts (8 bytes, rows padded by kVecR0Pad) share the region.
  const int64_t r0ld = static_cast<int64_t>(m) + kModeChunk;
  const int64_t r0vld = static_cast<int64_t>(m) + kVecR0Pad;
  const size_t r0bytes = sizeof(uint32_t) * static_cast<size_t>(L.max_cells * r0ld);
  const size_t r0vbytes = sizeof(uint2) * static_cast<size_t>(L.max_cells * r0vld);
  L.r0ld = r0bytes <= kR0TableCap ? r0ld : 0;
  L.r0vld = r0vbytes <= kR0TableCap ? r0vld : 0;
  const size_t region = std::max(L.r0ld > 0 ? r0bytes : 0, L.r0vld > 0 ? r0vbytes : 0);
  L.r0w = region > 0 ? take(region) : 0;
  L.bytes = off;
  return L;
}

template <typename T>
T* at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(base) + off);
}

// ---- pre-pass: bounding box ------------------------------------------------
// Blocks [0, kBoxBlocks): partial bounding boxes; the blocks after them (when
// pa.ntiles > 0) prepare the mode table, 256 modes each (prep_mode).
template <int D>
__global__ void __launch_bounds__(256)
k_box_partial(const double* __restrict__ pos, int64_t ld, int64_t N, double* __restrict__ box,
              PrepArgs pa) {
  if (blockIdx.x >= kBoxBlocks) {  // uniform over the block
    const int64_t m = static_cast<int64_t>(blockIdx.x - kBoxBlocks) * 256 + threadIdx.x;
    if (m < static_cast<int64_t>(pa.ntiles) * 16)
      prep_mode(pa.kind, pa.dim, pa.M, m, pa.cov, pa.z1, pa.z2, pa.sf, pa.table, pa.a_tab,
                pa.amp_tab);
    return;
  }
  __shared__ double red[2 * D][256];
  double lo[D], hi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    lo[k] = INFINITY;
    hi[k] = -INFINITY;
  }
  // kBoxUnroll points in flight per thread; out-of-range ones repeat point
  // N - 1, which changes no extreme
  const int64_t stride = static_cast<int64_t>(kBoxBlocks) * 256;
  for (int64_t pb = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; pb < N;
       pb += kBoxUnroll * stride) {
    double x[kBoxUnroll][D];
#pragma unroll
    for (int u = 0; u < kBoxUnroll; ++u) {
      const int64_t p = min(pb + u * stride, N - 1);
#pragma unroll
      for (int k = 0; k < D; ++k) x[u][k] = pos[k * ld + p];
    }
#pragma unroll
    for (int u = 0; u < kBoxUnroll; ++u)
#pragma unroll
      for (int k = 0; k < D; ++k)
        if (isfinite(x[u][k])) {
          lo[k] = fmin(lo[k], x[u][k]);
          hi[k] = fmax(hi[k], x[u][k]);
        }
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    red[2 * k][threadIdx.x] = lo[k];
    red[2 * k + 1][threadIdx.x] = hi[k];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        red[2 * k][threadIdx.x] = fmin(red[2 * k][threadIdx.x], red[2 * k][threadIdx.x + s]);
        red[2 * k + 1][threadIdx.x] =
            fmax(red[2 * k + 1][threadIdx.x], red[2 * k + 1][threadIdx.x + s]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < D; ++k) {
      box[blockIdx.x * 6 + 2 * k] = red[2 * k][0];
      box[blockIdx.x * 6 + 2 * k + 1] = red[2 * k + 1][0];
    }
}

// Modes of a 256-thread block held in registers: mode 256 r + thread for
// r < kPreModes (M <= 1024 in one load per thread, issued before anything
// waits on it); later modes are read from the table where they are used.
constexpr int kPreModes = 4;
struct PreModes {
  ModeRec r[kPreModes];
};
__device__ __forceinline__ void pre_load(const ModeRec* __restrict__ modes, int M, PreModes& p) {
#pragma unroll
  for (int r = 0; r < kPreModes; ++r) {
    const int m = 256 * r + static_cast<int>(threadIdx.x);
    p.r[r] = m < M ? modes[m] : ModeRec{};
  }
}
// Record of mode 256 rr + thread (m < M); rr is uniform over the block, so
// the register copy is chosen by a scalar branch.
__device__ __forceinline__ ModeRec pre_get(const ModeRec* __restrict__ modes, const PreModes& p,
                                           int rr, int m) {
  switch (rr) {
    case 0: return p.r[0];
    case 1: return p.r[1];
    case 2: return p.r[2];
    case 3: return p.r[3];
    default: return modes[m];
  }
}
static_assert(kPreModes == 4, "pre_get names every register copy");

// One block of 256 threads: modes in order, split into smooth (FP32 local
// phase) and rough (FP64 phase) lists by a block scan, so both keep the
// original mode order; h is the cell width.  Thread 0 receives the list
// lengths.
template <int D>
__device__ void split_modes(const ModeRec* __restrict__ modes, const PreModes& pre, int M,
                            const double* h,
                            uint4* __restrict__ btpl, uint4* __restrict__ btpl32,
                            float4* __restrict__ amp,
                            SmoothRec* __restrict__ sm, ModeRec* __restrict__ rough,
                            int* n_smooth, int* n_rough) {
  __shared__ int wcnt[4];  // smooth modes per wave of the current 256
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int ns_s = 0, nr_s = 0;   // list lengths so far (the same in every thread)
  for (int base = 0; base < M; base += 256) {
    const int m = base + threadIdx.x;
    ModeRec r{};
    bool smooth = false;
    if (m < M) {
      r = pre_get(modes, pre, base / 256, m);
      double b = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) b += fabs(r.kap[k]) * 0.5 * h[k];
      smooth = b <= kRoughRev && isfinite(r.th);
#pragma unroll
      for (int c = 0; c < D; ++c) smooth = smooth && isfinite(static_cast<double>(r.amp[c]));
    }
    // order-keeping split: a ballot per wave, then the waves' counts
    const unsigned long long bal = __ballot(smooth);
    if (lane == 0) wcnt[wv] = __popcll(bal);
    __syncthreads();
    int before = __popcll(bal & ((1ull << lane) - 1));  // smooth modes before m in block
    int total = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = wcnt[q];
      if (q < wv) before += c;
      total += c;
    }
    const int ns0 = ns_s, nr0 = nr_s;
    if (m < M) {
      if (smooth) {
        const int i = ns0 + before;
        // B template of the mode: the kap parts in their slots, r0 slots zero
        uint32_t w[16];
#pragma unroll
        for (int q = 0; q < 16; ++q) w[q] = 0;
#pragma unroll
        for (int d = 0; d < D; ++d) {
          uint16_t kp[3];
          float kf = static_cast<float>(ldexp(r.kap[d], half_cell_exp(h[d])));
          __asm__ volatile("" : "+v"(kf));  // FP32 rounding step kept
          split3(kf, kp);
#pragma unroll
          for (int t = 0; t < 6; ++t) {
            const int k = 8 * d + t;
            w[k >> 1] |= static_cast<uint32_t>(kp[kKPart[t]]) << (16 * (k & 1));
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          btpl[4 * i + q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
        // fp16 form: slots 0..7 (lower half-wave), then kap'_2 hi in slot 8
        uint16_t kh[3], kl[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          float kf = d < D ? static_cast<float>(ldexp(r.kap[d], half_cell_exp(h[d]))) : 0.0f;
          __asm__ volatile("" : "+v"(kf));  // FP32 rounding step kept (see k_direct_local32)
          split2h(kf, kh[d], kl[d]);
        }
        auto pk = [](uint16_t a, uint16_t b) {
          return static_cast<uint32_t>(a) | (static_cast<uint32_t>(b) << 16);
        };
        btpl32[2 * i] = make_uint4(pk(kh[0], kl[0]), pk(kh[0], kh[1]), pk(kl[1], kh[1]),
                                   pk(kh[2], kl[2]));
        btpl32[2 * i + 1] = make_uint4(pk(kh[2], 0), 0u, 0u, 0u);
        amp[i] = make_float4(r.amp[0], r.amp[1], D == 3 ? r.amp[2] : 0.0f, 0.0f);
        sm[i] = SmoothRec{{r.kap[0], r.kap[1], r.kap[2]}, r.th};
      } else {
        rough[nr0 + (threadIdx.x - before)] = r;
      }
    }
    ns_s = ns0 + total;
    nr_s = nr0 + min(256, M - base) - total;
    __syncthreads();  // wcnt is read by every thread before the next 256 overwrite it
  }
  if (threadIdx.x == 0) {
    *n_smooth = ns_s;
    *n_rough = nr_s;
  }
  // the main kernels read the tables up to kModeChunk modes past the last
  // smooth one: zero phase data, zero amplitude and finite B fragments there
  for (int i = threadIdx.x; i < 4 * kModeChunk; i += 256)
    btpl[4 * ns_s + i] = make_uint4(0u, 0u, 0u, 0u);
  for (int i = threadIdx.x; i < 2 * kModeChunk; i += 256)
    btpl32[2 * ns_s + i] = make_uint4(0u, 0u, 0u, 0u);
  for (int i = threadIdx.x; i < kModeChunk; i += 256) {
    amp[ns_s + i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    sm[ns_s + i] = SmoothRec{{0.0, 0.0, 0.0}, 0.0};
  }
}

// Cell grid from the bounding box: near-cubic cells over the dimensions with
// a non-zero extent, about N / kCellPts of them, or more where the modes
// need it: the cell width h is kept below 2 kCellRev / s_q, s_q the (1 -
// kBandSkip) quantile of s_m = sum_k |kap_mk| over the finite modes (to a
// quarter octave), so the local phase of those modes stays below kCellRev
// revolutions.  At most max_cells cells in all.
template <int D>
__global__ void __launch_bounds__(256)
k_bin_params(const double* __restrict__ box, int64_t N, int64_t max_cells,
             const ModeRec* __restrict__ modes, int M, BinParams* __restrict__ prm,
             uint4* __restrict__ btpl, uint4* __restrict__ btpl32, float4* __restrict__ amp,
             SmoothRec* __restrict__ sm, ModeRec* __restrict__ rough) {
  __shared__ BinParams Ps;
  __shared__ double red[2 * D][256];
  __shared__ int qh[kQBins];
  // every global read this single block needs is issued first: the mode
  // records (up to 1024) and the partial boxes
  PreModes pre;
  pre_load(modes, M, pre);
  constexpr int kBoxPer = kBoxBlocks / 256;
  static_assert(kBoxPer * 256 == kBoxBlocks, "partial boxes per thread");
  double bx[kBoxPer][2 * D];
#pragma unroll
  for (int u = 0; u < kBoxPer; ++u)
#pragma unroll
    for (int c = 0; c < 2 * D; ++c) bx[u][c] = box[(threadIdx.x + 256 * u) * 6 + c];
  for (int b = threadIdx.x; b < kQBins; b += 256) qh[b] = 0;
  __syncthreads();
  auto quant = [&](const ModeRec& r) {
    double sk = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) sk += fabs(r.kap[k]);
    if (!isfinite(sk)) return;
    const int b = sk > 0.0 ? static_cast<int>(fmin(fmax(floor(4.0 * log2(sk)) + kQBins / 2, 0.0),
                                                   kQBins - 1.0))
                           : 0;
    atomicAdd(&qh[b], 1);
  };
#pragma unroll
  for (int r = 0; r < kPreModes; ++r)
    if (256 * r + static_cast<int>(threadIdx.x) < M) quant(pre.r[r]);
  for (int m = 256 * kPreModes + threadIdx.x; m < M; m += 256) quant(modes[m]);
  __syncthreads();
  // the quantile's bin: inclusive scan of bin pairs, then the pair that
  // crosses the wanted count
  __shared__ int qs[256];
  __shared__ double s_q_s;
  const int c0 = qh[2 * threadIdx.x], c1 = qh[2 * threadIdx.x + 1];
  qs[threadIdx.x] = c0 + c1;
  if (threadIdx.x == 0) s_q_s = 0.0;
  __syncthreads();
  for (int st = 1; st < 256; st <<= 1) {
    const int v = threadIdx.x >= st ? qs[threadIdx.x - st] : 0;
    __syncthreads();
    qs[threadIdx.x] += v;
    __syncthreads();
  }
  const int nfin = qs[255];
  if (nfin > 0) {
    const int want = static_cast<int>(ceil((1.0 - kBandSkip) * nfin));
    const int before = qs[threadIdx.x] - c0 - c1;
    if (before < want && want <= before + c0 + c1) {
      const int b = before + c0 >= want ? 2 * threadIdx.x : 2 * threadIdx.x + 1;
      s_q_s = b == 0 ? 0.0 : exp2((b + 1 - kQBins / 2) / 4.0);  // upper edge of the bin
    }
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double a = INFINITY, b = -INFINITY;
#pragma unroll
    for (int u = 0; u < kBoxPer; ++u) {
      a = fmin(a, bx[u][2 * k]);
      b = fmax(b, bx[u][2 * k + 1]);
    }
    red[2 * k][threadIdx.x] = a;
    red[2 * k + 1][threadIdx.x] = b;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) {
#pragma unroll
      for (int k = 0; k < D; ++k) {
        red[2 * k][threadIdx.x] = fmin(red[2 * k][threadIdx.x], red[2 * k][threadIdx.x + st]);
        red[2 * k + 1][threadIdx.x] =
            fmax(red[2 * k + 1][threadIdx.x], red[2 * k + 1][threadIdx.x + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double s_q = s_q_s;
    double lo[3] = {0.0, 0.0, 0.0}, ext[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < D; ++k) {
      double a = red[2 * k][0], b = red[2 * k + 1][0];
      if (!(a <= b)) a = b = 0.0;  // no finite coordinate
      lo[k] = a;
      ext[k] = b - a;
    }
    double vol = 1.0;
    int active = 0;
    for (int k = 0; k < D; ++k)
      if (ext[k] > 0.0) {
        vol *= ext[k];
        ++active;
      }
    double target = fmax(1.0, static_cast<double>(N) / kCellPts);
    if (active > 0 && s_q > 0.0)
      target = fmax(target, vol / pow(2.0 * kCellRev / s_q, static_cast<double>(active)));
    target = fmin(target, static_cast<double>(max_cells));
    int n[3] = {1, 1, 1};
    if (active > 0) {
      const double h = pow(vol / target, 1.0 / active);
      for (int k = 0; k < D; ++k)
        if (ext[k] > 0.0) n[k] = static_cast<int>(fmin(fmax(ceil(ext[k] / h), 1.0), 65535.0));
    }
    // coarse cells of f_k cells per dimension, at most 8 per dimension; the
    // numbering covers whole coarse cells, so it may exceed n0 n1 n2
    int f[3], C[3];
    for (;;) {
      int64_t total = 1;
      for (int k = 0; k < 3; ++k) {
        f[k] = (n[k] + 7) / 8;
        C[k] = (n[k] + f[k] - 1) / f[k];
        total *= static_cast<int64_t>(f[k]) * C[k];
      }
      if (total <= max_cells) break;
      for (int k = 0; k < 3; ++k) n[k] = n[k] > 1 ? n[k] * 4 / 5 : 1;
    }
    BinParams P{};
    for (int k = 0; k < 3; ++k) {
      P.lo[k] = lo[k];
      P.n[k] = n[k];
      P.f[k] = f[k];
      P.C[k] = C[k];
      // a dimension without extent (a planar slice) gets width 0: every
      // point sits at the cell centre there (xi = 0) and the dimension adds
      // nothing to a mode's local-phase bound (half_cell_exp(0) scales by 1)
      P.h[k] = ext[k] > 0.0 ? ext[k] / n[k] : 0.0;
      P.inv_h[k] = ext[k] > 0.0 ? 1.0 / P.h[k] : 0.0;
      P.inv_f[k] = 1.0 / f[k];
    }
    P.F = f[0] * f[1] * f[2];
    P.ncoarse = C[0] * C[1] * C[2];
    P.ncells = P.ncoarse * P.F;
    Ps = P;
  }
  __syncthreads();
  // the modes' smooth / rough split depends on the cell width
  split_modes<D>(modes, pre, M, Ps.h, btpl, btpl32, amp, sm, rough, &Ps.n_smooth, &Ps.n_rough);
  __syncthreads();
  if (threadIdx.x == 0) *prm = Ps;
}

// Cell of point x: index i_k = floor((x_k - lo_k) / h_k) per dimension
// (NaN and x < lo land in cell 0, x >= hi in the last cell), numbered coarse
// cell (c0 C1 + c1) C2 + c2 first, then the cell within it (l0 f1 + l1) f2 +
// l2, where i_k = c_k f_k + l_k.  c_k = floor((i_k + 0.5) / f_k) is taken in
// FP64 (exact: the quotient is at least 0.5 / f_k from an integer) instead
// of an integer division.  X receives the cell centre lo + (i + 1/2) h.
template <int D>
__device__ __forceinline__ int cell_locate(const BinParams& P, const double* x, double (&X)[3]) {
  int cc = 0, l = 0;
  X[0] = X[1] = X[2] = 0.0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double fr = (x[k] - P.lo[k]) * P.inv_h[k];
    int ik = 0;
    if (fr >= 0.0) ik = fr < static_cast<double>(P.n[k]) ? static_cast<int>(fr) : P.n[k] - 1;
    const int ck = static_cast<int>((ik + 0.5) * P.inv_f[k]);
    cc = cc * P.C[k] + ck;
    l = l * P.f[k] + (ik - ck * P.f[k]);
    X[k] = P.lo[k] + (ik + 0.5) * P.h[k];
  }
  return cc * P.F + l;
}

// Positions of points pb + u kSortThreads (u < kSortUnroll), clamped to
// plast so that every load is in bounds.
template <int D>
__device__ __forceinline__ void load_points(const double* __restrict__ pos, int64_t ld, int64_t pb,
                                            int64_t plast, double (&x)[kSortUnroll][D]) {
#pragma unroll
  for (int u = 0; u < kSortUnroll; ++u) {
    const int64_t p = min(pb + u * kSortThreads, plast);
#pragma unroll
    for (int k = 0; k < D; ++k) x[u][k] = pos[k * ld + p];
  }
}

// Point block b (points [b blk_pts, (b + 1) blk_pts)) counts its points per
// cell in LDS; bcount[b][c] receives the counts.  1024 threads, kSortUnroll
// points per thread and round.  Whole rounds (every point in range) run
// without branches and load the next round's positions before the current
// round's atomics; the last, partial round checks each point.
template <int D>
__global__ void __launch_bounds__(kSortThreads)
k_bin_hist(const double* __restrict__ pos, int64_t ld, int64_t N, int64_t blk_pts,
           const BinParams* __restrict__ prm, int* __restrict__ bcount) {
  __shared__ int hist[kMaxCells];
  const BinParams P = *prm;
  for (int c = threadIdx.x; c < P.ncells; c += kSortThreads) hist[c] = 0;
  __syncthreads();
  constexpr int64_t kRound = kSortThreads * kSortUnroll;
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * blk_pts;
  const int64_t p1 = min(p0 + blk_pts, N);
  int64_t base = p0;
  if (base + kRound <= p1) {
    double x[kSortUnroll][D], xn[kSortUnroll][D];
    load_points<D>(pos, ld, base + threadIdx.x, p1 - 1, xn);
    for (; base + kRound <= p1; base += kRound) {
#pragma unroll
      for (int u = 0; u < kSortUnroll; ++u)
#pragma unroll
        for (int k = 0; k < D; ++k) x[u][k] = xn[u][k];
      load_points<D>(pos, ld, base + kRound + threadIdx.x, p1 - 1, xn);
#pragma unroll
      for (int u = 0; u < kSortUnroll; ++u) {
        double X[3];
        atomicAdd(&hist[cell_locate<D>(P, x[u], X)], 1);
      }
    }
  }
#pragma unroll
  for (int u = 0; u < kSortUnroll; ++u) {
    const int64_t p = base + threadIdx.x + u * kSortThreads;
    if (p < p1) {
      double x[D], X[3];
#pragma unroll
      for (int k = 0; k < D; ++k) x[k] = pos[k * ld + p];
      atomicAdd(&hist[cell_locate<D>(P, x, X)], 1);
    }
  }
  __syncthreads();
  int* row = bcount + static_cast<int64_t>(blockIdx.x) * P.ncells;
  for (int c = threadIdx.x; c < P.ncells; c += kSortThreads) row[c] = hist[c];
}

// Exclusive scan of every column of bcount over the point blocks, and the
// cell totals, in one launch: a block takes 16 cells x kColGroups groups of
// kColBlocks point blocks; each thread sums its group, the 16 group sums of a
// cell are scanned in LDS, and each thread rewrites its group from its offset.
// Neighbouring threads take neighbouring cells, so every access to a row of
// bcount is a run of 16 consecutive words.
constexpr int kColCells = 256 / kColGroups;
static_assert(kColCells * kColGroups == 256, "column scan block shape");
__global__ void __launch_bounds__(256)
k_col_scan(const BinParams* __restrict__ prm, int nblk, int* __restrict__ bcount,
           int* __restrict__ counts) {
  __shared__ int gs[kColCells][kColGroups + 1];
  const int ncells = prm->ncells;
  const int cl = threadIdx.x % kColCells, g = threadIdx.x / kColCells;
  const int c = blockIdx.x * kColCells + cl;
  const bool on = c < ncells;
  const int b0 = g * kColBlocks, b1 = min(nblk, b0 + kColBlocks);
  int run = 0;
  if (on)
    for (int b = b0; b < b1; ++b) run += bcount[static_cast<int64_t>(b) * ncells + c];
  gs[cl][g] = run;
  __syncthreads();
  if (g == 0 && on) {
    int sum = 0;
    for (int q = 0; q < kColGroups; ++q) {
      const int v = gs[cl][q];
      gs[cl][q] = sum;
      sum += v;
    }
    counts[c] = sum;
  }
  __syncthreads();
  if (on) {
    run = gs[cl][g];
    for (int b = b0; b < b1; ++b) {
      int* e = bcount + static_cast<int64_t>(b) * ncells + c;
      const int v = *e;
      *e = run;
      run += v;
    }
  }
}

// Exclusive scan of a[0, n) in place (n <= 4 kSortThreads) by all
// kSortThreads threads: four entries per thread, wave scans with shuffles,
// one LDS word per wave.  Returns the total.
__device__ __forceinline__ int block_scan4(int* a, int n, int* wsum) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int v[4], s = 0;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int i = 4 * t + u;
    v[u] = i < n ? a[i] : 0;
    s += v[u];
  }
  int x = s;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int pre = 0, total = 0;
#pragma unroll
  for (int k = 0; k < kSortThreads / 64; ++k) {
    pre += k < w ? wsum[k] : 0;
    total += wsum[k];
  }
  int run = pre + x - s;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int i = 4 * t + u;
    if (i < n) a[i] = run;
    run += v[u];
  }
  __syncthreads();
  return total;
}

// One block: cell c gets ceil(count / kWavePts) waves starting at slot
// offsets[c]; offsets[ncells] and the wave count close the list.
__global__ void __launch_bounds__(kSortThreads)
k_bin_offsets(BinParams* __restrict__ prm, const int* __restrict__ counts,
              int* __restrict__ offsets) {
  __shared__ int wv[4 * kSortThreads];
  __shared__ int wsum[kSortThreads / 64];
  const int ncells = prm->ncells;
  int carry = 0;  // uniform
  for (int base = 0; base < ncells; base += 4 * kSortThreads) {
    const int n = min(4 * kSortThreads, ncells - base);
    for (int i = threadIdx.x; i < n; i += kSortThreads)
      wv[i] = (counts[base + i] + kWavePts - 1) / kWavePts;
    __syncthreads();
    const int tw = block_scan4(wv, n, wsum);
    for (int i = threadIdx.x; i < n; i += kSortThreads) offsets[base + i] = (carry + wv[i]) * kWavePts;
    carry += tw;
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    offsets[ncells] = carry * kWavePts;
    prm->nwaves = carry;
  }
}

// Point record of a slot: local coordinates xi = x - X_cell (FP32) and the
// point index (-1 for padding) in the last lane.
__device__ __forceinline__ float4 make_rec(float a, float b, float c, int idx) {
  return make_float4(a, b, c, __builtin_bit_cast(float, idx));
}

// Centre of cell number e (FP64).
template <int D>
__device__ __forceinline__ void cell_centre(const BinParams& P, int e, double (&X)[3]) {
  X[0] = X[1] = X[2] = 0.0;
  int cc = e / P.F, l = e - (e / P.F) * P.F;
#pragma unroll
  for (int k = D - 1; k >= 0; --k) {
    const int ck = cc % P.C[k], lk = l % P.f[k];
    cc /= P.C[k];
    l /= P.f[k];
    X[k] = P.lo[k] + (ck * P.f[k] + lk + 0.5) * P.h[k];
  }
}

// Up to kWaveSlices waves per cell while the cells are few (a part of a row
// per wave keeps the dependent loads per wave short, so a few cells still
// spread over many waves; with thousands of cells one wave per cell is
// faster, every slice repeating the cell's set-up), over a grid-stride loop
// of (cell, slice) items.  Slice 0 records the
// cell's centre and its waves in the wave -> cell map and marks the slots
// past its points as padding; when r0w is given, every slice
// writes its part of the cell's row of r0 (the reduced FP64 phase at the
// centre, rounded to FP32) for every smooth mode, which the main kernels
// read instead of evaluating r0 in every wave of the cell:
// - scalar form (vec = false): words r0_hi | r0_lo << 16 (fp16 parts) up to
//   the end of the last chunk, zero phase data past n_smooth;
// - vector form: pairs (p0 | p1 << 16, p2) of the three bf16 parts, r0ld of
//   them, zero past n_smooth.
constexpr int kWaveSlices = 4;  // waves per cell for its row, at most
constexpr int kWavesMax = 8192; // waves for the cell rows, for extra slices
constexpr int kSliceWaves = 2048;  // cells x slices up to which cells are sliced
template <int D>
__device__ __forceinline__ void cell_slice(const BinParams& P, int e, int slice, int slices,
                                           int lane,
                                           const int* __restrict__ counts,
                                           const int* __restrict__ offsets,
                                           double* __restrict__ centre,
                                           int* __restrict__ wave_cell,
                                           float4* __restrict__ recs,
                                           const SmoothRec* __restrict__ sm,
                                           uint32_t* __restrict__ r0w, int64_t r0ld, bool vec) {
  double X[3];
  cell_centre<D>(P, e, X);
  const int s0 = offsets[e], s1 = offsets[e + 1];
  if (slice == 0) {
    if (lane == 0)
#pragma unroll
      for (int k = 0; k < 3; ++k) centre[4 * e + k] = X[k];
    for (int s = s0 + kWavePts * lane; s < s1; s += kWavePts * 64) wave_cell[s / kWavePts] = e;
    const int sp = s0 + counts[e] + lane;  // padding: fewer than kWavePts slots
    if (sp < s1) recs[sp] = make_rec(0.0f, 0.0f, 0.0f, -1);
  }
  if (r0w == nullptr || s1 == s0) return;  // no table, or no wave reads this row
  if (vec) {
    uint2* row = reinterpret_cast<uint2*>(r0w) + e * r0ld;
    // the slice's entries: a quarter of the row, whole 64-entry steps
    const int64_t per = (r0ld + 64 * slices - 1) / (64 * slices) * 64;
    const int64_t i1 = min(r0ld, per * (slice + 1));
#pragma unroll 4
    for (int64_t i = per * slice + lane; i < i1; i += 64) {
      uint2 v = make_uint2(0u, 0u);
      if (i < P.n_smooth) {
        const SmoothRec sr = sm[i];
        double t = __builtin_fma(sr.kap[0], X[0], sr.th);
        t = __builtin_fma(sr.kap[1], X[1], t);
        if constexpr (D == 3) t = __builtin_fma(sr.kap[2], X[2], t);
        float r0 = static_cast<float>(t - rint(t));
        __asm__ volatile("" : "+v"(r0));  // FP32 rounding step kept
        uint16_t p[3];
        split3(r0, p);
        v = make_uint2(static_cast<uint32_t>(p[0]) | (static_cast<uint32_t>(p[1]) << 16), p[2]);
      }
      row[i] = v;
    }
    return;
  }
  uint32_t* row = r0w + e * r0ld;
  const int end = (P.n_smooth + kModeChunk - 1) / kModeChunk * kModeChunk;
  const int per = (end + 64 * slices - 1) / (64 * slices) * 64;
  const int i1 = min(end, per * (slice + 1));
#pragma unroll 4
  for (int i = per * slice + lane; i < i1; i += 64) {
    const SmoothRec sr = sm[i];
    double t = __builtin_fma(sr.kap[0], X[0], sr.th);
    t = __builtin_fma(sr.kap[1], X[1], t);
    if constexpr (D == 3) t = __builtin_fma(sr.kap[2], X[2], t);
    float r0 = static_cast<float>(t - rint(t));
    __asm__ volatile("" : "+v"(r0));  // FP32 rounding step kept (see k_direct_local32)
    uint16_t rh, rl;
    split2h(r0, rh, rl);
    row[i] = static_cast<uint32_t>(rh) | (static_cast<uint32_t>(rl) << 16);
  }
}


// Point block b again (same threads and rounds as k_bin_hist): ranks its
// points within each cell with LDS atomics (starting from the block's base in
// the cell) and writes their records (local coordinates, index) to their
// slots.
template <int D>
__device__ __forceinline__ void scatter_point(const BinParams& P, const double* x, int64_t p,
                                              int* rank, float4* __restrict__ recs) {
  double X[3];
  const int e = cell_locate<D>(P, x, X);  // the cell k_bin_hist counted it in
  float xi[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int k = 0; k < D; ++k) xi[k] = static_cast<float>(x[k] - X[k]);
  recs[atomicAdd(&rank[e], 1)] = make_rec(xi[0], xi[1], xi[2], static_cast<int>(p));
}

// The two placement steps in one launch; they write disjoint memory (the
// scatter the slots below each cell's count, cell_slice the padding slots
// above it, the centres, the wave -> cell map and the r0 rows).  Blocks
// [0, nblk) scatter point block b; the blocks after them take the (cell,
// slice) items of cell_slice, kSortThreads / 64 waves per block, in a
// grid-stride loop over nwv waves.
template <int D>
__global__ void __launch_bounds__(kSortThreads)
k_bin_scatter(const double* __restrict__ pos, int64_t ld, int64_t N, int64_t blk_pts, int nblk,
              const BinParams* __restrict__ prm, const int* __restrict__ bcount,
              const int* __restrict__ counts, const int* __restrict__ offsets,
              float4* __restrict__ recs, double* __restrict__ centre, int* __restrict__ wave_cell,
              const SmoothRec* __restrict__ sm, uint32_t* __restrict__ r0w, int64_t r0ld,
              bool vec) {
  __shared__ int rank[kMaxCells];
  const BinParams P = *prm;
  if (static_cast<int>(blockIdx.x) >= nblk) {  // uniform over the block
    constexpr int kWpb = kSortThreads / 64;
    const int nwv = (static_cast<int>(gridDim.x) - nblk) * kWpb;
    const int slices = max(1, min(kWaveSlices, min(nwv, kSliceWaves) / max(P.ncells, 1)));
    const int items = P.ncells * slices;
    for (int it = (static_cast<int>(blockIdx.x) - nblk) * kWpb + static_cast<int>(threadIdx.x >> 6);
         it < items; it += nwv)  // uniform per wave
      cell_slice<D>(P, it / slices, it % slices, slices, threadIdx.x & 63, counts, offsets, centre,
                    wave_cell, recs, sm, r0w, r0ld, vec);
    return;
  }
  const int* cbase = bcount + static_cast<int64_t>(blockIdx.x) * P.ncells;
  for (int e = threadIdx.x; e < P.ncells; e += kSortThreads) rank[e] = offsets[e] + cbase[e];
  __syncthreads();
  constexpr int64_t kRound = kSortThreads * kSortUnroll;
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * blk_pts;
  const int64_t p1 = min(p0 + blk_pts, N);
  int64_t base = p0;
  if (base + kRound <= p1) {
    double x[kSortUnroll][D], xn[kSortUnroll][D];
    load_points<D>(pos, ld, base + threadIdx.x, p1 - 1, xn);
    for (; base + kRound <= p1; base += kRound) {
#pragma unroll
      for (int u = 0; u < kSortUnroll; ++u)
#pragma unroll
        for (int k = 0; k < D; ++k) x[u][k] = xn[u][k];
      load_points<D>(pos, ld, base + kRound + threadIdx.x, p1 - 1, xn);
#pragma unroll
      for (int u = 0; u < kSortUnroll; ++u)
        scatter_point<D>(P, x[u], base + threadIdx.x + u * kSortThreads, rank, recs);
    }
  }
#pragma unroll
  for (int u = 0; u < kSortUnroll; ++u) {
    const int64_t p = base + threadIdx.x + u * kSortThreads;
    if (p < p1) {
      double x[D];
#pragma unroll
      for (int k = 0; k < D; ++k) x[k] = pos[k * ld + p];
      scatter_point<D>(P, x, p, rank, recs);
    }
  }
}

// ---- main kernel -------------------------------------------------------------
// 32 x 32 form of the local kernel: one v_mfma_f32_32x32x16_f16 gives the
// phases of 32 points x 32 modes (the fp16 product above), about half the
// matrix-core time per point-mode of four 16 x 16 x 32 bf16 tiles; the VALU
// work (one sine and one FMA per point-mode) is the same.  One wave per
// record group (kWavePts points, two row groups of 32), as k_direct_local,
// with its own LDS table per chunk of kModeChunk modes: per mode the word
// r0_hi | r0_lo << 16, which the upper half-wave ORs into the B fragment read
// from the per-mode table (slots 10 and 11), and the amplitude, which every
// lane takes for its mode column (8-byte records: 32 distinct ones per read,
// no bank conflicts).  The two row groups' MFMAs of tile q + 1 are issued
// right after the sines of their tile q, so each result set has a whole
// group's VALU work to complete in.  Output element v of lane l in group g is
// point 32 g + 8 (v / 4) + 4 (l / 32) + v % 4, mode column l % 32.
constexpr int kRed32Ld = 33;  // LDS row of the cross-lane sums (floats)
template <int D, typename OutT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_direct_local32(const BinParams* __restrict__ prm, const double* __restrict__ centre,
                 const int* __restrict__ wave_cell, const float4* __restrict__ recs,
                 const uint4* __restrict__ btpl32, const float4* __restrict__ ampt,
                 const SmoothRec* __restrict__ sm, const uint32_t* __restrict__ r0w,
                 int64_t r0ld, const ModeRec* __restrict__ rough,
                 const double* __restrict__ pos, int64_t ld, OutT* __restrict__ out,
                 double scale) {
  constexpr int kWaveLds = kModeChunk * 8 > 64 * kRed32Ld * 4 ? kModeChunk * 8
                                                              : 64 * kRed32Ld * 4;
  __shared__ __attribute__((aligned(16))) char lds[4][kWaveLds];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = blockIdx.x * 4 + wv;
  const BinParams P = *prm;
  if (w >= P.nwaves) return;
  const int n_smooth = P.n_smooth, n_rough = P.n_rough;
  const int cell = wave_cell[w];
  double X[3];
  {
    const double* xc = centre + 4 * static_cast<int64_t>(cell);
    X[0] = xc[0];
    X[1] = xc[1];
    X[2] = xc[2];
  }
  const uint32_t* r0row = r0w != nullptr ? r0w + cell * r0ld : nullptr;
  const int hl = lane >> 5, lr = lane & 31;
  const float* rf = reinterpret_cast<const float*>(recs + static_cast<int64_t>(w) * kWavePts);
  float usc[3];
#pragma unroll
  for (int k = 0; k < 3; ++k)
    usc[k] = k < D ? static_cast<float>(ldexp(1.0, -half_cell_exp(P.h[k]))) : 0.0f;
  h8 a[2];
#pragma unroll
  for (int g = 0; g < 2; ++g) a[g] = a_fragment32<D>(rf + 4 * (32 * g + lr), usc, hl);
  uint2* wt = reinterpret_cast<uint2*>(lds[wv]);
  // upper half-wave: r0 slots from the table (opaque to the compiler, so the
  // merge below stays one v_and_or_b32 instead of a select and an or)
  unsigned mu = hl ? ~0u : 0u;
  __asm__ volatile("" : "+v"(mu));
  f16x zero;
#pragma unroll
  for (int v = 0; v < 16; ++v) zero[v] = 0.0f;
  float acc[2][16], tot[2][16];
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      acc[g][v] = 0.0f;
      tot[g][v] = 0.0f;
    }
  for (int c0 = 0; c0 < n_smooth; c0 += kModeChunk) {
    // the wave's table: r0 of its cell (FP64 phase at the centre, reduced
    // exactly) split into two fp16 parts, and the amplitude; zero phase data
    // and amplitude past n_smooth.  The wave reads only what it wrote, in
    // order, so no barrier is needed.
    __asm__ volatile("" ::: "memory");
    if (r0row != nullptr) {  // precomputed per cell: a copy
#pragma unroll 4
      for (int r = 0; r < kModeChunk / 64; ++r) {
        const int i = c0 + 64 * r + lane;
        wt[64 * r + lane] = make_uint2(r0row[i], __builtin_bit_cast(uint32_t, ampt[i].x));
      }
    } else {
#pragma unroll 2
    for (int r = 0; r < kModeChunk / 64; ++r) {
      const int i = c0 + 64 * r + lane;
      const SmoothRec sr = sm[i];
      double t = __builtin_fma(sr.kap[0], X[0], sr.th);
      t = __builtin_fma(sr.kap[1], X[1], t);
      if constexpr (D == 3) t = __builtin_fma(sr.kap[2], X[2], t);
      // the empty asm keeps the FP32 rounding step: without it the compiler
      // folds FP64 -> FP32 -> fp16 into a long software FP64 -> fp16 sequence
      float r0 = static_cast<float>(t - rint(t));
      __asm__ volatile("" : "+v"(r0));
      uint16_t rh, rl;
      split2h(r0, rh, rl);
      wt[64 * r + lane] = make_uint2(static_cast<uint32_t>(rh) | (static_cast<uint32_t>(rl) << 16),
                                     __builtin_bit_cast(uint32_t, ampt[i].x));
    }
    }
    __asm__ volatile("" ::: "memory");
    // tiles of 32 modes; B fragments one tile ahead of their MFMAs (reads at
    // most 63 modes past n_smooth, inside the zeroed padding); table records
    // of the next tile clamped to the chunk
    const int ntile = min(kModeChunk / 32, (n_smooth - c0 + 31) / 32);
    // per-mode B table: a wave-uniform tile base (scalar registers) plus this
    // lane's fixed offset, so the loop advances the base with scalar adds
    const uint4* bp = btpl32 + 2 * c0;
    const unsigned bo = 2u * lr + hl;
    auto fragment = [&](const uint4& bq, const uint2& wq) {
      uint4 b = bq;
      b.y |= wq.x & mu;
      return __builtin_bit_cast(h8, b);
    };
    uint4 bq = bp[bo];
    uint2 wq = wt[lr];
    float am = __builtin_bit_cast(float, wq.y);
    f16x t0, t1;
    {
      const h8 b = fragment(bq, wq);
      t0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0], b, zero, 0, 0, 0);
      t1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1], b, zero, 0, 0, 0);
    }
    // the table of the tile after the last is read and never used: its
    // entries 32 ntile + lr stay below 32 (kModeChunk / 32) + 32 <= kWaveLds / 8
    static_assert(32 * (kModeChunk / 32) + 32 <= kWaveLds / 8, "prefetch inside the table");
    bq = bp[bo + 64];
    wq = wt[32 + lr];
#pragma unroll 2
    for (int q = 0; q + 1 < ntile; ++q) {
      const h8 b = fragment(bq, wq);
      const float am_next = __builtin_bit_cast(float, wq.y);
      bq = bp[64 * (q + 2) + bo];
      wq = wt[32 * (q + 2) + lr];
#pragma unroll
      for (int v = 0; v < 16; ++v)
        acc[0][v] = __builtin_fmaf(am, __builtin_amdgcn_sinf(t0[v]), acc[0][v]);
      t0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0], b, zero, 0, 0, 0);
#pragma unroll
      for (int v = 0; v < 16; ++v)
        acc[1][v] = __builtin_fmaf(am, __builtin_amdgcn_sinf(t1[v]), acc[1][v]);
      t1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1], b, zero, 0, 0, 0);
      am = am_next;
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
      acc[0][v] = __builtin_fmaf(am, __builtin_amdgcn_sinf(t0[v]), acc[0][v]);
      acc[1][v] = __builtin_fmaf(am, __builtin_amdgcn_sinf(t1[v]), acc[1][v]);
    }
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int v = 0; v < 16; ++v) {
        tot[g][v] += acc[g][v];
        acc[g][v] = 0.0f;
      }
  }
  // cross-lane sums over the 32 mode columns; the table and the
  // sums share the wave's LDS region, the compiler barrier keeps the last
  // table reads before these writes
  __asm__ volatile("" ::: "memory");
  float* red = reinterpret_cast<float*>(lds[wv]);
#pragma unroll
  for (int g = 0; g < 2; ++g)
#pragma unroll
    for (int v = 0; v < 16; ++v)
      red[(32 * g + 8 * (v >> 2) + 4 * hl + (v & 3)) * kRed32Ld + lr] = tot[g][v];
  __asm__ volatile("" ::: "memory");
  // groups of four columns in FP32 (error below 2^-23 of their magnitudes,
  // under 1e-6 sigma in all), the eight group sums in FP64
  double sum = 0.0;
#pragma unroll
  for (int c = 0; c < 32; c += 4) {
    const float* rr = red + lane * kRed32Ld + c;
    sum += static_cast<double>((rr[0] + rr[1]) + (rr[2] + rr[3]));
  }
  const int p = __builtin_bit_cast(int, rf[lane * 4 + 3]);
  if (p < 0) return;
  if (n_rough > 0) {
    double x[1][D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[0][k] = pos[k * ld + p];
    double rt[1];
    sum_modes_scalar<D, 1>(rough, n_rough, x, rt);
    sum += rt[0];
  }
  store(out + p, sum * scale);
}

// FP64-phase sum over the rough modes for the D components of one point.
template <int D>
__device__ __forceinline__ void sum_modes_vec(const ModeRec* __restrict__ modes, int M,
                                              const double (&x)[D], double (&tot)[D]) {
#pragma unroll
  for (int c = 0; c < D; ++c) tot[c] = 0.0;
  for (int m0 = 0; m0 < M; m0 += kChunk) {
    const int m1 = min(m0 + kChunk, M);
    float acc[D];
#pragma unroll
    for (int c = 0; c < D; ++c) acc[c] = 0.0f;
    for (int m = m0; m < m1; ++m) {
      const ModeRec md = modes[m];
      double t = __builtin_fma(md.kap[0], x[0], md.th);
      t = __builtin_fma(md.kap[1], x[1], t);
      if constexpr (D > 2) t = __builtin_fma(md.kap[2], x[2], t);
      const float sv = sin_rev(t);
#pragma unroll
      for (int c = 0; c < D; ++c) acc[c] = __builtin_fmaf(md.amp[c], sv, acc[c]);
    }
#pragma unroll
    for (int c = 0; c < D; ++c) tot[c] += static_cast<double>(acc[c]);
  }
}

// Vector-valued form (incompressible fields): component c of point x is
// sum_m amp_c(m) sin(2 pi t_m(x)), one sine per point-mode shared by the D
// components.  The phases come from 16 x 16 x 32 bf16 products (three-part
// splits, 24 significant bits).  Wave w takes record group w (64 points,
// four MFMA row groups on every 16-mode B fragment), so each tile's B read
// and amplitude read serve 1024 point-modes; its FP32 accumulators (D x 16
// per lane) go into FP64 per point every kVecChunk modes through the wave's
// LDS rows.
//
// K slots 24 .. 26 (lane group kg = 3) carry r0: A holds 1 there and B the
// three bf16 parts of the cell's r0, read from the per-cell table that
// cell_slice writes, so the matrix instruction returns the whole local
// phase.  The amplitudes of kVecAmpChunk modes at a time sit in the block's
// LDS.  Without a per-cell table (too large) k_direct_local_vec_eval below
// evaluates r0 per wave pair and chunk instead.
constexpr int kVecChunk = 1024;     // modes per FP32 accumulation interval
constexpr int kVecEvalChunk = 512;  // modes per (r0, amplitudes) table of a wave pair (eval form)
constexpr int kVecAmpChunk = 1024;  // modes per amplitude table of a block (16 KB)
constexpr int kVecRedLd = 33;       // LDS row of the cross-lane sums (doubles)
static_assert(kVecChunk <= kModeChunk && kVecEvalChunk <= kModeChunk,
              "the mode tables are zero-padded by kModeChunk entries");
static_assert(kVecAmpChunk <= kModeChunk, "the amplitude table reads stay in the padding");
static_assert(kVecAmpChunk % kVecChunk == 0, "accumulation intervals split the table");
// B fragment reads run two tiles (32 modes) past the last tile issued, and the
// r0 lanes read 16 bytes (two entries) from an 8-byte entry
static_assert(kVecR0Pad >= 32 + 16, "r0 row padding covers the reads past the smooth modes");

// A fragment of the r0-in-K form: slots as a_fragment, plus ones in slots
// 24 .. 26 (lane group 3).
template <int D>
__device__ __forceinline__ bf8 a_fragment_r0(const float* rec, int kg, float sc) {
  if (kg == 3) {
    constexpr uint16_t kOne = 0x3F80;  // bf16 1.0
    const uint16_t v[8] = {kOne, kOne, kOne, 0, 0, 0, 0, 0};
    return __builtin_bit_cast(bf8, v);
  }
  return a_fragment<D>(rec, kg, sc);
}

constexpr int kVecSumLd = 65;  // LDS row (floats) of the vector kernel's cross-lane sums
template <int D, typename OutT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_direct_local_vec(const BinParams* __restrict__ prm, const int* __restrict__ wave_cell,
                     const float4* __restrict__ recs, const uint4* __restrict__ btpl,
                     const float4* __restrict__ ampt, const uint2* __restrict__ r0v,
                     int64_t r0ld, const ModeRec* __restrict__ rough,
                     const double* __restrict__ pos, int64_t ld, OutT* __restrict__ out,
                     int64_t out_ld, double scale) {
  __shared__ __attribute__((aligned(16))) f4 s_amp[kVecAmpChunk];
  __shared__ float s_red[4][16 * kVecSumLd];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = blockIdx.x * 4 + wv;
  const BinParams P = *prm;
  if (blockIdx.x * 4 >= P.nwaves) return;  // the whole block is past the list
  // a wave past the list (in the last block) runs the loops on the last
  // record group, for the block's barriers, and stores nothing
  const bool live = w < P.nwaves;
  const int n_smooth = P.n_smooth, n_rough = P.n_rough;
  const int rg = live ? w : P.nwaves - 1;
  const int kg = lane >> 4, j = lane & 15;
  const float* rf = reinterpret_cast<const float*>(recs + static_cast<int64_t>(rg) * kWavePts);
  bf8 a[4];
  // (a select chain, not P.h[kg]: a lane-dependent index would put P in scratch)
  const double hk = kg == 0 ? P.h[0] : (kg == 1 ? P.h[1] : P.h[2]);
  const float sck = kg < D ? static_cast<float>(ldexp(1.0, -half_cell_exp(hk))) : 0.0f;
#pragma unroll
  for (int g = 0; g < 4; ++g) a[g] = a_fragment_r0<D>(rf + (16 * g + j) * 4, kg, sck);
  // B fragment source of this lane, mode j of a tile: the mode template
  // (lane groups 0..2) or the cell's r0 parts (group 3; the 16-byte read
  // also takes the next entry, which meets zero slots of A)
  const char* bsrc =
      kg < 3 ? reinterpret_cast<const char*>(btpl + 4 * j + kg)
             : reinterpret_cast<const char*>(r0v + static_cast<int64_t>(wave_cell[rg]) * r0ld + j);
  const int64_t bstep = kg < 3 ? 16 * 4 * sizeof(uint4) : 16 * sizeof(uint2);  // per tile
  const f4 zero = {0.0f, 0.0f, 0.0f, 0.0f};
  float acc[D][4][4];
  double comp[D];
#pragma unroll
  for (int c = 0; c < D; ++c) {
    comp[c] = 0.0;
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[c][g][v] = 0.0f;
  }
  // acc[c][g][v] is slot 16 g + 4 (lane / 16) + v of the record group, mode
  // column j; lane l sums slot l's 16 columns in FP64
  float* red = s_red[wv];
  auto flush = [&]() {
#pragma unroll
    for (int c = 0; c < D; ++c) {
      __asm__ volatile("" ::: "memory");
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          red[j * kVecSumLd + 16 * g + 4 * kg + v] = acc[c][g][v];
          acc[c][g][v] = 0.0f;
        }
      __asm__ volatile("" ::: "memory");
      double sum = 0.0;
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) sum += static_cast<double>(red[jj * kVecSumLd + lane]);
      comp[c] += sum;
    }
    __asm__ volatile("" ::: "memory");
  };
  for (int a0 = 0; a0 < n_smooth; a0 += kVecAmpChunk) {
    __syncthreads();  // the previous amplitude table is read
    // zero amplitudes past n_smooth (the padding of ampt)
    for (int i = threadIdx.x; i < kVecAmpChunk; i += 256) {
      const float4 am = ampt[a0 + i];
      s_amp[i] = f4{am.x, am.y, am.z, 0.0f};
    }
    __syncthreads();
    const int a1 = min(a0 + kVecAmpChunk, n_smooth);
    for (int c0 = a0; c0 < a1; c0 += kVecChunk) {
      const int ntile = min(kVecChunk / 16, (a1 - c0 + 15) / 16);
      const char* bp = bsrc + static_cast<int64_t>(c0 / 16) * bstep;
      const f4* mp = s_amp + (c0 - a0) + j;
      uint4 ba = *reinterpret_cast<const uint4*>(bp);
      uint4 bb = *reinterpret_cast<const uint4*>(bp + bstep);
      bp += 2 * bstep;
      f4 ta[4], tb[4];
      f4 ama, amb;
      auto issue = [&](f4 (&t)[4], f4& am, uint4& bq) {
        am = *mp;
        mp += 16;
#pragma unroll
        for (int g = 0; g < 4; ++g)
          t[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[g], __builtin_bit_cast(bf8, bq), zero,
                                                         0, 0, 0);
        bq = *reinterpret_cast<const uint4*>(bp);
        bp += bstep;
      };
      auto consume = [&](const f4 (&t)[4], const f4& am) {
#pragma unroll
        for (int g = 0; g < 4; ++g)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float sv = __builtin_amdgcn_sinf(t[g][v]);
#pragma unroll
            for (int c = 0; c < D; ++c) acc[c][g][v] = __builtin_fmaf(am[c], sv, acc[c][g][v]);
          }
      };
      issue(ta, ama, ba);
      int q = 0;
      for (; q + 2 < ntile; q += 2) {
        issue(tb, amb, bb);
        consume(ta, ama);
        issue(ta, ama, ba);
        consume(tb, amb);
      }
      if (q + 1 < ntile) {
        issue(tb, amb, bb);
        consume(ta, ama);
        consume(tb, amb);
      } else {
        consume(ta, ama);
      }
      flush();
    }
  }
  const int p = live ? __builtin_bit_cast(int, rf[lane * 4 + 3]) : -1;
  if (p < 0) return;
  if (n_rough > 0) {
    double x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = pos[k * ld + p];
    double rt[D];
    sum_modes_vec<D>(rough, n_rough, x, rt);
#pragma unroll
    for (int c = 0; c < D; ++c) comp[c] += rt[c];
  }
#pragma unroll
  for (int c = 0; c < D; ++c) store(out + c * out_ld + p, comp[c] * scale);
}

// The vector form with r0 evaluated in the kernel: per wave pair (the two
// halves of one record group, 32 slots per wave, one cell) an LDS table of
// (r0, amplitudes) for kVecEvalChunk modes, r0 entering the product as the
// accumulator C.
template <int D, typename OutT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4)))
k_direct_local_vec_eval(const BinParams* __restrict__ prm, const double* __restrict__ centre,
                   const int* __restrict__ wave_cell, const float4* __restrict__ recs,
                   const uint4* __restrict__ btpl, const float4* __restrict__ ampt,
                   const SmoothRec* __restrict__ sm, const ModeRec* __restrict__ rough,
                   const double* __restrict__ pos, int64_t ld, OutT* __restrict__ out,
                   int64_t out_ld, double scale) {
  // per wave pair (the two halves of one record group, one cell): the
  // (r0, amplitude) table of one chunk of modes, each wave filling half of
  // it, and afterwards the pair's cross-lane sums
  constexpr int kPairLds = kVecEvalChunk * 32 > 2 * 16 * kVecRedLd * 8 ? kVecEvalChunk * 32
                                                                    : 2 * 16 * kVecRedLd * 8;
  __shared__ __attribute__((aligned(16))) char lds[2][kPairLds];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = blockIdx.x * 4 + wv;
  const BinParams P = *prm;
  if (blockIdx.x * 4 >= 2 * P.nwaves) return;  // the whole block is past the list
  // a pair past the list (in the last block) runs the loops on the last
  // record group, for the block's barriers, and stores nothing
  const bool live = w < 2 * P.nwaves;
  const int n_smooth = P.n_smooth, n_rough = P.n_rough;
  const int rg = live ? (w >> 1) : P.nwaves - 1;
  const int par = w & 1;
  double X[3];
  {
    const double* xc = centre + 4 * static_cast<int64_t>(wave_cell[rg]);
    X[0] = xc[0];
    X[1] = xc[1];
    X[2] = xc[2];
  }
  const int kg = lane >> 4, j = lane & 15;
  const float* rf =
      reinterpret_cast<const float*>(recs + static_cast<int64_t>(rg) * kWavePts + 32 * (w & 1));
  bf8 a[2];
  // (a select chain, not P.h[kg]: a lane-dependent index would put P in scratch)
  const double hk = kg == 0 ? P.h[0] : (kg == 1 ? P.h[1] : P.h[2]);
  const float sck = kg < D ? static_cast<float>(ldexp(1.0, -half_cell_exp(hk))) : 0.0f;
#pragma unroll
  for (int g = 0; g < 2; ++g) a[g] = a_fragment<D>(rf + (16 * g + j) * 4, kg, sck);
  f4* rq = reinterpret_cast<f4*>(lds[wv >> 1]);
  f4* ram = rq + kVecEvalChunk;
  float acc[D][2][4], tot[D][2][4];
#pragma unroll
  for (int c = 0; c < D; ++c)
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        acc[c][g][v] = 0.0f;
        tot[c][g][v] = 0.0f;
      }
  for (int c0 = 0; c0 < n_smooth; c0 += kVecEvalChunk) {
#pragma unroll 2
    for (int r = par; r < kVecEvalChunk / 64; r += 2) {  // this wave's half of the table
      const int i = c0 + 64 * r + lane;  // past n_smooth: zero padding
      const SmoothRec sr = sm[i];
      double t = __builtin_fma(sr.kap[0], X[0], sr.th);
      t = __builtin_fma(sr.kap[1], X[1], t);
      if constexpr (D == 3) t = __builtin_fma(sr.kap[2], X[2], t);
      const float r0 = static_cast<float>(t - rint(t));
      const float4 am = ampt[i];
      rq[64 * r + lane] = f4{r0, r0, r0, r0};
      ram[64 * r + lane] = f4{am.x, am.y, am.z, 0.0f};
    }
    __syncthreads();  // both halves of the table written
    const int ntile = min(kVecEvalChunk / 16, (n_smooth - c0 + 15) / 16);
    const uint4* bp = btpl + 4 * c0;
    const unsigned bo = 4u * j + kg;
    const f4* rp = rq + j;
    const f4* mp = ram + j;
    uint4 ba = bp[bo], bb = bp[bo + 64];
    bp += 128;
    f4 ta[2], tb[2];
    f4 ama, amb;
    auto issue = [&](f4 (&t)[2], f4& am, uint4& bq) {
      const f4 cq = *rp;
      am = *mp;
      rp += 16;
      mp += 16;
#pragma unroll
      for (int g = 0; g < 2; ++g)
        t[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[g], __builtin_bit_cast(bf8, bq), cq, 0, 0,
                                                       0);
      bq = bp[bo];
      bp += 64;
    };
    auto consume = [&](const f4 (&t)[2], const f4& am) {
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float sv = __builtin_amdgcn_sinf(t[g][v]);
#pragma unroll
          for (int c = 0; c < D; ++c) acc[c][g][v] = __builtin_fmaf(am[c], sv, acc[c][g][v]);
        }
    };
    issue(ta, ama, ba);
    int q = 0;
    for (; q + 2 < ntile; q += 2) {
      issue(tb, amb, bb);
      consume(ta, ama);
      issue(ta, ama, ba);
      consume(tb, amb);
    }
    if (q + 1 < ntile) {
      issue(tb, amb, bb);
      consume(ta, ama);
      consume(tb, amb);
    } else {
      consume(ta, ama);
    }
#pragma unroll
    for (int c = 0; c < D; ++c)
#pragma unroll
      for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          tot[c][g][v] += acc[c][g][v];
          acc[c][g][v] = 0.0f;
        }
    __syncthreads();  // the pair's table is read; the next chunk (or the sums) overwrites it
  }
  // cross-lane sums, one component at a time: tot[c][g][v] is slot
  // 16 g + 4 (lane / 16) + v of the half, mode column j
  const int p = live && lane < 32 ? __builtin_bit_cast(int, rf[lane * 4 + 3]) : -1;
  double comp[D];
  double* red = reinterpret_cast<double*>(lds[wv >> 1]) + par * 16 * kVecRedLd;
#pragma unroll
  for (int c = 0; c < D; ++c) {
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        red[j * kVecRedLd + 16 * g + 4 * kg + v] = static_cast<double>(tot[c][g][v]);
    double sum = 0.0;
    if (lane < 32)
#pragma unroll
      for (int jj = 0; jj < 16; ++jj) sum += red[jj * kVecRedLd + lane];
    comp[c] = sum;
    __asm__ volatile("" ::: "memory");
  }
  if (p < 0) return;
  if (n_rough > 0) {
    double x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = pos[k * ld + p];
    double rt[D];
    sum_modes_vec<D>(rough, n_rough, x, rt);
#pragma unroll
    for (int c = 0; c < D; ++c) comp[c] += rt[c];
  }
#pragma unroll
  for (int c = 0; c < D; ++c) store(out + c * out_ld + p, comp[c] * scale);
}

}  // namespace

size_t binned_work_bytes(int dim, int64_t M, int64_t N) {
  if (dim < 2 || dim > 3 || N <= 0) return 0;
  return bin_layout(M, N).bytes;
}

bool binned_supported(int dim, int64_t N) {
  // slot and point indices are 32-bit: max_waves * kWavePts < 2^31
  return (dim == 2 || dim == 3) && N > 0 && N < (static_cast<int64_t>(1) << 30);
}

hipError_t launch_direct_binned(int dim, int64_t M, int64_t N, const void* table,
                                const double* pos, int64_t pos_ld, void* out, int64_t out_ld,
                                bool vec, int out_dtype, double scale, void* bin_work,
                                hipStream_t s, hipEvent_t split, const PrepArgs* prep) {
  if (N <= 0) return prep ? launch_prep(prep->kind, prep->dim, prep->M, prep->cov, prep->z1,
                                        prep->z2, prep->sf, prep->table, s)
                          : hipSuccess;
  // the mode table's preparation rides on the bounding-box launch
  PrepArgs pa{};
  if (prep) pa = *prep;
  const unsigned box_grid = static_cast<unsigned>(
      kBoxBlocks + (static_cast<int64_t>(pa.ntiles) * 16 + 255) / 256);
  const BinLayout L = bin_layout(M, N);
  BinParams* prm = at<BinParams>(bin_work, L.params);
  double* box = at<double>(bin_work, L.box);
  double* centre = at<double>(bin_work, L.centre);
  int* bcount = at<int>(bin_work, L.bcount);
  int* counts = at<int>(bin_work, L.counts);
  int* offsets = at<int>(bin_work, L.offsets);
  int* wave_cell = at<int>(bin_work, L.wave_cell);
  float4* recs = at<float4>(bin_work, L.recs);

  uint4* btpl = at<uint4>(bin_work, L.btpl);
  uint4* btpl32 = at<uint4>(bin_work, L.btpl32);
  float4* ampt = at<float4>(bin_work, L.amp);
  SmoothRec* sm = at<SmoothRec>(bin_work, L.sm);
  ModeRec* rough = at<ModeRec>(bin_work, L.rough);
  uint32_t* r0w = L.r0ld > 0 || L.r0vld > 0 ? at<uint32_t>(bin_work, L.r0w) : nullptr;
  // the r0 rows cell_slice writes: the scalar or the vector form, if any
  uint32_t* r0rows = vec ? (L.r0vld > 0 ? r0w : nullptr) : (L.r0ld > 0 ? r0w : nullptr);
  const int64_t r0rld = vec ? L.r0vld : L.r0ld;
  const ModeRec* modes = static_cast<const ModeRec*>(table);
  const int nblk = static_cast<int>(L.nblk);
  const unsigned cb = static_cast<unsigned>((L.max_cells + kColCells - 1) / kColCells);
  // placement: the scatter's point blocks, then the cell rows' blocks (a
  // wave per cell at least, up to kWaveSlices while within kWavesMax waves)
  constexpr int kWpb = kSortThreads / 64;
  const unsigned pb = static_cast<unsigned>(
      nblk + (std::max<int64_t>(L.max_cells,
                                std::min<int64_t>(kWavesMax, L.max_cells * kWaveSlices)) +
              kWpb - 1) / kWpb);
  const unsigned mb = static_cast<unsigned>((L.max_waves + 3) / 4);
  const unsigned vb = static_cast<unsigned>((2 * L.max_waves + 3) / 4);
  const int Mi = static_cast<int>(M);
  const bool f32 = out_dtype == 1;
#define GSC_BIN(D)                                                                               \
  do {                                                                                           \
    k_box_partial<D><<<box_grid, 256, 0, s>>>(pos, pos_ld, N, box, pa);                          \
    k_bin_params<D><<<1, 256, 0, s>>>(box, N, L.max_cells, modes, Mi, prm, btpl, btpl32, ampt,   \
                                      sm, rough);                                                \
    k_bin_hist<D><<<nblk, kSortThreads, 0, s>>>(pos, pos_ld, N, L.blk_pts, prm, bcount);         \
    k_col_scan<<<cb, 256, 0, s>>>(prm, nblk, bcount, counts);                                    \
    k_bin_offsets<<<1, kSortThreads, 0, s>>>(prm, counts, offsets);                              \
    k_bin_scatter<D><<<pb, kSortThreads, 0, s>>>(pos, pos_ld, N, L.blk_pts, nblk, prm, bcount,   \
                                                 counts, offsets, recs, centre, wave_cell, sm,   \
                                                 r0rows, r0rld, vec);                            \
    if (split) (void)hipEventRecord(split, s);                                                   \
    if (vec && r0rows != nullptr && f32)                                                         \
      k_direct_local_vec<D, float><<<mb, 256, 0, s>>>(                                         \
          prm, wave_cell, recs, btpl, ampt, reinterpret_cast<const uint2*>(r0rows), r0rld, rough, \
          pos, pos_ld, static_cast<float*>(out), out_ld, scale);                                 \
    else if (vec && r0rows != nullptr)                                                           \
      k_direct_local_vec<D, double><<<mb, 256, 0, s>>>(                                        \
          prm, wave_cell, recs, btpl, ampt, reinterpret_cast<const uint2*>(r0rows), r0rld, rough, \
          pos, pos_ld, static_cast<double*>(out), out_ld, scale);                                \
    else if (vec && f32)                                                                         \
      k_direct_local_vec_eval<D, float><<<vb, 256, 0, s>>>(prm, centre, wave_cell, recs, btpl,   \
                                                           ampt, sm, rough, pos, pos_ld,         \
                                                           static_cast<float*>(out), out_ld,     \
                                                           scale);                               \
    else if (vec)                                                                                \
      k_direct_local_vec_eval<D, double><<<vb, 256, 0, s>>>(prm, centre, wave_cell, recs, btpl,  \
                                                            ampt, sm, rough, pos, pos_ld,        \
                                                            static_cast<double*>(out), out_ld,   \
                                                            scale);                              \
    else if (f32)                                                                                \
      k_direct_local32<D, float><<<mb, 256, 0, s>>>(prm, centre, wave_cell, recs, btpl32, ampt,  \
                                                    sm, r0w, L.r0ld, rough, pos, pos_ld,         \
                                                    static_cast<float*>(out), scale);            \
    else                                                                                         \
      k_direct_local32<D, double><<<mb, 256, 0, s>>>(prm, centre, wave_cell, recs, btpl32, ampt, \
                                                     sm, r0w, L.r0ld, rough, pos, pos_ld,        \
                                                     static_cast<double*>(out), scale);          \
  } while (0)
  if (dim == 2) GSC_BIN(2);
  else GSC_BIN(3);
#undef GSC_BIN
  return hipGetLastError();
}

}  // namespace gsc
