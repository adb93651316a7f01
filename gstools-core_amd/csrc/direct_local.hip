// Direct summation with local phases on the matrix cores (point lists, dim 2
// and 3): the "binned" direct variant.
//
// The phase of point x and mode m splits at the centre X_c of a spatial cell:
//   t = kap . x + th = (kap . X_c + th) + kap . xi,   xi = x - X_c.
// r0(c, m) = (kap . X_c + th) - rint(...) is evaluated in FP64 once per
// (wave, mode) and rounded to FP32; kap . xi is small (|xi| <= half a cell)
// and is computed, together with + r0, on the matrix cores for 16 points x
// 16 modes per v_mfma_f32_16x16x32_bf16:
//   D[i][j] = sum_k A[i][k] B[k][j]
// with xi and kap split into three bf16 parts each (x = h + m + l, 24
// significant bits) and K holding, per dimension, the six products of parts
// whose orders sum to at most 2 (h h, h m, m h, h l, m m, l h): 18 of the 32
// slots in 3-D.  r0 enters as the FP32 accumulator input C.  Every product of
// two bf16 parts is exact in FP32; the dropped products are below 2^-24 of
// the phase.  The B fragments depend on the mode only and are read from one
// table shared by all waves; a wave keeps (r0, amplitude) of its cell for
// one chunk of modes in LDS.  The VALU is left with one v_sin_f32 (argument in
// revolutions) and one FMA per point-mode (on gfx950 the matrix instructions
// do not run beside the VALU, so the 16 cycles of one MFMA per 256
// point-modes are the whole cost of the phases).  FP32 partial sums over
// 128-mode chunks are added into FP64 totals.
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

// K slots of the split phase product, per dimension d: slots 6 d + s hold
// (x part kXPart[s]) * (kap part kKPart[s]); the rest are zero.
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

constexpr int kCellPts = 512;       // target points per cell
constexpr int kMaxCells = 16384;    // cell counters of one block fill 64 KB of LDS
constexpr int kMaxSortBlocks = 1024;
constexpr int kColBlocks = 32;      // point blocks per group of the column scan
constexpr int kColGroups = kMaxSortBlocks / kColBlocks;
constexpr int kWavePts = 64;        // point slots per wave (4 MFMA groups of 16)
constexpr int kModeChunk = 256;     // modes per (r0, amplitude) table of a wave (16 tiles)
constexpr double kRoughRev = 1.0;   // bound on |kap . xi| (revolutions) of the FP32 phase
constexpr int kBoxBlocks = 1024;    // partial bounding boxes
constexpr int kRedLd = 65;          // LDS row of the final cross-lane sums (doubles)

struct alignas(16) BinParams {
  double lo[3];
  double h[3];
  double inv_h[3];
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
  size_t params, box, cell_of, bcount, gsum, counts, offsets, wave_cell, recs, btpl, amp, sm, rough,
      bytes;
};

BinLayout bin_layout(int64_t M, int64_t N) {
  BinLayout L{};
  L.max_cells = std::min<int64_t>(kMaxCells, N / (kCellPts / 2) + 1);
  L.max_waves = (N + kWavePts - 1) / kWavePts + L.max_cells;
  L.nblk = std::max<int64_t>(1, std::min<int64_t>(kMaxSortBlocks, (N + 4095) / 4096));
  L.blk_pts = (N + L.nblk - 1) / L.nblk;
  const size_t m = static_cast<size_t>(M > 0 ? M : 1);
  size_t off = 0;
  auto take = [&](size_t b) { const size_t o = off; off += al256(b); return o; };
  L.params = take(sizeof(BinParams));
  L.box = take(sizeof(double) * 6 * kBoxBlocks);
  L.cell_of = take(sizeof(int) * static_cast<size_t>(N));
  L.bcount = take(sizeof(int) * static_cast<size_t>(L.nblk * L.max_cells));
  L.gsum = take(sizeof(int) * static_cast<size_t>(kColGroups * L.max_cells));
  L.counts = take(sizeof(int) * static_cast<size_t>(L.max_cells));
  L.offsets = take(sizeof(int) * static_cast<size_t>(L.max_cells + 1));

  L.wave_cell = take(sizeof(int) * static_cast<size_t>(L.max_waves));
  L.recs = take(sizeof(float4) * static_cast<size_t>(L.max_waves) * kWavePts);
  L.btpl = take(sizeof(uint4) * 4 * (m + kModeChunk));
  L.amp = take(sizeof(float) * (m + kModeChunk));
  L.sm = take(sizeof(SmoothRec) * (m + kModeChunk));
  L.rough = take(sizeof(ModeRec) * m);
  L.bytes = off;
  return L;
}

template <typename T>
T* at(void* base, size_t off) {
  return reinterpret_cast<T*>(static_cast<char*>(base) + off);
}

// ---- pre-pass: bounding box ------------------------------------------------
template <int D>
__global__ void __launch_bounds__(256)
k_box_partial(const double* __restrict__ pos, int64_t ld, int64_t N, double* __restrict__ box) {
  __shared__ double red[2 * D][256];
  double lo[D], hi[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    lo[k] = INFINITY;
    hi[k] = -INFINITY;
  }
  for (int64_t p = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; p < N;
       p += static_cast<int64_t>(gridDim.x) * 256) {
#pragma unroll
    for (int k = 0; k < D; ++k) {
      const double x = pos[k * ld + p];
      if (isfinite(x)) {
        lo[k] = fmin(lo[k], x);
        hi[k] = fmax(hi[k], x);
      }
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

// Cell grid from the bounding box: about N / kCellPts near-cubic cells over
// the dimensions with a non-zero extent, at most max_cells in all.
template <int D>
__global__ void __launch_bounds__(256)
k_bin_params(const double* __restrict__ box, int64_t N, int64_t max_cells,
             BinParams* __restrict__ prm) {
  __shared__ double red[2 * D][256];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double a = INFINITY, b = -INFINITY;
    for (int i = threadIdx.x; i < kBoxBlocks; i += 256) {
      a = fmin(a, box[i * 6 + 2 * k]);
      b = fmax(b, box[i * 6 + 2 * k + 1]);
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
  if (threadIdx.x != 0) return;
  double lo[3] = {0.0, 0.0, 0.0}, ext[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < D; ++k) {
    double a = red[2 * k][0], b = red[2 * k + 1][0];
    if (!(a <= b)) a = b = 0.0;  // no finite coordinate
    lo[k] = a;
    ext[k] = b - a;
  }
  const double target = fmax(1.0, static_cast<double>(N) / kCellPts);
  double vol = 1.0;
  int active = 0;
  for (int k = 0; k < D; ++k)
    if (ext[k] > 0.0) {
      vol *= ext[k];
      ++active;
    }
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
    P.h[k] = ext[k] > 0.0 ? ext[k] / n[k] : 1.0;
    P.inv_h[k] = ext[k] > 0.0 ? 1.0 / P.h[k] : 0.0;
  }
  P.F = f[0] * f[1] * f[2];
  P.ncoarse = C[0] * C[1] * C[2];
  P.ncells = P.ncoarse * P.F;
  *prm = P;
}

// Number of the cell of point x: coarse cell (c0 C1 + c1) C2 + c2, then the
// cell within it (l0 f1 + l1) f2 + l2, where cell index i_k = c_k f_k + l_k.
template <int D>
__device__ __forceinline__ int cell_of_point(const BinParams& P, const double* x) {
  int cc = 0, l = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) {
    const double fr = (x[k] - P.lo[k]) * P.inv_h[k];
    int ik = 0;  // NaN and x < lo land in cell 0, x >= hi in the last cell
    if (fr >= 0.0) ik = fr < static_cast<double>(P.n[k]) ? static_cast<int>(fr) : P.n[k] - 1;
    cc = cc * P.C[k] + ik / P.f[k];
    l = l * P.f[k] + ik % P.f[k];
  }
  return cc * P.F + l;
}

// Point block b (points [b blk_pts, (b + 1) blk_pts)) counts its points per
// cell in LDS; bcount[b][c] receives the counts.
template <int D>
__global__ void __launch_bounds__(256)
k_bin_hist(const double* __restrict__ pos, int64_t ld, int64_t N, int64_t blk_pts,
           const BinParams* __restrict__ prm, int* __restrict__ cell_of,
           int* __restrict__ bcount) {
  __shared__ int hist[kMaxCells];
  const BinParams P = *prm;
  for (int c = threadIdx.x; c < P.ncells; c += 256) hist[c] = 0;
  __syncthreads();
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * blk_pts;
  const int64_t p1 = min(p0 + blk_pts, N);
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    double x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = pos[k * ld + p];
    const int c = cell_of_point<D>(P, x);
    cell_of[p] = c;
    atomicAdd(&hist[c], 1);
  }
  __syncthreads();
  int* row = bcount + static_cast<int64_t>(blockIdx.x) * P.ncells;
  for (int c = threadIdx.x; c < P.ncells; c += 256) row[c] = hist[c];
}

// Exclusive scan of every column of bcount over the point blocks, in three
// passes parallel over (cell, group of kColBlocks blocks):
//   sums of each group, scan of the group sums per cell (and the cell total),
//   rescan of each group from its offset.
__global__ void __launch_bounds__(256)
k_col_sum(const BinParams* __restrict__ prm, int nblk, const int* __restrict__ bcount,
          int* __restrict__ gsum) {
  const int ncells = prm->ncells;
  const int c = blockIdx.x * 256 + threadIdx.x, g = blockIdx.y;
  if (c >= ncells) return;
  int run = 0;
  for (int b = g * kColBlocks; b < min(nblk, (g + 1) * kColBlocks); ++b)
    run += bcount[static_cast<int64_t>(b) * ncells + c];
  gsum[g * ncells + c] = run;
}

__global__ void __launch_bounds__(256)
k_col_groups(const BinParams* __restrict__ prm, int ngroups, int* __restrict__ gsum,
             int* __restrict__ counts) {
  const int ncells = prm->ncells;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= ncells) return;
  int run = 0;
  for (int g = 0; g < ngroups; ++g) {
    const int v = gsum[g * ncells + c];
    gsum[g * ncells + c] = run;
    run += v;
  }
  counts[c] = run;
}

__global__ void __launch_bounds__(256)
k_col_fix(const BinParams* __restrict__ prm, int nblk, int* __restrict__ bcount,
          const int* __restrict__ gsum) {
  const int ncells = prm->ncells;
  const int c = blockIdx.x * 256 + threadIdx.x, g = blockIdx.y;
  if (c >= ncells) return;
  int run = gsum[g * ncells + c];
  for (int b = g * kColBlocks; b < min(nblk, (g + 1) * kColBlocks); ++b) {
    int* e = bcount + static_cast<int64_t>(b) * ncells + c;
    const int v = *e;
    *e = run;
    run += v;
  }
}

// One block: cell c gets ceil(count / kWavePts) waves starting at slot
// offsets[c]; offsets[ncells] and the wave count close the list.
__global__ void __launch_bounds__(1024)
k_bin_offsets(BinParams* __restrict__ prm, const int* __restrict__ counts,
              int* __restrict__ offsets) {
  __shared__ int part[1024];
  __shared__ int carry_s;
  const int ncells = prm->ncells;
  if (threadIdx.x == 0) carry_s = 0;
  __syncthreads();
  for (int base = 0; base < ncells; base += 1024) {
    const int c = base + threadIdx.x;
    const int waves = c < ncells ? (counts[c] + kWavePts - 1) / kWavePts : 0;
    part[threadIdx.x] = waves;
    __syncthreads();
    for (int st = 1; st < 1024; st <<= 1) {  // inclusive scan (Hillis-Steele)
      const int v = threadIdx.x >= st ? part[threadIdx.x - st] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    const int carry = carry_s;
    if (c < ncells) offsets[c] = (carry + part[threadIdx.x] - waves) * kWavePts;
    __syncthreads();
    if (threadIdx.x == 1023) carry_s = carry + part[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    offsets[ncells] = carry_s * kWavePts;
    prm->nwaves = carry_s;
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

// Thread per cell: records its waves in the wave -> cell map and marks the
// slots past its points as padding.
__global__ void __launch_bounds__(256)
k_bin_waves(const BinParams* __restrict__ prm, const int* __restrict__ counts,
            const int* __restrict__ offsets, int* __restrict__ wave_cell,
            float4* __restrict__ recs) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= prm->ncells) return;
  const int s0 = offsets[e], s1 = offsets[e + 1];
  for (int s = s0; s < s1; s += kWavePts) wave_cell[s / kWavePts] = e;
  for (int s = s0 + counts[e]; s < s1; ++s) recs[s] = make_rec(0.0f, 0.0f, 0.0f, -1);
}

// Point block b again: ranks its points within each cell with LDS atomics
// (starting from the block's base in the cell) and writes their records
// (local coordinates, index) to their slots.
template <int D>
__global__ void __launch_bounds__(256)
k_bin_scatter(const double* __restrict__ pos, int64_t ld, int64_t N, int64_t blk_pts,
              const BinParams* __restrict__ prm, const int* __restrict__ cell_of,
              const int* __restrict__ bcount, const int* __restrict__ offsets,
              float4* __restrict__ recs) {
  __shared__ int rank[kMaxCells];
  const BinParams P = *prm;
  const int* base = bcount + static_cast<int64_t>(blockIdx.x) * P.ncells;
  for (int e = threadIdx.x; e < P.ncells; e += 256) rank[e] = offsets[e] + base[e];
  __syncthreads();
  const int64_t p0 = static_cast<int64_t>(blockIdx.x) * blk_pts;
  const int64_t p1 = min(p0 + blk_pts, N);
  for (int64_t p = p0 + threadIdx.x; p < p1; p += 256) {
    const int e = cell_of[p];
    double X[3];
    cell_centre<D>(P, e, X);
    float xi[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int k = 0; k < D; ++k) xi[k] = static_cast<float>(pos[k * ld + p] - X[k]);
    recs[atomicAdd(&rank[e], 1)] = make_rec(xi[0], xi[1], xi[2], static_cast<int>(p));
  }
}

// One block: modes in order, split into smooth (FP32 local phase) and rough
// (FP64 phase) lists by a block scan, so both keep the original mode order.
template <int D>
__global__ void __launch_bounds__(256)
k_mode_split(const ModeRec* __restrict__ modes, int M, BinParams* __restrict__ prm,
             uint4* __restrict__ btpl, float* __restrict__ amp, SmoothRec* __restrict__ sm,
             ModeRec* __restrict__ rough) {
  __shared__ int part[256];
  __shared__ int ns_s, nr_s;
  const BinParams P = *prm;
  if (threadIdx.x == 0) ns_s = nr_s = 0;
  __syncthreads();
  for (int base = 0; base < M; base += 256) {
    const int m = base + threadIdx.x;
    ModeRec r{};
    bool smooth = false;
    if (m < M) {
      r = modes[m];
      double b = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) b += fabs(r.kap[k]) * 0.5 * P.h[k];
      smooth = b <= kRoughRev && isfinite(r.th) && isfinite(static_cast<double>(r.amp[0]));
    }
    part[threadIdx.x] = smooth ? 1 : 0;
    __syncthreads();
    for (int s = 1; s < 256; s <<= 1) {
      const int v = threadIdx.x >= s ? part[threadIdx.x - s] : 0;
      __syncthreads();
      part[threadIdx.x] += v;
      __syncthreads();
    }
    const int ns0 = ns_s, nr0 = nr_s;
    const int before = part[threadIdx.x] - (smooth ? 1 : 0);  // smooth modes before m in block
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
          split3(static_cast<float>(r.kap[d]), kp);
#pragma unroll
          for (int t = 0; t < 6; ++t) {
            const int k = 6 * d + t;
            w[k >> 1] |= static_cast<uint32_t>(kp[kKPart[t]]) << (16 * (k & 1));
          }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
          btpl[4 * i + q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
        amp[i] = r.amp[0];
        sm[i] = SmoothRec{{r.kap[0], r.kap[1], r.kap[2]}, r.th};
      } else {
        rough[nr0 + (threadIdx.x - before)] = r;
      }
    }
    __syncthreads();
    if (threadIdx.x == 255) {
      const int blk = min(256, M - base);
      ns_s = ns0 + part[255];
      nr_s = nr0 + blk - part[255];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    prm->n_smooth = ns_s;
    prm->n_rough = nr_s;
  }
  // the main kernel reads B fragments of up to kModeChunk modes past the
  // last smooth one (with zero amplitude): keep them finite
  for (int i = threadIdx.x; i < 4 * kModeChunk; i += 256)
    btpl[4 * ns_s + i] = make_uint4(0u, 0u, 0u, 0u);
}

// ---- main kernel -------------------------------------------------------------
// One wave per kWavePts slots of one cell; four independent waves per block,
// each with its own LDS region (no block barriers).
template <int D, typename OutT>
__global__ void __launch_bounds__(256)
k_direct_local(const BinParams* __restrict__ prm, const int* __restrict__ wave_cell,
               const float4* __restrict__ recs, const uint4* __restrict__ btpl,
               const float* __restrict__ ampt, const SmoothRec* __restrict__ sm,
               const ModeRec* __restrict__ rough, const double* __restrict__ pos, int64_t ld,
               OutT* __restrict__ out, double scale) {
  // per wave: (r0, amplitude) of one chunk of modes, and afterwards the
  // 16 x 64 partial sums (row stride kRedLd doubles)
  constexpr int kWaveLds = kModeChunk * 8 > 16 * kRedLd * 8 ? kModeChunk * 8 : 16 * kRedLd * 8;
  __shared__ __attribute__((aligned(16))) char lds[4][kWaveLds];
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int w = blockIdx.x * 4 + wv;
  const BinParams P = *prm;
  if (w >= P.nwaves) return;
  const int n_smooth = P.n_smooth, n_rough = P.n_rough;
  double X[3];
  cell_centre<D>(P, wave_cell[w], X);
  // A operands: lane l holds slots 8 (l / 16) .. + 7 of point l % 16 of group g
  const int kg = lane >> 4, j = lane & 15;
  const float* rf = reinterpret_cast<const float*>(recs + static_cast<int64_t>(w) * kWavePts);
  bf8 a[4];
#pragma unroll
  for (int g = 0; g < 4; ++g) {
    uint16_t xp[3][3];
#pragma unroll
    for (int d = 0; d < 3; ++d) split3(d < D ? rf[(16 * g + j) * 4 + d] : 0.0f, xp[d]);
    uint16_t sl[32];  // all 32 slots of the point, then the lane's eight
#pragma unroll
    for (int k = 0; k < 32; ++k) {
      sl[k] = k < 6 * D ? xp[k / 6][kXPart[k % 6]] : 0;
    }
    uint16_t v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e)
      v[e] = kg == 0 ? sl[e] : (kg == 1 ? sl[8 + e] : (kg == 2 ? sl[16 + e] : sl[24 + e]));
    a[g] = __builtin_bit_cast(bf8, v);
  }
  float2* ra = reinterpret_cast<float2*>(lds[wv]);
  const uint4* bfrag = btpl + kg;  // lane's k-range of mode j: btpl[4 m + kg]
  float acc[4][4];
  double tot[4][4];
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      acc[g][v] = 0.0f;
      tot[g][v] = 0.0;
    }
  for (int c0 = 0; c0 < n_smooth; c0 += kModeChunk) {
    // (r0, amplitude) of modes c0 .. c0 + kModeChunk - 1 for this cell: the
    // FP64 phase at the cell centre, reduced exactly (zeros past n_smooth)
#pragma unroll
    for (int r = 0; r < kModeChunk / 64; ++r) {
      const int i = c0 + 64 * r + lane;
      float2 e = make_float2(0.0f, 0.0f);
      if (i < n_smooth) {
        const SmoothRec sr = sm[i];
        double t = __builtin_fma(sr.kap[0], X[0], sr.th);
        t = __builtin_fma(sr.kap[1], X[1], t);
        if constexpr (D == 3) t = __builtin_fma(sr.kap[2], X[2], t);
        e = make_float2(static_cast<float>(t - rint(t)), ampt[i]);
      }
      ra[64 * r + lane] = e;
    }
    // software pipeline: the phases of tile q + 1 are issued before the
    // sines of tile q, so the MFMA results are ready when they are read; the
    // loop is unrolled by two so the two result sets keep their registers
    constexpr int kTiles = kModeChunk / 16;
    const int ntile = min(kTiles, (n_smooth - c0 + 15) / 16);
    const int mmax = n_smooth + kModeChunk - 1;  // the B table is padded by a chunk
    f4 ta[4], tb[4];
    float ama, amb;
    // B fragments are requested two tiles ahead of their MFMAs
    auto bload = [&](int q) { return bfrag[4 * min(c0 + 16 * q + j, mmax)]; };
    uint4 ba = bload(0), bb = bload(1);
    auto issue = [&](f4 (&t)[4], float& am, uint4& bq, int q) {
      const float2 e = ra[16 * min(q, kTiles - 1) + j];
      am = e.y;
#pragma unroll
      for (int g = 0; g < 4; ++g)
        t[g] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[g], __builtin_bit_cast(bf8, bq),
                                                       f4{e.x, e.x, e.x, e.x}, 0, 0, 0);
      bq = bload(q + 2);
    };
    auto consume = [&](const f4 (&t)[4], float am) {
#pragma unroll
      for (int g = 0; g < 4; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          acc[g][v] = __builtin_fmaf(am, __builtin_amdgcn_sinf(t[g][v]), acc[g][v]);
    };
    issue(ta, ama, ba, 0);
    for (int q = 0; q < ntile; q += 2) {
      issue(tb, amb, bb, q + 1);
      consume(ta, ama);
      if (q + 1 >= ntile) break;
      issue(ta, ama, ba, q + 2);
      consume(tb, amb);
    }
#pragma unroll
    for (int g = 0; g < 4; ++g)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        tot[g][v] += static_cast<double>(acc[g][v]);
        acc[g][v] = 0.0f;
      }
  }
  // cross-lane sum: tot[g][v] is slot 16 g + 4 (lane / 16) + v, mode column j.
  // The table and the sums share the wave's LDS region under different
  // types; the compiler barrier keeps the last table reads before these writes.
  __asm__ volatile("" ::: "memory");
  double* red = reinterpret_cast<double*>(lds[wv]);
#pragma unroll
  for (int g = 0; g < 4; ++g)
#pragma unroll
    for (int v = 0; v < 4; ++v) red[j * kRedLd + 16 * g + 4 * kg + v] = tot[g][v];
  double sum = 0.0;
#pragma unroll
  for (int jj = 0; jj < 16; ++jj) sum += red[jj * kRedLd + lane];
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
                                const double* pos, int64_t pos_ld, void* out, int out_dtype,
                                double scale, void* bin_work, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  const BinLayout L = bin_layout(M, N);
  BinParams* prm = at<BinParams>(bin_work, L.params);
  double* box = at<double>(bin_work, L.box);
  int* cell_of = at<int>(bin_work, L.cell_of);
  int* bcount = at<int>(bin_work, L.bcount);
  int* gsum = at<int>(bin_work, L.gsum);
  int* counts = at<int>(bin_work, L.counts);
  int* offsets = at<int>(bin_work, L.offsets);
  int* wave_cell = at<int>(bin_work, L.wave_cell);
  float4* recs = at<float4>(bin_work, L.recs);

  uint4* btpl = at<uint4>(bin_work, L.btpl);
  float* ampt = at<float>(bin_work, L.amp);
  SmoothRec* sm = at<SmoothRec>(bin_work, L.sm);
  ModeRec* rough = at<ModeRec>(bin_work, L.rough);
  const ModeRec* modes = static_cast<const ModeRec*>(table);
  const int nblk = static_cast<int>(L.nblk);
  const int ngroups = (nblk + kColBlocks - 1) / kColBlocks;
  const unsigned cb = static_cast<unsigned>((L.max_cells + 255) / 256);
  const unsigned mb = static_cast<unsigned>((L.max_waves + 3) / 4);
  const int Mi = static_cast<int>(M);
  const bool f32 = out_dtype == 1;
#define GSC_BIN(D)                                                                               \
  do {                                                                                           \
    k_box_partial<D><<<kBoxBlocks, 256, 0, s>>>(pos, pos_ld, N, box);                            \
    k_bin_params<D><<<1, 256, 0, s>>>(box, N, L.max_cells, prm);                                 \
    k_bin_hist<D><<<nblk, 256, 0, s>>>(pos, pos_ld, N, L.blk_pts, prm, cell_of, bcount);         \
    k_col_sum<<<dim3(cb, ngroups), 256, 0, s>>>(prm, nblk, bcount, gsum);                        \
    k_col_groups<<<cb, 256, 0, s>>>(prm, ngroups, gsum, counts);                                 \
    k_col_fix<<<dim3(cb, ngroups), 256, 0, s>>>(prm, nblk, bcount, gsum);                        \
    k_bin_offsets<<<1, 1024, 0, s>>>(prm, counts, offsets);                                      \
    k_bin_waves<<<cb, 256, 0, s>>>(prm, counts, offsets, wave_cell, recs);                       \
    k_bin_scatter<D><<<nblk, 256, 0, s>>>(pos, pos_ld, N, L.blk_pts, prm, cell_of, bcount,       \
                                          offsets, recs);                                        \
    k_mode_split<D><<<1, 256, 0, s>>>(modes, Mi, prm, btpl, ampt, sm, rough);                    \
    if (f32)                                                                                     \
      k_direct_local<D, float><<<mb, 256, 0, s>>>(prm, wave_cell, recs, btpl, ampt, sm, rough,   \
                                                  pos, pos_ld, static_cast<float*>(out), scale); \
    else                                                                                         \
      k_direct_local<D, double><<<mb, 256, 0, s>>>(prm, wave_cell, recs, btpl, ampt, sm, rough,  \
                                                   pos, pos_ld, static_cast<double*>(out),       \
                                                   scale);                                       \
  } while (0)
  if (dim == 2) GSC_BIN(2);
  else GSC_BIN(3);
#undef GSC_BIN
  return hipGetLastError();
}

}  // namespace gsc
