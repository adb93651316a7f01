// Factored summation on structured grids: host orchestration (workspace plan,
// launches) and the ensemble kernels that do not depend on the operand format.
// The contraction itself (per-axis tables, split-precision MFMA GEMM) is in
// gsc_gemm.h, instantiated per operand format in gemm_*.hip.
#include "gsc_device.h"
#include "gsc_factored.h"
#include "gsc_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace gsc {
namespace {

// ---------------------------------------------------------------------------
// Ensembles: many realisations (own modes, own grid) in one graph of launches.
// Every realisation has a descriptor; factored realisations contribute
// GEMM block tiles, direct ones blocks of kEnsPts points, to two flat work
// lists consumed by persistent kernels through atomic counters.
// ---------------------------------------------------------------------------
__global__ void k_ens_reset(int* __restrict__ counters, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) counters[i] = 0;
}

__global__ void __launch_bounds__(256)
k_ens_prep(const EnsReal* __restrict__ d, int count) {
  const EnsReal& r = d[blockIdx.y];
  const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m >= static_cast<int64_t>(r.ntiles) * 16) return;
  prep_mode(KIND_SCALAR, r.dim, r.M, m, r.cov, r.z1, r.z2, nullptr, r.table, r.a_tab, r.amp_tab);
}


constexpr int kEnsP = 8;              // points per thread in the direct ensemble kernel
constexpr int kEnsPts = 256 * kEnsP;  // points per work item
static_assert(kEnsPts == kEnsPoints, "direct work item size");

template <int D, typename OutT>
__global__ void __launch_bounds__(256)
k_ens_direct(const EnsReal* __restrict__ d, const int* __restrict__ dir,
             const int64_t* __restrict__ blk_start, int ndir, int* __restrict__ counter) {
  __shared__ int64_t item_s;
  const int64_t total = blk_start[ndir];
  for (;;) {
    if (threadIdx.x == 0) item_s = atomicAdd(counter, 1);
    __syncthreads();
    const int64_t t = item_s;
    __syncthreads();
    if (t >= total) break;
    const int i = find_item(blk_start, ndir, t);
    const EnsReal& r = d[dir[i]];
    // the descriptor's pointers are device (global) memory: their global
    // views let the compiler use global instead of flat accesses
    const auto ax0 = global_ptr(r.ax0), ax1 = global_ptr(r.ax1), ax2 = global_ptr(r.ax2);
    const int64_t base = (t - blk_start[i]) * kEnsPts + threadIdx.x;
    const int64_t N = r.npts;
    double x[kEnsP][D];
#pragma unroll
    for (int p = 0; p < kEnsP; ++p) {
      int64_t j = base + p * 256;
      j = j < N ? j : N - 1;
      if constexpr (D == 2) {
        x[p][0] = ax0[j / r.s1];
        x[p][1] = ax1[j % r.s1];
      } else {
        const int64_t i0 = j / (r.s1 * r.s2), rem = j - i0 * (r.s1 * r.s2);
        x[p][0] = ax0[i0];
        x[p][1] = ax1[rem / r.s2];
        x[p][2] = ax2[rem % r.s2];
      }
    }
    double tot[kEnsP];
    sum_modes_scalar<D, kEnsP>(global_ptr(r.table), static_cast<int>(r.M), x, tot);
#pragma unroll
    for (int p = 0; p < kEnsP; ++p) {
      const int64_t j = base + p * 256;
      if (j < N) *(global_ptr(static_cast<OutT*>(r.out)) + j) = static_cast<OutT>(tot[p] * r.scale);
    }
  }
}

template <int D>
__global__ void __launch_bounds__(256)
k_expand_grid(const double* __restrict__ ax0, const double* __restrict__ ax1,
              const double* __restrict__ ax2, int64_t s1, int64_t s2, int64_t N,
              double* __restrict__ pos) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (j >= N) return;
  if constexpr (D == 1) {
    pos[j] = ax0[j];
  } else if constexpr (D == 2) {
    pos[j] = ax0[j / s1];
    pos[N + j] = ax1[j % s1];
  } else {
    const int64_t i0 = j / (s1 * s2), r = j - i0 * (s1 * s2);
    pos[j] = ax0[i0];
    pos[N + j] = ax1[r / s2];
    pos[2 * N + j] = ax2[r % s2];
  }
}

}  // namespace

hipError_t launch_expand_grid(int dim, const int64_t* shape, const double* const* axes,
                              double* pos, hipStream_t s) {
  int64_t N = 1;
  for (int i = 0; i < dim; ++i) N *= shape[i];
  if (N <= 0) return hipSuccess;
  const int64_t s1 = dim > 1 ? shape[1] : 1, s2 = dim > 2 ? shape[2] : 1;
  const unsigned g = static_cast<unsigned>((N + 255) / 256);
  if (dim == 1) k_expand_grid<1><<<g, 256, 0, s>>>(axes[0], nullptr, nullptr, s1, s2, N, pos);
  else if (dim == 2) k_expand_grid<2><<<g, 256, 0, s>>>(axes[0], axes[1], nullptr, s1, s2, N, pos);
  else k_expand_grid<3><<<g, 256, 0, s>>>(axes[0], axes[1], axes[2], s1, s2, N, pos);
  return hipGetLastError();
}

FactoredPlan plan_factored(int dim, int64_t M, const int64_t* shape, int precision, int ncomp,
                           int tile) {
  FactoredPlan p{};
  p.dim = dim;
  p.ncomp = ncomp;
  p.brows = tile_block_rows(tile);
  const int64_t br = p.brows;
  p.ns = precision_splits(precision);
  p.bf16 = precision_is_bf16(precision);
  p.mp = static_cast<int>(((M > 0 ? M : 1) + 31) / 32 * 32);
  p.nch = p.mp / 16;
  p.kp = 2 * static_cast<int64_t>(p.mp);
  if (dim == 2) {
    // rows are padded to the block rows and columns to up to kMaxTileCols:
    // take axis 1 as the row (A) axis when that pads less; the GEMM then
    // stores the C-order output transposed (row stride 1, column stride nrow)
    auto work = [br](int64_t r, int64_t c) {
      return (r + br - 1) / br * br * ((c + kMaxTileCols - 1) / kMaxTileCols);
    };
    p.trans = work(shape[1], shape[0]) < work(shape[0], shape[1]) ? 1 : 0;
    p.na = 1;
    p.nrow = shape[p.trans ? 1 : 0];
    p.nb_out = 0;
    p.ncols = shape[p.trans ? 0 : 1];
  } else {
    p.na = shape[0];
    p.nrow = shape[1];
    p.nb_out = shape[1];
    p.ncols = shape[2];
  }
  p.arows = (p.nrow + br - 1) / br * br;
  p.ncp = (p.ncols + kMaxTileCols - 1) / kMaxTileCols * kMaxTileCols;
  p.ncpt = p.ncp * ncomp;
  p.row_tiles = (dim == 3 ? p.na : 1) * (p.arows / br);
  auto al = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
  size_t off = al(mode_table_bytes(3, M));
  p.off_s = off;
  if (dim == 3) off += al(sizeof(float2) * static_cast<size_t>(p.na) * p.mp);
  p.off_a = off;
  off += al(sizeof(uint32_t) * 16 * static_cast<size_t>(p.ns) * p.nch * p.arows);
  p.off_c = off;
  off += al(sizeof(float2) * static_cast<size_t>(p.ncpt) * p.mp);
  p.bytes = off;
  return p;
}

FactoredPlan plan_factored_product(const FactoredPlan& p, int tcols, const double* org,
                                   const double* step) {
  FactoredPlan q = p;
  q.prod = 1;
  q.tcols = tcols;
  q.col_tiles = (p.ncols + tcols - 1) / tcols;
  q.row_tiles = p.arows / p.brows;
  q.na = q.row_tiles * q.col_tiles;
  q.arows = p.brows;
  q.ncp = kMaxTileCols;
  q.ncpt = q.ncp;
  for (int i = 0; i < 2; ++i) {
    q.org[i] = org[i];
    q.step[i] = step[i];
  }
  auto al = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
  size_t off = p.off_s;  // after the mode table
  q.off_s = off;
  off += al(sizeof(float2) * static_cast<size_t>(q.na) * q.mp);
  q.off_a = off;
  off += al(sizeof(uint32_t) * 16 * static_cast<size_t>(q.ns) * q.nch * q.arows);
  q.off_c = off;
  off += al(sizeof(float2) * static_cast<size_t>(q.ncpt) * q.mp);
  q.bytes = off;
  return q;
}

// Dispatches a launcher template on the operand format of a plan.
template <template <typename, int> class F, typename... Args>
hipError_t by_format(bool bf16, int ns, Args&&... args) {
  if (bf16) {
    if (ns == 1) return F<__bf16, 1>::run(args...);
    if (ns == 2) return F<__bf16, 2>::run(args...);
    return F<__bf16, 3>::run(args...);
  }
  if (ns == 1) return F<_Float16, 1>::run(args...);
  if (ns == 2) return F<_Float16, 2>::run(args...);
  return F<_Float16, 3>::run(args...);
}
template <typename T, int NS> struct TablesF {
  static hipError_t run(const TableArgs& a, dim3 g, hipStream_t s) { return tables_launch<T, NS>(a, g, s); }
};
template <typename T, int NS> struct GemmF {
  static hipError_t run(const GemmArgs& a, dim3 g, hipStream_t s) {
    return factored_gemm_launch<T, NS>(a, g, s);
  }
};
template <typename T, int NS> struct EnsF {
  static hipError_t run(const EnsLaunch& e, int part, int accum, bool out_f64, hipStream_t s) {
    return ens_launch<T, NS>(e, part, accum, out_f64, s);
  }
};

hipError_t launch_factored(const FactoredPlan& p, int64_t M, const double* const* axes,
                           void* work, void* out, int out_dtype, double scale, int accum,
                           hipStream_t s) {
  if (p.row_tiles == 0 || p.ncols == 0) return hipSuccess;  // an empty grid
  char* w = static_cast<char*>(work);
  const ModeRec* modes = reinterpret_cast<const ModeRec*>(w);
  float2* S = reinterpret_cast<float2*>(w + p.off_s);
  uint32_t* A = reinterpret_cast<uint32_t*>(w + p.off_a);
  float2* C = reinterpret_cast<float2*>(w + p.off_c);
  const int Mi = static_cast<int>(M);
  // 3-D, and the 2-D product form, build B = S(block) C
  const bool kr = p.dim == 3 || p.prod;
  // tables (the mode table was written by launch_prep): one launch
  const int64_t nmax = std::max(std::max(kr ? p.na : 0, p.arows), p.ncpt);
  const dim3 tg(static_cast<unsigned>((nmax + 63) / 64), static_cast<unsigned>(p.nch), 3);
  const double* xa = axes[p.dim == 3 ? 1 : (p.trans ? 1 : 0)];
  const double* xc = axes[p.dim == 3 ? 2 : (p.trans ? 0 : 1)];
  const ProdGeom pg{p.prod, p.tcols, p.col_tiles, static_cast<int64_t>(p.brows),
                    p.org[0], p.org[1], p.step[0], p.step[1]};
  TableArgs ta{modes, Mi,      p.dim,   p.nch, axes[0], xa, xc, p.na, p.nrow,
               p.arows, p.ncols, p.ncp, S,     A,       C,  p.trans, p.ncomp, pg};
  hipError_t e = by_format<TablesF>(p.bf16, p.ns, ta, tg, s);
  if (e != hipSuccess) return e;
  // output: (ncomp, grid) dense; a transposed 2-D plan stores rows with
  // stride 1 and columns with stride nrow
  const int64_t npts = (p.dim == 3 ? p.na : 1) * p.nrow * p.ncols;
  GemmArgs ga{A,       p.arows, p.nch, S,     p.na,  C,  p.ncp, p.nrow, p.nb_out, out,
              out_dtype != 1, p.trans ? 1 : p.ncols, p.ncols, p.trans ? p.nrow : 1, scale, kr,
              accum, p.ncpt, npts, p.brows, p.prod != 0, p.col_tiles};
  const int64_t tile_cols = gemm_tile_cols(p.ns, accum);
  const int64_t tpc = (p.ncols + tile_cols - 1) / tile_cols;
  const dim3 grid(static_cast<unsigned>(p.row_tiles), static_cast<unsigned>(tpc * p.ncomp));
  return by_format<GemmF>(p.bf16, p.ns, ga, grid, s);
}

hipError_t launch_ensemble(const EnsLaunch* batches, int nbatch, int precision, int out_dtype,
                           int accum, hipStream_t s) {
  const EnsLaunch& e = batches[0];
  k_ens_reset<<<1, 256, 0, s>>>(e.counters, e.ncounters);
  if (e.max_tiles > 0)
    k_ens_prep<<<dim3(static_cast<unsigned>((static_cast<int64_t>(e.max_tiles) * 16 + 255) / 256),
                      static_cast<unsigned>(e.count)),
                 256, 0, s>>>(e.d, e.count);
  const bool f32 = out_dtype == 1;
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  // the batches one after the other on this stream: each one's tables, then
  // its GEMM, which finishes before the next batch rewrites the shared region
  const bool bf = precision_is_bf16(precision);
  const int ns = precision_splits(precision);
  for (int b = 0; b < nbatch; ++b) {
    if (batches[b].nfact == 0) continue;
    err = by_format<EnsF>(bf, ns, batches[b], 0, accum, !f32, s);
    if (err == hipSuccess) err = by_format<EnsF>(bf, ns, batches[b], 1, accum, !f32, s);
    if (err != hipSuccess) return err;
  }
  if (e.ndir > 0) {
    int* c1 = e.dir_counter;
    if (e.dim == 3) {
      if (f32) k_ens_direct<3, float><<<e.direct_blocks, 256, 0, s>>>(e.d, e.dir, e.blk_start, e.ndir, c1);
      else k_ens_direct<3, double><<<e.direct_blocks, 256, 0, s>>>(e.d, e.dir, e.blk_start, e.ndir, c1);
    } else {
      if (f32) k_ens_direct<2, float><<<e.direct_blocks, 256, 0, s>>>(e.d, e.dir, e.blk_start, e.ndir, c1);
      else k_ens_direct<2, double><<<e.direct_blocks, 256, 0, s>>>(e.d, e.dir, e.blk_start, e.ndir, c1);
    }
  }
  return hipGetLastError();
}

}  // namespace gsc