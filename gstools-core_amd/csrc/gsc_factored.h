// Structured-grid (factored) path: workspace plan and launchers.
#pragma once

#include <hip/hip_runtime.h>

#include "gsc_internal.h"

#include <cstddef>
#include <cstdint>

namespace gsc {

// Split-precision formats, numbered as gsc_precision in the public header.
inline int precision_splits(int prec) {
  switch (prec) {
    case 3: return 3;          // bf16x6
    case 4: case 5: return 1;  // fp16x1, bf16x1
    default: return 2;         // auto (fp16x3), fp16x3, bf16x3
  }
}
inline bool precision_is_bf16(int prec) { return prec == 2 || prec == 3 || prec == 5; }

// Operand roles.  The GEMM rows run over the A axis (axis 0 in 2-D, axis 1
// in 3-D), its columns over the last axis; K = 2M interleaves the real and
// imaginary parts of every mode.  In 3-D a block works on one index a of
// axis 0 and multiplies that axis' factor into the B operand.
//   S: axis 0 of a 3-D grid (amplitude and theta folded in): na x mp float2
//   A: the A axis, split into ns parts: ns x nch x arows x 16 packed pairs
//      (2-D: amplitude and theta folded in); rows >= the axis length are zero
//   C: last axis, ncpt x mp float2; columns >= ncols of each component are
//      zero.  The incompressible field (ncomp = dim) has one block of ncp
//      columns per component i, its entries multiplied by the projector
//      p_i(k_m) = delta_i0 - k_i k_0 / |k|^2, so the GEMM's columns run over
//      (component, last-axis index) and its output over (component, grid point).
// Workspace layout (bytes from the start of the caller's workspace):
//   [0, off_s)       ModeRec table (M records)
//   [off_s, off_a)   S (3-D only)
//   [off_a, off_c)   A
//   [off_c, bytes)   C
// (all chunk-major: the 16 modes of one row are contiguous, rows of one
// chunk follow each other).
// GEMM block rows per tile shape (gsc_opts.tile): 256 (4 waves x 64 rows,
// one wave per SIMD) or 128 (4 waves x 32 rows, two blocks per CU).
constexpr int kBlockRowsTall = 256;
constexpr int kBlockRowsShort = 128;
inline int tile_block_rows(int tile) { return tile == 2 ? kBlockRowsShort : kBlockRowsTall; }
constexpr int kMaxTileCols = 128;
struct FactoredPlan {
  int dim;
  int ns;
  bool bf16;
  int mp;            // modes padded to a multiple of 32 (an even number of 16-mode chunks)
  int nch;           // mp / 16
  int64_t kp;        // 2 * mp
  int64_t na;        // 3-D: axis 0 (one block row band per index); 2-D: 1
  int64_t nrow;      // length of the A axis
  int brows;         // GEMM block rows (kBlockRowsTall or kBlockRowsShort)
  int64_t arows;     // nrow padded to brows
  int64_t nb_out;    // 3-D: output rows per axis-0 index (= nrow); 2-D: 0
  int64_t ncols;     // last-axis length
  int64_t ncp;       // ncols padded to kMaxTileCols (one field component)
  int ncomp;         // field components: 1, or dim for the incompressible field
  int64_t ncpt;      // ncomp * ncp: columns of the C table and of the GEMM
  int64_t row_tiles; // 3-D: na * arows / brows; 2-D: arows / brows
  int trans;         // 2-D with axis 1 as the A axis (it pads less); the output is stored transposed
  size_t off_s, off_a, off_c, bytes;
  // 2-D product form (uniform axes, see plan_factored_product): S over the
  // tiles (na = row_tiles col_tiles), A over the brows row offsets of a tile,
  // C over the kMaxTileCols column offsets; org / step of the row (A) axis
  // and the column axis
  int prod;
  int tcols;          // GEMM tile columns
  int64_t col_tiles;
  double org[2], step[2];
};

// Columns of a GEMM block tile for an operand format and accumulation mode:
// the wave keeps 2 x tile/32 accumulator blocks, plus a second set for
// the two-level modes, within its register file.
constexpr int gemm_tile_cols(int ns, int accum) {
  if (accum == 0) return 32;  // FP64 second set
  if (ns == 3) return 64;
  return 128;
}

FactoredPlan plan_factored(int dim, int64_t M, const int64_t* shape, int precision,
                           int ncomp = 1, int tile = 0);
// The product form of a 2-D scalar plan p whose row (A) axis is org[0] +
// i step[0] and column axis org[1] + j step[1] (phase of point (r, c) of tile
// (rh, ch) = tile corner part + row offset part + column offset part): S per
// tile with amplitude and theta folded in, A and C one tile each, the GEMM as
// in 3-D (B = S(tile) C).  Its workspace follows the same mode table.
FactoredPlan plan_factored_product(const FactoredPlan& p, int tcols, const double* org,
                                   const double* step);

// One realisation of an ensemble (device-resident descriptor).
struct EnsReal {
  const double* cov;  // (dim, M) row-major
  const double* z1;
  const double* z2;
  const double* ax0;
  const double* ax1;
  const double* ax2;
  void* out;          // dense C-order grid
  ModeRec* table;     // mode records, then the matrix-core tile tables
  double* a_tab;
  float* amp_tab;
  float2* S;          // factored tables (roles as in FactoredPlan)
  uint32_t* A;
  float2* C;
  int64_t M;
  int ntiles;         // 16-mode tiles of the prep kernel
  int dim;
  int path;           // GSC_PATH_DIRECT or GSC_PATH_FACTORED
  int mp;             // factored: modes padded to 32
  int trans;          // factored 2-D: rows are y and columns x (the padding is smaller)
  // Product form (factored 2-D on uniform axes): row r = brows rh + a and
  // column c = tcols ch + b of tile (rh, ch) have the phase
  //   th + k_r (org_r + brows rh step_r) + k_c (org_c + tcols ch step_c)
  //      + k_r a step_r + k_c b step_c,
  // so S holds the first line per tile (na = row_tiles col_tiles rows,
  // amplitude folded in), A the row offsets a < brows and C the column
  // offsets b < ncp (one tile each), and the GEMM runs every tile as a 3-D
  // block (B = S(tile) C).  The tables shrink from the whole grid to a tile.
  int prod;
  int tcols;          // product form: GEMM tile columns
  double org[2], step[2];  // product form: origin and spacing of the row and column axes
  int64_t na, nrow, arows, nb_out, ncols, ncp, row_tiles, col_tiles;  // factored geometry
  int64_t s0, s1, s2, npts;                     // grid shape
  double scale;
};

// Launch sequence of an ensemble (all pointers device-resident).  fact / dir
// list the realisations on each path, task_start / blk_start their prefix
// sums of table tasks and direct work items.  The factored realisations run in
// up to kEnsGemmGroups persistent GEMM launches, one per tile shape and
// operand form (tables of the whole grid, or the 2-D product form): each has
// its own realisation list, tile prefix sums, XCD groups and counters.
// max_tiles bounds the batched prep grid; the *_blocks are the persistent
// grid sizes.
constexpr int kXcdGroups = 8;     // ensemble GEMM work groups (one per XCD of an MI355X)
constexpr int kEnsGemmGroups = 4; // GEMM launches: tile shape (256, 128 rows) x operand form
struct EnsGemmGroup {
  const int* fact;             // realisations of this group, longest first per XCD group
  int nfact;
  const int64_t* tile_start;   // nfact + 1 prefix sums of their GEMM tiles
  const int64_t* group_start;  // tile ranges of the kXcdGroups groups (kXcdGroups + 1)
  int* counters;               // kXcdGroups work counters
  int gemm_blocks;
  int brows;                   // block rows of the tile shape (256 or 128)
  bool kr;                     // B built as S(block) C: 3-D, or the 2-D product form
};
struct EnsLaunch {
  const EnsReal* d;
  int count;
  const int* fact;             // every factored realisation (table tasks)
  int nfact;
  const int64_t* task_start;   // table tasks: 3 nfact + 1 prefix sums (roles S, A, C)
  int64_t n_table_tasks;
  EnsGemmGroup gemm[kEnsGemmGroups];
  int ngemm;
  const int* dir;
  int ndir;
  const int64_t* blk_start;
  int* dir_counter;
  int* counters;               // all counters of the plan (reset at the start of a run)
  int ncounters;
  int dim;
  int max_tiles;
  int direct_blocks;
};
constexpr int kEnsPoints = 2048;  // points per direct work item
// counters: kXcdGroups per GEMM launch of each table batch, then the
// direct-kernel counter
constexpr int kEnsBatchCounters = kEnsGemmGroups * kXcdGroups;
// Default bound on an ensemble's factored-table region (gsc_opts.table_bytes
// = 0): plans whose tables exceed the bound build and consume them in
// batches of realisations that reuse one region.
constexpr size_t kEnsTableBytes = size_t(4) << 30;
// batches[0] holds the plan-wide fields (descriptors, mode preparation,
// direct realisations, all counters); every batch its table tasks and GEMM
// groups; the batches run in order on `stream`.
hipError_t launch_ensemble(const EnsLaunch* batches, int nbatch, int precision, int out_dtype,
                           int accum, hipStream_t stream);


// Coordinates of the product form's tables (FactoredPlan.prod): S rows are
// tiles a = rh col_tiles + ch at (org0 + rh brows step0, org1 + ch tcols
// step1), A rows and C columns the offsets i step0 and j step1.
struct ProdGeom {
  int prod;
  int tcols;
  int64_t col_tiles, brows;
  double org0, org1, step0, step1;
};

// Arguments of the per-operand-format launchers (gsc_gemm.h, gemm_*.hip).
struct TableArgs {
  const ModeRec* modes;
  int M, dim, nch;
  const double* xs;  // axis of S (3-D), A and C
  const double* xa;
  const double* xc;
  int64_t na, nrow, arows, ncols, ncp;
  float2* S;
  uint32_t* A;
  float2* C;
  int trans;  // 2-D: A from axis 1, C from axis 0
  int ncomp;  // C components (1, or dim with the projector folded in)
  ProdGeom pg;  // product form (pg.prod != 0): coordinates from origin and spacing
};
struct GemmArgs {
  const uint32_t* A;
  int64_t arows;
  int nch;
  const float2* S;
  int64_t na;
  const float2* C;
  int64_t ncp;
  int64_t nrow;    // valid rows of the A axis
  int64_t nb_out;  // 3-D: output row of (a, b) is a * nb_out + b
  void* out;
  bool out_f64;
  int64_t ldo, ncols;
  int64_t ldc;  // column stride of the output (1, or nrow for a transposed 2-D plan)
  double scale;
  bool kr;
  int accum;
  int64_t ncpt;        // columns of the C table: ncomp blocks of ncp
  int64_t comp_stride; // output elements between field components
  int brows;           // block rows: 256 or 128 (the plan's tile shape)
  bool prod;           // 2-D product form: tile (x, y) of the grid is S row x col_tiles + y
  int64_t col_tiles;
};

// Expects the mode table at the start of `work` (launch_prep, KIND_SCALAR).
// axes: device pointers, one per dimension.  out: dense C-order grid.
// accum: 0 = FP64 second level, 1 = two-level FP32 (default), 2 = plain FP32.
hipError_t launch_factored(const FactoredPlan& p, int64_t M, const double* const* axes,
                           void* work, void* out, int out_dtype, double scale, int accum,
                           hipStream_t stream);

// The grid's points as a dense (dim, N) point list, C order: pos[i * N + j]
// is coordinate i of flat grid index j.
hipError_t launch_expand_grid(int dim, const int64_t* shape, const double* const* axes,
                              double* pos, hipStream_t stream);

// Direct kernel on a grid: point coordinates come from the axes (device
// pointers) by the C-order flat index; nothing is materialised.
hipError_t launch_direct_grid(int dim, int64_t M, const int64_t* shape, const double* const* axes,
                              const void* table, void* out, int out_dtype, double scale,
                              int variant, hipStream_t stream);

// Per operand format (T = _Float16 or __bf16, NS splits), instantiated in gemm_*.hip.
template <typename T, int NS>
hipError_t tables_launch(const TableArgs& a, dim3 grid, hipStream_t s);
template <typename T, int NS>
hipError_t factored_gemm_launch(const GemmArgs& a, dim3 grid, hipStream_t s);
template <typename T, int NS>
hipError_t ens_launch(const EnsLaunch& e, int part, int accum, bool out_f64, hipStream_t s);

}  // namespace gsc
