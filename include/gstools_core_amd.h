/*
 * gstools_core_amd.h -- C-ABI of the MI355X (gfx950) randomization-method
 * summators.
 *
 * The reference crate (GSTools-Core 1.1.0) exposes these computations as Rust
 * functions over ndarray views and as PyO3 functions of the Python module
 * `gstools_core`.  Each entry point below replaces one of them; its comment
 * names the reference interface (file:line) it stands in for.  The
 * signatures use plain pointers, sizes and signed element strides only, so a
 * Rust `extern "C"` block, ctypes or cgo can bind them directly.
 *
 * Conventions
 *  - Arrays are float64 unless stated.  A 2-D array (rows, cols) is addressed
 *    as p[r * s_row + c * s_col] with signed strides counted in elements, so
 *    numpy/ndarray views of any layout (including negative strides) can be
 *    passed without copying on the caller side.
 *  - `cov`  is the (dim, M) matrix of spectral samples k_m (one mode per
 *    column), `z1`, `z2` the M standard-normal weights, `pos` the (dim, N)
 *    matrix of points (one point per column).
 *  - Every function returns a gsc_status.  Nothing unwinds or aborts across
 *    this boundary; the message of the last failure on the calling thread is
 *    available from gsc_last_error().
 *  - Results are NOT scaled (as in the reference) unless gsc_opts.scale is
 *    set; GSTools' RandMeth multiplies by sqrt(var / M) itself.
 *  - The host-pointer entry points are synchronous and thread-safe (one lock
 *    per device).  The *_dev entry points take device pointers, enqueue work
 *    on the given HIP stream (a hipStream_t passed as void*, NULL = the null
 *    stream) and return without synchronising.
 */
#ifndef GSTOOLS_CORE_AMD_H
#define GSTOOLS_CORE_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gsc_status {
  GSC_OK = 0,
  GSC_ERR_INVALID = 1,   /* bad sizes, shapes, strides, null pointers or options */
  GSC_ERR_DIM = 2,       /* unsupported spatial dimension */
  GSC_ERR_NODEVICE = 3,  /* no usable gfx950 device */
  GSC_ERR_HIP = 4,       /* a HIP runtime call failed */
  GSC_ERR_ALLOC = 5,     /* host or device allocation failed */
  GSC_ERR_WORKSPACE = 6  /* caller workspace too small */
} gsc_status;

typedef enum gsc_dtype { GSC_F64 = 0, GSC_F32 = 1 } gsc_dtype;

/* Evaluation path for structured grids (gsc_summate_grid*). */
typedef enum gsc_path {
  GSC_PATH_AUTO = 0,      /* library rule (factored when both grid extents are >= 16) */
  GSC_PATH_DIRECT = 1,    /* per-point kernel (FP64 phase, FP32 sin) */
  GSC_PATH_FACTORED = 2   /* per-axis tables contracted on the MFMA cores */
} gsc_path;

/* Split-precision format of the factored path's MFMA products. */
typedef enum gsc_precision {
  GSC_PREC_AUTO = 0,      /* most accurate format that meets 1e-5 sigma: fp16x3 */
  GSC_PREC_F16X3 = 1,     /* fp16 hi/lo parts, 3 products, FP32 accumulate */
  GSC_PREC_BF16X3 = 2,    /* bf16 hi/lo parts, 3 products */
  GSC_PREC_BF16X6 = 3,    /* bf16 hi/mid/lo parts, 6 products */
  GSC_PREC_F16X1 = 4,     /* fp16, 1 product (fast, ~1e-3 sigma) */
  GSC_PREC_BF16X1 = 5     /* bf16, 1 product (fast, ~1e-2 sigma) */
} gsc_precision;

/* Accumulation of the factored path's FP32 MFMA partial sums, which are
 * flushed every 256 modes into a second accumulator set.  The second set's
 * registers limit the GEMM tile: 128 columns per block for F32X2 and F32
 * (64 for bf16x6), 32 for F64. */
typedef enum gsc_accum {
  GSC_ACCUM_F64 = 0,    /* second set in FP64 (narrow tiles, slowest) */
  GSC_ACCUM_F32X2 = 1,  /* second set in FP32 (default; ~1e-6 sigma at M = 4000) */
  GSC_ACCUM_F32 = 2     /* one FP32 set over all modes (error grows with M) */
} gsc_accum;

/* How the direct (per-point) kernels evaluate the phases k . x.
 * LOCAL sorts the points into spatial cells and splits every phase at the
 * cell centre X: (k . X + theta) in FP64 once per cell and mode, k . (x - X)
 * on the matrix cores, from two-part fp16 splits for scalar fields
 * (v_mfma_f32_32x32x16_f16, 1024 point-modes per instruction) and three-part
 * bf16 splits for incompressible fields (v_mfma_f32_16x16x32_bf16, 256
 * point-modes per instruction); modes whose local phase could exceed one
 * revolution keep the FP64 phase.  It serves all three summators in 2-D and
 * 3-D (N < 2^30); in other dimensions LOCAL runs the VALU kernel.
 * On gfx950 matrix instructions do not overlap with the VALU sine work (the
 * measured time is the sum of both parts). */
typedef enum gsc_direct_variant {
  GSC_DIRECT_MFMA = 0,  /* v_mfma_f64_16x16x4 tiles of 16 modes x 16 points (dim <= 3) */
  GSC_DIRECT_VALU = 1,  /* FP64 FMA chains on the vector ALUs */
  GSC_DIRECT_LOCAL = 2, /* cell-local phases from split-bf16 matrix-core products */
  GSC_DIRECT_AUTO = 3   /* LOCAL from 256 modes up, VALU below, where the cell sort costs
                           more than it saves (default) */
} gsc_direct_variant;

/* Block tile of the factored path's GEMM: 256 rows (4 waves of 64 rows, one
 * wave per SIMD) or 128 rows (4 waves of 32 rows, two blocks per CU: one
 * block's waits are covered by the other's matrix instructions; less row
 * padding on small grids).  Columns: 128 (fp16x3 / bf16x3), 64 (bf16x6),
 * 32 (FP64 flush). */
typedef enum gsc_tile {
  GSC_TILE_AUTO = 0,  /* library default (256 rows) */
  GSC_TILE_256 = 1,
  GSC_TILE_128 = 2
} gsc_tile;

typedef struct gsc_opts {
  int32_t out_dtype;       /* gsc_dtype of `out` (default GSC_F64, as the reference) */
  int32_t path;            /* gsc_path, structured entry points only */
  int32_t precision;       /* gsc_precision, factored path only */
  int32_t device;          /* HIP device ordinal, -1 = current device */
  int32_t num_threads;     /* host threads for staging strided arrays, 0 = all */
  int32_t accum;           /* gsc_accum, factored path only */
  int32_t direct_variant;  /* gsc_direct_variant, direct kernels only */
  int32_t blocks_per_cu;   /* persistent factored-GEMM blocks per CU in ensembles, 0 = auto
                              (2 with GSC_ACCUM_F64, else 1) */
  double scale;            /* multiplies every output value (1.0 = reference) */
  int32_t tile;            /* gsc_tile, factored path only */
  int64_t table_bytes;     /* ensembles: bound on the factored tables held at once; plans
                              whose tables exceed it build and use them in batches of
                              realisations that reuse one region (0 = 4 GiB) */
} gsc_opts;

/* Fills *opts with the defaults: f64 output, auto path/precision, device -1, scale 1. */
void gsc_opts_init(gsc_opts* opts);

/* Library version string ("1.1.0-mi355x.<n>"); the reference exports
 * gstools_core.__version__ (lib.rs:31). */
const char* gsc_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* gsc_last_error(void);

/* Number of visible HIP devices (0 when none; never fails). */
int gsc_device_count(void);

/*
 * Scalar randomization summator.
 *   out[j] = scale * sum_{m<M} z1[m] cos(k_m . x_j) + z2[m] sin(k_m . x_j)
 * Replaces field::summator (src/field.rs:37-65) and the PyO3 function
 * gstools_core.summate (src/lib.rs:33-48).  dim >= 1 (1..3 use specialised
 * kernels, 4..16 a generic one).  Host pointers; `out` has N elements of
 * opts->out_dtype with stride out_s.  opts may be NULL.
 */
int gsc_summate(int dim, int64_t M, int64_t N,
                const double* cov, int64_t cov_sd, int64_t cov_sm,
                const double* z1, int64_t z1_s,
                const double* z2, int64_t z2_s,
                const double* pos, int64_t pos_sd, int64_t pos_sp,
                void* out, int64_t out_s, const gsc_opts* opts);

/*
 * Incompressible (divergence-free) vector summator, dim in {2, 3}:
 *   out[i][j] = scale * sum_m (delta_i0 - k_mi k_m0 / |k_m|^2)
 *                        (z1[m] cos(k_m . x_j) + z2[m] sin(k_m . x_j))
 * Replaces field::summator_incompr (src/field.rs:97-182) and
 * gstools_core.summate_incompr (src/lib.rs:50-65).  `out` is (dim, N) with
 * strides (out_sd, out_sp); the reference returns Fortran order, i.e.
 * out_sd = 1, out_sp = dim.  Returns GSC_ERR_DIM for other dimensions (the
 * reference panics, field.rs:180).
 */
int gsc_summate_incompr(int dim, int64_t M, int64_t N,
                        const double* cov, int64_t cov_sd, int64_t cov_sm,
                        const double* z1, int64_t z1_s,
                        const double* z2, int64_t z2_s,
                        const double* pos, int64_t pos_sd, int64_t pos_sp,
                        void* out, int64_t out_sd, int64_t out_sp, const gsc_opts* opts);

/*
 * Fourier-method summator:
 *   out[j] = scale * sum_m sf[m] (z1[m] cos(k_m . x_j) + z2[m] sin(k_m . x_j))
 * with k_m the columns of `modes`.  Replaces field::summator_fourier
 * (src/field.rs:219-249) and gstools_core.summate_fourier (src/lib.rs:67-84).
 */
int gsc_summate_fourier(int dim, int64_t M, int64_t N,
                        const double* sf, int64_t sf_s,
                        const double* modes, int64_t md_sd, int64_t md_sm,
                        const double* z1, int64_t z1_s,
                        const double* z2, int64_t z2_s,
                        const double* pos, int64_t pos_sd, int64_t pos_sp,
                        void* out, int64_t out_s, const gsc_opts* opts);

/*
 * Device-resident forms of the three summators (same arithmetic).  All
 * arrays are device pointers in dense row-major layout: cov/modes (dim, M),
 * z1/z2/sf (M), pos (dim, N) with row pitch pos_ld >= N elements, out (N) for
 * the scalar forms and (dim, N) with row pitch out_ld >= N for incompr.
 * `work` is device scratch of at least gsc_workspace_bytes(dim, M, N, opts)
 * bytes (the per-mode table and, for the LOCAL variant, the cell sort); it
 * must stay untouched until the stream has run the call.  `sf` is used by the Fourier form only and may be NULL otherwise.
 * These have no counterpart in the reference (which has no device side);
 * they are the device half of the three functions above.
 */
size_t gsc_workspace_bytes(int dim, int64_t M, int64_t N, const gsc_opts* opts);
int gsc_summate_dev(int dim, int64_t M, int64_t N, const double* cov, const double* z1,
                    const double* z2, const double* pos, int64_t pos_ld, void* out,
                    void* work, size_t work_bytes, const gsc_opts* opts, void* stream);
int gsc_summate_incompr_dev(int dim, int64_t M, int64_t N, const double* cov, const double* z1,
                            const double* z2, const double* pos, int64_t pos_ld, void* out,
                            int64_t out_ld, void* work, size_t work_bytes, const gsc_opts* opts,
                            void* stream);

/*
 * gsc_summate_dev without a workspace argument: the same kernels and the
 * same result, with the scratch taken from a device array of
 * GSC_NOALLOC_SCRATCH_BYTES that the library holds on every device it runs on
 * until the process ends.  It calls no hipMalloc / hipFree, so callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid allocation) can use it.  No counterpart in the
 * reference.  Calls whose workspace (gsc_workspace_bytes) exceeds the array
 * return GSC_ERR_WORKSPACE.  The array belongs to the device opts->device (-1:
 * the current device), which must be the device of `stream`.  All calls on a
 * device share its one array and leave their data in it: calls that overlap
 * on one device (different streams or threads without ordering) race.
 */
#define GSC_NOALLOC_SCRATCH_BYTES ((size_t)512 << 20)
int gsc_summate_dev_noalloc(int dim, int64_t M, int64_t N, const double* cov, const double* z1,
                            const double* z2, const double* pos, int64_t pos_ld, void* out,
                            const gsc_opts* opts, void* stream);
int gsc_summate_fourier_dev(int dim, int64_t M, int64_t N, const double* sf, const double* modes,
                            const double* z1, const double* z2, const double* pos, int64_t pos_ld,
                            void* out, void* work, size_t work_bytes, const gsc_opts* opts,
                            void* stream);

/*
 * Phase timing of the device summators (no counterpart in the reference):
 * runs the *_dev call of `field` (a gsc_field: scalar, incompressible with
 * out_ld, or Fourier with sf) `repeats` times on `stream`, with HIP events
 * before it, between the mode preparation and cell sort and the summation
 * kernel, and after it; returns the mean ms of the preparation and sort in
 * ms[0], of the summation kernel in ms[1] and of the whole call in ms[2].
 */
int gsc_direct_phase_times(int field, int dim, int64_t M, int64_t N, const double* sf,
                           const double* cov, const double* z1, const double* z2,
                           const double* pos, int64_t pos_ld, void* out, int64_t out_ld,
                           void* work, size_t work_bytes, const gsc_opts* opts, void* stream,
                           int repeats, double* ms);

/*
 * Structured grids.  The field on the C-order grid axes[0] x ... x
 * axes[dim-1] (shape[i] coordinates on axis i, dim in {1, 2, 3}) equals
 * gsc_summate on the expanded point list
 *   pos[:, j] = (axes[0][i0], axes[1][i1], axes[2][i2]),  j = (i0 * shape[1] + i1) * shape[2] + i2
 * which is how GSTools calls summate for structured meshes (it expands the
 * mesh in Python first).  These entry points have no counterpart in the
 * reference: they are the grid front-end of gsc_summate that avoids the
 * expansion.  opts->path selects the per-point kernel (GSC_PATH_DIRECT) or
 * the factored MFMA contraction (GSC_PATH_FACTORED, dim 2 or 3, precision
 * from opts->precision); GSC_PATH_AUTO lets the library choose.  `out` is the
 * dense C-order grid (prod(shape) values of opts->out_dtype).
 *
 * gsc_grid_path returns the path (GSC_PATH_DIRECT / GSC_PATH_FACTORED) the
 * library takes for these arguments, or -1 on invalid arguments.
 */
int gsc_grid_path(int dim, int64_t M, const int64_t* shape, const gsc_opts* opts);
size_t gsc_grid_workspace_bytes(int dim, int64_t M, const int64_t* shape, const gsc_opts* opts);
int gsc_summate_grid(int dim, int64_t M,
                     const double* cov, int64_t cov_sd, int64_t cov_sm,
                     const double* z1, int64_t z1_s,
                     const double* z2, int64_t z2_s,
                     const double* const* axes, const int64_t* axis_s, const int64_t* shape,
                     void* out, const gsc_opts* opts);
/* Device form: cov (dim, M), z1, z2, every axes[i] and out are device
 * pointers (`axes` itself is a host array of dim device pointers). */
int gsc_summate_grid_dev(int dim, int64_t M, const double* cov, const double* z1,
                         const double* z2, const double* const* axes, const int64_t* shape,
                         void* out, void* work, size_t work_bytes, const gsc_opts* opts,
                         void* stream);

/*
 * Structured grids for every field of the reference's summators:
 *   GSC_FIELD_SCALAR   field::summator (field.rs:37-65), as gsc_summate_grid;
 *   GSC_FIELD_INCOMPR  field::summator_incompr (field.rs:97-182), dim 2 or 3;
 *                      `out` holds dim component grids back to back
 *                      (out[i * prod(shape) + j], j the C-order grid index);
 *   GSC_FIELD_FOURIER  field::summator_fourier (field.rs:219-249) with the
 *                      spectrum factors sf (M values, stride sf_s; `cov` holds
 *                      the modes).
 * Each equals its point-list summator on the expanded grid.  On the factored
 * path the incompressible field folds the projector delta_i0 - k_i k_0 / |k|^2
 * into the last axis' table, so one contraction with dim times the columns
 * gives all components.  No counterpart in the reference (its callers expand
 * structured meshes into point lists).  sf is ignored unless FOURIER.
 */
typedef enum gsc_field {
  GSC_FIELD_SCALAR = 0,
  GSC_FIELD_INCOMPR = 1,
  GSC_FIELD_FOURIER = 2
} gsc_field;
size_t gsc_grid_field_workspace_bytes(int field, int dim, int64_t M, const int64_t* shape,
                                      const gsc_opts* opts);
int gsc_summate_grid_field(int field, int dim, int64_t M, const double* sf, int64_t sf_s,
                           const double* cov, int64_t cov_sd, int64_t cov_sm,
                           const double* z1, int64_t z1_s, const double* z2, int64_t z2_s,
                           const double* const* axes, const int64_t* axis_s,
                           const int64_t* shape, void* out, const gsc_opts* opts);
int gsc_summate_grid_field_dev(int field, int dim, int64_t M, const double* sf,
                               const double* cov, const double* z1, const double* z2,
                               const double* const* axes, const int64_t* shape, void* out,
                               void* work, size_t work_bytes, const gsc_opts* opts,
                               void* stream);

/*
 * Autotuner support (no counterpart in the reference): the mean time in ms
 * of one gsc_summate_grid_field_dev call with these sizes and options, over
 * `repeats` calls after one warm-up call, from HIP events on a private
 * stream.  The library allocates its own synthetic inputs (unit-spaced axes,
 * seeded Gaussian modes, spectrum factors of one) and frees them.
 */
int gsc_tune_grid_time(int field, int dim, int64_t M, const int64_t* shape,
                       const gsc_opts* opts, int repeats, double* ms);

/*
 * Ensembles of realisations on structured grids (no counterpart in the
 * reference, whose callers loop over realisations in Python and call summate
 * once per field).  `count` realisations of one dimension dim in {2, 3};
 * realisation r has m_counts[r] modes and the grid shapes[r * dim + i].
 * Device arrays hold the realisations back to back:
 *   cov  (dim, M_r) row-major blocks;  z1, z2  M_r values each;
 *   axes for each r its dim axes back to back;
 *   out  for each r its dense C-order grid (opts->out_dtype).
 * paths[r] is GSC_PATH_DIRECT, GSC_PATH_FACTORED or GSC_PATH_AUTO (NULL: all
 * AUTO, the rule of gsc_grid_path); scales[r] multiplies realisation r
 * (NULL: opts->scale for all).  The plan owns its device workspace and a HIP
 * graph captured at creation: a counter reset, one mode-preparation launch,
 * one table launch, one persistent factored-GEMM kernel and one persistent
 * direct kernel that take 128x128 tiles / 2048-point blocks of all
 * realisations from flat work lists (longest modes first).  gsc_ensemble_run
 * replays the graph on `stream`; runs of one plan must not overlap.
 */
typedef struct gsc_ensemble gsc_ensemble;
int gsc_ensemble_create(gsc_ensemble** plan, int count, int dim, const int64_t* shapes,
                        const int64_t* m_counts, const int32_t* paths, const double* scales,
                        const double* cov, const double* z1, const double* z2,
                        const double* axes, void* out, const gsc_opts* opts);
int gsc_ensemble_run(gsc_ensemble* plan, void* stream);
/* Mean ms of one replay of the plan over `repeats` replays after a warm-up
 * replay, from HIP events on a private stream (autotuner support). */
int gsc_ensemble_time(gsc_ensemble* plan, int repeats, double* ms);
/* n_factored / n_direct realisations, workspace bytes, total point-modes. */
int gsc_ensemble_info(const gsc_ensemble* plan, int64_t* n_factored, int64_t* n_direct,
                      size_t* workspace_bytes, double* point_modes);
void gsc_ensemble_destroy(gsc_ensemble* plan);

/*
 * Seeded spectral-mode sampling: the caller side of the summators (GSTools'
 * RandMeth draws k from the model's spectral density and z1, z2 ~ N(0, 1) on
 * the host before calling summate, gen_benchmark_inputs.py:17-29; GSTools
 * itself is not part of the reference).  No counterpart in the reference.
 * Fills cov (dim, M) row-major, z1 and z2 (M), dim 1..3:
 *   z1, z2 ~ N(0, 1);
 *   GSC_MODEL_GAUSSIAN:    k ~ N(0, len_scale^-2 I);
 *   GSC_MODEL_EXPONENTIAL: k = g / (len_scale |w|), g ~ N(0, I), w ~ N(0, 1)
 *                          (isotropic multivariate t with one degree of freedom).
 * Random bits: Philox4x64-10 keyed (seed, stream_id), word for word the
 * stream of numpy.random.Philox with key words (seed, stream_id); mode m takes words
 * 8m .. 8m + 7 (four Box-Muller pairs), so mode m does not depend on M.
 */
typedef enum gsc_model { GSC_MODEL_GAUSSIAN = 0, GSC_MODEL_EXPONENTIAL = 1 } gsc_model;
/* Host arrays (synchronous). */
int gsc_sample_modes(int model, int dim, int64_t M, double len_scale, uint64_t seed,
                     uint64_t stream_id, double* cov, double* z1, double* z2,
                     const gsc_opts* opts);
/* Device arrays, enqueued on `stream`. */
int gsc_sample_modes_dev(int model, int dim, int64_t M, double len_scale, uint64_t seed,
                         uint64_t stream_id, double* cov, double* z1, double* z2,
                         const gsc_opts* opts, void* stream);
/* Ensemble whose modes are drawn by the sampler above at the start of every
 * replay (realisation r: model, len_scales[r], seeds[r], stream_id 0), into
 * plan-owned device arrays; otherwise as gsc_ensemble_create.  No counterpart
 * in the reference.  gsc_ensemble_reseed copies `count` new seeds (host
 * array) into the plan, ordered on `stream` before the next run. */
int gsc_ensemble_create_seeded(gsc_ensemble** plan, int count, int dim, const int64_t* shapes,
                               const int64_t* m_counts, const int32_t* paths,
                               const double* scales, int model, const double* len_scales,
                               const uint64_t* seeds, const double* axes, void* out,
                               const gsc_opts* opts);
int gsc_ensemble_reseed(gsc_ensemble* plan, const uint64_t* seeds, void* stream);

/* Either creator with a GEMM tile shape per realisation (gsc_tile, e.g. the
 * autotuner's choice for its shape class): tiles[r] for r < count, or NULL
 * (opts->tile for all).  The factored realisations of each tile shape share
 * one persistent GEMM launch of the plan's graph.  seeds == NULL: the modes
 * come from cov, z1, z2 (model and len_scales unused), as
 * gsc_ensemble_create; otherwise they are drawn as gsc_ensemble_create_seeded
 * does (cov, z1, z2 unused).  No counterpart in the reference. */
int gsc_ensemble_create_tuned(gsc_ensemble** plan, int count, int dim, const int64_t* shapes,
                              const int64_t* m_counts, const int32_t* paths, const int32_t* tiles,
                              const double* scales, const double* cov, const double* z1,
                              const double* z2, int model, const double* len_scales,
                              const uint64_t* seeds, const double* axes, void* out,
                              const gsc_opts* opts);

/*
 * Kriging calculators, n conditions and N targets:
 *   krig_fac_i(j) = sum_k mat[k][i] vecs[k][j]
 *   field[j] = sum_i cond[i] krig_fac_i(j),   error[j] = sum_i vecs[i][j] krig_fac_i(j)
 * Replaces krige::calculator_field_krige_and_variance (src/krige.rs:24-70),
 * gstools_core.calc_field_krige_and_variance (src/lib.rs:86-103) and, with
 * error == NULL, krige::calculator_field_krige (krige.rs:88-118) /
 * gstools_core.calc_field_krige (lib.rs:105-118).  mat is (n, n), vecs
 * (n, N), cond (n), all FP64 with signed element strides; field and error
 * receive N values with strides field_s / error_s.  With error, the 2 n^2 N
 * products run as FP64 matrix-core tiles; the field alone costs n^2 + 2 n N
 * (mat cond first).  Host pointers, synchronous.
 */
int gsc_krige(int64_t n, int64_t N, const double* mat, int64_t mat_s0, int64_t mat_s1,
              const double* vecs, int64_t vecs_s0, int64_t vecs_s1, const double* cond,
              int64_t cond_s, double* field, int64_t field_s, double* error, int64_t error_s,
              const gsc_opts* opts);
/* Workspace of gsc_krige_dev: 2 * 8 * ceil(n / 128) * N bytes with the error,
 * 8 n bytes for the field alone. */
size_t gsc_krige_workspace_bytes(int64_t n, int64_t N, int with_error);
/* Device form: mat dense (n, n) row-major, vecs (n, N) with row pitch
 * vecs_ld >= N, cond (n), field and error (N, dense; error may be NULL);
 * enqueued on `stream`. */
int gsc_krige_dev(int64_t n, int64_t N, const double* mat, const double* vecs, int64_t vecs_ld,
                  const double* cond, double* field, double* error, void* work,
                  size_t work_bytes, const gsc_opts* opts, void* stream);

/* Roofline calibration of the direct kernels (no counterpart in the
 * reference): the v_sin_f32 rate the current device sustains, alone
 * (*sin_per_s) and paired with one dependent FP32 FMA per sine
 * (*sin_fma_per_s), from HIP events over `repeats` launches of a kernel with
 * eight independent chains per lane and eight waves per SIMD.  Returns a
 * gsc_status (no message is set). */
int gsc_calibrate_sine(int repeats, double* sin_per_s, double* sin_fma_per_s);

/* Ceiling of the cell-local kernels' instruction mix (no counterpart in the
 * reference): point-modes per second of a kernel that runs, per 1024
 * point-modes of a wave, the scalar kernel's mix (one 32x32x16 fp16 matrix
 * product, 16 sines, 16 FMAs; *scalar_pm_per_s) or the three-component
 * vector kernel's (four 16x16x32 bf16 products, 16 sines, 48 FMAs;
 * *vector_pm_per_s), four waves per SIMD and no memory traffic, from HIP
 * events over `repeats` launches.  Returns a gsc_status (no message). */
int gsc_calibrate_local_mix(int repeats, double* scalar_pm_per_s, double* vector_pm_per_s);

/* Device memory plumbing for hosts without their own allocator (e.g. plain
 * numpy callers of the device-resident entry points). */
void* gsc_device_alloc(size_t bytes, int device);
void gsc_device_free(void* ptr);
int gsc_memcpy_to_device(void* dst, const void* src, size_t bytes);
int gsc_memcpy_to_host(void* dst, const void* src, size_t bytes);
int gsc_synchronize(void);

#ifdef __cplusplus
}
#endif

#endif /* GSTOOLS_CORE_AMD_H */
