// Structured-grid (factored) path: workspace plan and launchers.
#pragma once

#include <hip/hip_runtime.h>

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

// Workspace layout (bytes from the start of the caller's workspace):
//   [0, off_p)     ModeRec table (M records)
//   [off_p, ...)   P: na x mp float2   (axis 0, amplitude and theta folded in)
//   [off_q, ...)   Q: nb x mp float2   (axis 1, 3-D only)
//   [off_bt, ...)  B operand (last axis, split): ns x ncp x kp elements of fp16/bf16
// (all chunk-major: 16 modes of one row are contiguous, rows of one chunk
// follow each other).
struct FactoredPlan {
  int dim;
  int ns;
  bool bf16;
  int mp;          // modes padded to a multiple of 32 (an even number of 16-mode chunks)
  int64_t kp;      // 2 * mp
  int64_t na, nb;  // rows = na * nb (nb = 1 in 2-D)
  int64_t rows;
  int64_t ncols;   // last-axis length
  int64_t ncp;     // ncols padded to the column tile
  size_t off_p, off_q, off_bt, bytes;
};

FactoredPlan plan_factored(int dim, int64_t M, const int64_t* shape, int precision);

// Expects the mode table at the start of `work` (launch_prep, KIND_SCALAR).
// axes: device pointers, one per dimension.  out: dense C-order grid.
// accum: 0 = FP64 flush (default), 1 = two-level FP32, 2 = plain FP32.
hipError_t launch_factored(const FactoredPlan& p, int64_t M, const double* const* axes,
                           void* work, void* out, int out_dtype, double scale, int accum,
                           hipStream_t stream);

// Direct kernel on a grid: point coordinates come from the axes (device
// pointers) by the C-order flat index; nothing is materialised.
hipError_t launch_direct_grid(int dim, int64_t M, const int64_t* shape, const double* const* axes,
                              const void* table, void* out, int out_dtype, double scale,
                              int variant, hipStream_t stream);

}  // namespace gsc
