// Host-side staging of the synchronous entry points: a persistent pool of
// worker threads (the reference's `num_threads`, field.rs:48-51, sizes the
// share of it one call uses) and pinned double buffers that overlap the host
// copies of one chunk with the DMA of the previous one.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>

namespace gsc {

// Runs fn(t) for t in [0, n_tasks) on up to `threads` threads (the caller
// included; threads <= 0: every worker).  Returns when all tasks are done.
void parallel_for(int64_t n_tasks, int threads, const std::function<void(int64_t)>& fn);

// Pinned staging buffers of one device (two chunks and their events).
struct Staging {
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
};
constexpr size_t kStageChunk = size_t(16) << 20;  // bytes per pinned chunk

// Allocates the buffers on first use (hipHostMalloc); false on failure.
bool staging_ready(Staging& s);
void staging_release(Staging& s);

// Host (rows, cols) doubles at src[r * sr + c * sc] -> dense row-major
// device array d_dst, through the pinned chunks, enqueued on `stream`.
// Returns when the last chunk's DMA has been queued (src may then be reused
// only after the stream has run it: the caller synchronises).
hipError_t stage_upload(Staging& s, hipStream_t stream, double* d_dst, const double* src,
                        int64_t rows, int64_t cols, int64_t sr, int64_t sc, int threads);

// Dense row-major device (rows, cols) elements of `elt` bytes (4 or 8) ->
// host dst[r * sr + c * sc], through the pinned chunks.  Synchronous: it
// waits for the work already on `stream` and for its own copies.
hipError_t stage_download(Staging& s, hipStream_t stream, const void* d_src, void* dst,
                          int64_t rows, int64_t cols, int64_t sr, int64_t sc, size_t elt,
                          int threads);

// (ncomp, N) row-major device array -> (N, ncomp) row-major (the memory
// order of a Fortran-order (ncomp, N) array), elements of `elt` bytes.
hipError_t launch_transpose_cn(const void* in, void* out, int ncomp, int64_t N, size_t elt,
                               hipStream_t stream);

}  // namespace gsc
