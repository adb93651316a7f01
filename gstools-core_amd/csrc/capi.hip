// C-ABI entry points (include/gstools_core_amd.h).
//
// The host-pointer functions mirror the reference's contract
// (src/lib.rs:33-84 -> src/field.rs:37-249): read-only strided inputs, a
// freshly computed result, shape checks first.  Where the reference panics
// (field.rs:44-46, 104-106, 180; panic = "abort" in Cargo.toml:21-22) these
// return a status code and leave a message for gsc_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "gsc_factored.h"
#include "gsc_host.h"
#include "gsc_internal.h"
#include "gstools_core_amd.h"

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define GSC_HIP(call)                                                                  \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess)                                                              \
      return fail(GSC_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));         \
  } while (0)

// Per-device state for the synchronous host-pointer path: one stream and one
// growing device buffer, guarded by a mutex (calls on one device serialise).
struct DeviceCtx {
  std::mutex mu;
  std::once_flag arch_once;  // the gfx950 check runs once per device, from any thread
  bool ok = false;
  char arch[64] = {0};
  hipStream_t stream = nullptr;
  void* buf = nullptr;
  size_t cap = 0;
  gsc::Staging stage;  // pinned chunks of the host <-> device copies
};

// Host arrays of at least this many bytes go through the pinned chunks; below
// it, one pageable copy costs less than the pipeline's set-up.
constexpr size_t kStageMin = size_t(4) << 20;
constexpr int kMaxDevices = 64;
DeviceCtx g_dev[kMaxDevices];

const gsc_opts* opts_or_default(const gsc_opts* o, gsc_opts* tmp) {
  if (o) return o;
  gsc_opts_init(tmp);
  return tmp;
}

int resolve_device(const gsc_opts* o, int* dev) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(GSC_ERR_NODEVICE, "no HIP device is visible");
  int d = o->device;
  if (d < 0) {
    if (hipGetDevice(&d) != hipSuccess) d = 0;
  }
  if (d >= n || d >= kMaxDevices)
    return fail(GSC_ERR_INVALID, "device %d out of range (%d visible)", d, n);
  *dev = d;
  return GSC_OK;
}

// Checks once per device that it is a gfx950 (the only code object built).
// Safe to call from concurrent first calls on different threads.
int check_arch(int dev) {
  DeviceCtx& c = g_dev[dev];
  std::call_once(c.arch_once, [&c, dev] {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
      std::snprintf(c.arch, sizeof(c.arch), "(unknown)");
      return;
    }
    std::snprintf(c.arch, sizeof(c.arch), "%s", p.gcnArchName);
    c.ok = std::strncmp(p.gcnArchName, "gfx950", 6) == 0;
  });
  return c.ok ? GSC_OK
              : fail(GSC_ERR_NODEVICE, "device %d is %s, this library is built for gfx950", dev,
                     c.arch);
}

int validate_common(int dim, int64_t M, int64_t N, const double* cov, const double* z1,
                    const double* z2, const double* pos, const void* out) {
  if (dim < 1) return fail(GSC_ERR_INVALID, "dim must be >= 1 (got %d)", dim);
  if (dim > gsc::kMaxDim)
    return fail(GSC_ERR_DIM, "dim %d is above the supported maximum %d", dim, gsc::kMaxDim);
  if (M < 0 || N < 0) return fail(GSC_ERR_INVALID, "negative size (M=%lld, N=%lld)",
                                  (long long)M, (long long)N);
  if (M > (int64_t)0x7fffffff) return fail(GSC_ERR_INVALID, "M=%lld exceeds 2^31-1", (long long)M);
  if (M > 0 && (!cov || !z1 || !z2)) return fail(GSC_ERR_INVALID, "null mode array");
  if (N > 0 && (!pos || !out)) return fail(GSC_ERR_INVALID, "null position or output array");
  return GSC_OK;
}

// GEMM tile shape of the factored path (GSC_TILE_AUTO: 256-row blocks).
int grid_tile(const gsc_opts* o) { return o->tile == GSC_TILE_AUTO ? GSC_TILE_256 : o->tile; }

int validate_opts(const gsc_opts* o) {
  if (o->out_dtype != GSC_F64 && o->out_dtype != GSC_F32)
    return fail(GSC_ERR_INVALID, "out_dtype must be GSC_F64 or GSC_F32");
  if (o->direct_variant < GSC_DIRECT_MFMA || o->direct_variant > GSC_DIRECT_AUTO)
    return fail(GSC_ERR_INVALID, "unknown direct_variant %d", o->direct_variant);
  if (o->blocks_per_cu < 0 || o->blocks_per_cu > 64)
    return fail(GSC_ERR_INVALID, "blocks_per_cu must be in 0..64");
  if (o->tile < GSC_TILE_AUTO || o->tile > GSC_TILE_128)
    return fail(GSC_ERR_INVALID, "unknown tile shape %d", o->tile);
  if (o->table_bytes < 0) return fail(GSC_ERR_INVALID, "table_bytes must be >= 0");
  return GSC_OK;
}

// Gathers a strided 2-D host array (rows, cols) into dense row-major `dst`,
// in column ranges on the host pool (nthreads: 0 = all workers).
void gather2(const double* src, int64_t rows, int64_t cols, int64_t sr, int64_t sc, double* dst,
             int nthreads) {
  if (sc == 1 && sr == cols) {
    std::memcpy(dst, src, sizeof(double) * rows * cols);
    return;
  }
  constexpr int64_t kCols = int64_t(1) << 16;
  const int64_t pieces = rows * cols < (1 << 20) ? 1 : (cols + kCols - 1) / kCols;
  const int64_t step = (cols + pieces - 1) / pieces;
  gsc::parallel_for(pieces, nthreads, [&](int64_t p) {
    const int64_t c0 = p * step, c1 = std::min(cols, c0 + step);
    for (int64_t r = 0; r < rows; ++r)
      for (int64_t c = c0; c < c1; ++c) dst[r * cols + c] = src[r * sr + c * sc];
  });
}

void gather1(const double* src, int64_t n, int64_t s, double* dst) {
  if (s == 1) {
    std::memcpy(dst, src, sizeof(double) * n);
    return;
  }
  for (int64_t i = 0; i < n; ++i) dst[i] = src[i * s];
}

size_t align256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }

// Shared body of the three host-pointer summators.
//   kind: gsc::KIND_*; for KIND_INCOMPR `out` is (dim, N) with strides
//   (out_sd, out_sp), otherwise N values with stride out_s = out_sp.
// The local-phase (binned) direct kernel serves all three summators in 2-D
// and 3-D; other dimensions use the FP64-phase VALU kernel for that variant.
// GSC_DIRECT_AUTO: the local variant from kAutoLocalModes modes up; below,
// its cell sort (tens of microseconds plus ~33 ns per point) costs more than
// the ~0.13 ns per point-mode it saves over the FP64-phase VALU kernel.
constexpr int64_t kAutoLocalModes = 256;
int resolve_variant(const gsc_opts* o, int64_t M) {
  if (o->direct_variant != GSC_DIRECT_AUTO) return o->direct_variant;
  return M >= kAutoLocalModes ? GSC_DIRECT_LOCAL : GSC_DIRECT_VALU;
}

bool use_binned(int kind, int dim, int64_t M, int64_t N, const gsc_opts* o) {
  (void)kind;
  return resolve_variant(o, M) == GSC_DIRECT_LOCAL && gsc::binned_supported(dim, N);
}

size_t direct_work_bytes(int kind, int dim, int64_t M, int64_t N, const gsc_opts* o) {
  const size_t table = align256(gsc::mode_table_bytes(dim, M));
  return use_binned(kind, dim, M, N, o) ? table + gsc::binned_work_bytes(dim, M, N) : table;
}

// Launches the direct kernel of `kind` on device buffers; `work` holds the
// mode table (written by launch_prep) followed by the binned kernel's area.
// split: null, or an event recorded before the summation kernel (after the
// cell sort of the LOCAL variant).
hipError_t launch_direct(int kind, int dim, int64_t M, int64_t N, void* work, const double* pos,
                         int64_t pos_ld, void* out, int64_t out_ld, const gsc_opts* o,
                         hipStream_t s, hipEvent_t split = nullptr) {
  if (use_binned(kind, dim, M, N, o)) {
    void* bin = static_cast<char*>(work) + align256(gsc::mode_table_bytes(dim, M));
    return gsc::launch_direct_binned(dim, M, N, work, pos, pos_ld, out, out_ld,
                                     kind == gsc::KIND_INCOMPR, o->out_dtype, o->scale, bin, s,
                                     split);
  }
  if (split) {
    const hipError_t e = hipEventRecord(split, s);
    if (e != hipSuccess) return e;
  }
  if (kind == gsc::KIND_INCOMPR) {
    const int variant = o->direct_variant == GSC_DIRECT_MFMA ? GSC_DIRECT_MFMA : GSC_DIRECT_VALU;
    return gsc::launch_direct_incompr(dim, M, N, work, pos, pos_ld, out, out_ld, o->out_dtype,
                                      o->scale, variant, s);
  }
  const int variant = o->direct_variant == GSC_DIRECT_MFMA ? GSC_DIRECT_MFMA : GSC_DIRECT_VALU;
  return gsc::launch_direct_scalar(dim, M, N, work, pos, pos_ld, out, o->out_dtype, o->scale,
                                   variant, s);
}

// The mode table (from cov, z1, z2 and, for the Fourier summator, sf) and the
// direct kernel of `kind`: the local variant prepares the table in extra
// blocks of its first sort kernel, the others launch the preparation first.
hipError_t launch_prep_direct(int kind, int dim, int64_t M, int64_t N, const double* cov,
                              const double* z1, const double* z2, const double* sf, void* work,
                              const double* pos, int64_t pos_ld, void* out, int64_t out_ld,
                              const gsc_opts* o, hipStream_t s, hipEvent_t split = nullptr) {
  if (use_binned(kind, dim, M, N, o)) {
    const gsc::PrepArgs pa = gsc::prep_args(kind, dim, M, cov, z1, z2, sf, work);
    void* bin = static_cast<char*>(work) + align256(gsc::mode_table_bytes(dim, M));
    return gsc::launch_direct_binned(dim, M, N, work, pos, pos_ld, out, out_ld,
                                     kind == gsc::KIND_INCOMPR, o->out_dtype, o->scale, bin, s,
                                     split, &pa);
  }
  const hipError_t e = gsc::launch_prep(kind, dim, M, cov, z1, z2, sf, work, s);
  if (e != hipSuccess) return e;
  return launch_direct(kind, dim, M, N, work, pos, pos_ld, out, out_ld, o, s, split);
}

int host_summate(int kind, int dim, int64_t M, int64_t N, const double* sf, int64_t sf_s,
                 const double* cov, int64_t cov_sd, int64_t cov_sm, const double* z1,
                 int64_t z1_s, const double* z2, int64_t z2_s, const double* pos, int64_t pos_sd,
                 int64_t pos_sp, void* out, int64_t out_sd, int64_t out_sp,
                 const gsc_opts* opts_in) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts_in, &tmp);
  int rc = validate_common(dim, M, N, cov, z1, z2, pos, out);
  if (rc) return rc;
  if ((rc = validate_opts(o))) return rc;
  if (kind == gsc::KIND_INCOMPR && dim != 2 && dim != 3)
    return fail(GSC_ERR_DIM, "Only two- and three-dimensional problems are supported.");
  if (kind == gsc::KIND_FOURIER && M > 0 && !sf)
    return fail(GSC_ERR_INVALID, "null spectrum_factor array");
  if (N == 0) return GSC_OK;

  const int ncomp = kind == gsc::KIND_INCOMPR ? dim : 1;
  const size_t elt = o->out_dtype == GSC_F32 ? 4 : 8;
  // result view: (ncomp, N) with strides (out_sd, out_sp)
  const int64_t o_sd = kind == gsc::KIND_INCOMPR ? out_sd : N;
  // the reference's Fortran-order (d, N) result (field.rs:166-174) is the
  // device's (d, N) rows transposed: done on the device, then one flat copy
  const bool out_fortran = kind == gsc::KIND_INCOMPR && out_sd == 1 && out_sp == ncomp;
  try {
    // Dense host staging of the (small) mode arrays
    const int64_t Mm = std::max<int64_t>(M, 1);
    std::vector<double> modes((size_t)(dim * Mm + 3 * Mm), 0.0);
    if (M > 0) {
      gather2(cov, dim, M, cov_sd, cov_sm, modes.data(), 1);
      gather1(z1, M, z1_s, modes.data() + dim * M);
      gather1(z2, M, z2_s, modes.data() + (dim + 1) * M);
      if (kind == gsc::KIND_FOURIER) gather1(sf, M, sf_s, modes.data() + (dim + 2) * M);
    }
    int dev = 0;
    if ((rc = resolve_device(o, &dev))) return rc;
    DeviceCtx& c = g_dev[dev];
    std::lock_guard<std::mutex> lock(c.mu);
    GSC_HIP(hipSetDevice(dev));
    if ((rc = check_arch(dev))) return rc;
    if (!c.stream) GSC_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    const bool staged_in = sizeof(double) * (size_t)dim * (size_t)N >= kStageMin &&
                           gsc::staging_ready(c.stage);
    const bool staged_out = elt * (size_t)ncomp * (size_t)N >= kStageMin &&
                            gsc::staging_ready(c.stage);

    const size_t b_table = align256(direct_work_bytes(kind, dim, M, N, o));
    const size_t b_modes = align256(sizeof(double) * modes.size());
    const size_t b_pos = align256(sizeof(double) * dim * N);
    const size_t b_out = align256(elt * ncomp * N);
    const size_t need = b_table + b_modes + b_pos + b_out * (out_fortran ? 2 : 1);
    if (need > c.cap) {
      if (c.buf) GSC_HIP(hipFree(c.buf));
      c.buf = nullptr;
      c.cap = 0;
      if (hipMalloc(&c.buf, need) != hipSuccess)
        return fail(GSC_ERR_ALLOC, "hipMalloc of %zu bytes failed", need);
      c.cap = need;
    }
    char* base = static_cast<char*>(c.buf);
    void* d_table = base;
    double* d_modes = reinterpret_cast<double*>(base + b_table);
    double* d_pos = reinterpret_cast<double*>(base + b_table + b_modes);
    void* d_out = base + b_table + b_modes + b_pos;

    GSC_HIP(hipMemcpyAsync(d_modes, modes.data(), sizeof(double) * modes.size(),
                           hipMemcpyHostToDevice, c.stream));
    std::vector<double> pos_dense;
    if (staged_in) {
      GSC_HIP(gsc::stage_upload(c.stage, c.stream, d_pos, pos, dim, N, pos_sd, pos_sp,
                                o->num_threads));
    } else {
      const double* pos_src = pos;
      if (!(pos_sp == 1 && pos_sd == N)) {
        pos_dense.resize((size_t)(dim * N));
        gather2(pos, dim, N, pos_sd, pos_sp, pos_dense.data(), o->num_threads);
        pos_src = pos_dense.data();
      }
      GSC_HIP(hipMemcpyAsync(d_pos, pos_src, sizeof(double) * dim * N, hipMemcpyHostToDevice,
                             c.stream));
    }
    GSC_HIP(launch_prep_direct(kind, dim, M, N, d_modes, d_modes + dim * M,
                               d_modes + (dim + 1) * M,
                               kind == gsc::KIND_FOURIER ? d_modes + (dim + 2) * M : nullptr,
                               d_table, d_pos, N, d_out, N, o, c.stream));

    // result: a flat (1, ncomp N) copy for dense C-order and Fortran-order
    // views, else the strided (ncomp, N) scatter
    const void* d_res = d_out;
    int64_t rows = ncomp, cols = N, sr = o_sd, sc = out_sp;
    if (out_fortran) {
      void* d_t = static_cast<char*>(d_out) + b_out;
      GSC_HIP(gsc::launch_transpose_cn(d_out, d_t, ncomp, N, elt, c.stream));
      d_res = d_t;
      rows = 1;
      cols = ncomp * N;
      sr = cols;
      sc = 1;
    }
    if (staged_out) {
      GSC_HIP(gsc::stage_download(c.stage, c.stream, d_res, out, rows, cols, sr, sc, elt,
                                  o->num_threads));
    } else {
      const bool flat = sc == 1 && (rows == 1 || sr == cols);
      std::vector<char> out_tmp;
      void* host_dst = out;
      if (!flat) {
        out_tmp.resize(elt * ncomp * N);
        host_dst = out_tmp.data();
      }
      GSC_HIP(hipMemcpyAsync(host_dst, d_res, elt * ncomp * N, hipMemcpyDeviceToHost, c.stream));
      GSC_HIP(hipStreamSynchronize(c.stream));
      if (!flat) {
        for (int64_t r = 0; r < rows; ++r)
          for (int64_t j = 0; j < cols; ++j) {
            const int64_t src = r * cols + j;
            const int64_t dst = r * sr + j * sc;
            if (elt == 8)
              static_cast<double*>(out)[dst] = reinterpret_cast<const double*>(out_tmp.data())[src];
            else
              static_cast<float*>(out)[dst] = reinterpret_cast<const float*>(out_tmp.data())[src];
          }
      }
    }
  } catch (const std::bad_alloc&) {
    return fail(GSC_ERR_ALLOC, "host allocation failed");
  } catch (...) {
    return fail(GSC_ERR_INVALID, "unexpected host exception");
  }
  return GSC_OK;
}

int dev_common(int kind, int dim, int64_t M, int64_t N, const double* sf, const double* cov,
               const double* z1, const double* z2, const double* pos, int64_t pos_ld, void* out,
               int64_t out_ld, void* work, size_t work_bytes, const gsc_opts* opts_in,
               void* stream, hipEvent_t split = nullptr) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts_in, &tmp);
  int rc = validate_common(dim, M, N, cov, z1, z2, pos, out);
  if (rc) return rc;
  if ((rc = validate_opts(o))) return rc;
  if (kind == gsc::KIND_INCOMPR && dim != 2 && dim != 3)
    return fail(GSC_ERR_DIM, "Only two- and three-dimensional problems are supported.");
  if (kind == gsc::KIND_FOURIER && M > 0 && !sf)
    return fail(GSC_ERR_INVALID, "null spectrum_factor array");
  if (pos_ld < N) return fail(GSC_ERR_INVALID, "pos_ld (%lld) < N (%lld)", (long long)pos_ld,
                              (long long)N);
  if (kind == gsc::KIND_INCOMPR && out_ld < N)
    return fail(GSC_ERR_INVALID, "out_ld (%lld) < N (%lld)", (long long)out_ld, (long long)N);
  const size_t need = direct_work_bytes(kind, dim, M, N, o);
  if (!work || work_bytes < need)
    return fail(GSC_ERR_WORKSPACE, "workspace of %zu bytes is below the %zu required", work_bytes,
                need);
  if (N == 0) return GSC_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  GSC_HIP(launch_prep_direct(kind, dim, M, N, cov, z1, z2, sf, work, pos, pos_ld, out, out_ld, o,
                             s, split));
  return GSC_OK;
}

// ---- the no-allocation entry point ------------------------------------------------
// gsc_summate_dev_noalloc runs the kernels of gsc_summate_dev on this array
// instead of a caller workspace.  Every device the library runs on holds its
// own copy until the process ends.  In host code `scratch` is only a stand-in
// name: its device address on each device comes from hipGetSymbolAddress,
// looked up once per device (with that device current) and kept in
// g_scratch; kernels only ever receive that address.
}  // namespace

static __device__ __align__(256) unsigned char scratch[GSC_NOALLOC_SCRATCH_BYTES];

namespace {

struct ScratchSlot {
  std::once_flag once;
  void* ptr = nullptr;
  hipError_t err = hipSuccess;
};
ScratchSlot g_scratch[kMaxDevices];

int scratch_address(int dev, void** ptr) {
  ScratchSlot& sl = g_scratch[dev];
  std::call_once(sl.once, [&sl, dev] {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) cur = -1;
    if (cur != dev && (sl.err = hipSetDevice(dev)) != hipSuccess) return;
    void* p = nullptr;
    sl.err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (sl.err == hipSuccess) sl.ptr = p;
    if (cur >= 0 && cur != dev) (void)hipSetDevice(cur);
  });
  if (sl.err != hipSuccess || !sl.ptr)
    return fail(GSC_ERR_HIP, "device scratch of device %d is unavailable: %s", dev,
                hipGetErrorString(sl.err != hipSuccess ? sl.err : hipErrorInvalidSymbol));
  *ptr = sl.ptr;
  return GSC_OK;
}

// ---- structured grids -------------------------------------------------------

int validate_grid(int dim, int64_t M, const int64_t* shape, const void* axes) {
  if (dim < 1 || dim > 3)
    return fail(GSC_ERR_DIM, "grid entry points support dim 1..3 (got %d)", dim);
  if (M < 0) return fail(GSC_ERR_INVALID, "negative M");
  if (M > (int64_t)0x7fffffff) return fail(GSC_ERR_INVALID, "M=%lld exceeds 2^31-1", (long long)M);
  if (!shape || !axes) return fail(GSC_ERR_INVALID, "null shape or axes");
  for (int i = 0; i < dim; ++i)
    if (shape[i] < 0) return fail(GSC_ERR_INVALID, "negative axis length");
  return GSC_OK;
}

int64_t grid_points(int dim, const int64_t* shape) {
  int64_t n = 1;
  for (int i = 0; i < dim; ++i) n *= shape[i];
  return n;
}

// Path choice: the factored contraction needs a 2-D or 3-D grid; below a
// 16-point last axis or 16 rows its GEMM block tiles are mostly padding.
// Without modes the field is zero: the direct kernels write those zeros (the
// factored tables would hold nothing, so that path is never taken for M = 0,
// whatever was asked).
int choose_grid_path(int dim, int64_t M, const int64_t* shape, const gsc_opts* o) {
  if (dim == 1 || M == 0) return GSC_PATH_DIRECT;
  if (o->path == GSC_PATH_DIRECT || o->path == GSC_PATH_FACTORED) return o->path;
  const int64_t cols = shape[dim - 1];
  const int64_t rows = grid_points(dim, shape) / (cols > 0 ? cols : 1);
  return (cols >= 16 && rows >= 16) ? GSC_PATH_FACTORED : GSC_PATH_DIRECT;
}

// Field kinds of the grid entry points (gsc_field) and their mode kinds.
int field_kind(int field) {
  return field == GSC_FIELD_INCOMPR ? gsc::KIND_INCOMPR
                                    : (field == GSC_FIELD_FOURIER ? gsc::KIND_FOURIER
                                                                  : gsc::KIND_SCALAR);
}
int field_comps(int field, int dim) { return field == GSC_FIELD_INCOMPR ? dim : 1; }

// Workspace of a grid call.  Factored: the plan (mode table and per-axis
// tables; the incompressible field has dim component blocks of the last-axis
// table).  Direct: the mode table; the incompressible field also expands the
// grid into a (dim, N) point list and runs the point-list kernel on it.
size_t grid_work_bytes(int field, int dim, int64_t M, const int64_t* shape, const gsc_opts* o,
                       int path) {
  if (path == GSC_PATH_FACTORED)
    return gsc::plan_factored(dim, M, shape, o->precision, field_comps(field, dim),
                              grid_tile(o)).bytes;
  if (field != GSC_FIELD_INCOMPR) return gsc::mode_table_bytes(dim, M);
  const int64_t N = grid_points(dim, shape);
  return align256(sizeof(double) * (size_t)dim * (size_t)N) +
         direct_work_bytes(gsc::KIND_INCOMPR, dim, M, N, o);
}

// True when x[i] = org + i step for every i < n within 8 units in the last
// place of the axis' largest magnitude (np.arange and np.linspace axes): the
// product form evaluates the coordinates from org and step, so a phase moves
// by at most |k| 8 eps max|x| revolutions.
bool uniform_axis(const double* x, int64_t n, double* org, double* step) {
  if (n <= 0 || !std::isfinite(x[0]) || !std::isfinite(x[n - 1])) return false;
  const double d = n > 1 ? (x[n - 1] - x[0]) / (double)(n - 1) : 0.0;
  const double tol = 8.0 * DBL_EPSILON * std::max(std::fabs(x[0]), std::fabs(x[n - 1]));
  for (int64_t i = 1; i + 1 < n; ++i)
    if (!(std::fabs(x[i] - (x[0] + (double)i * d)) <= tol)) return false;
  *org = x[0];
  *step = d;
  return std::isfinite(d);
}

// The plan of a factored grid call: the product form when the grid is 2-D,
// one field component, both axes uniform (known on the host: haxes) and its
// tables take no more room than the whole-grid ones (every grid of more than
// one tile); the whole-grid tables otherwise.  Ensembles use the same rule,
// so a realisation equals its single-grid call.
gsc::FactoredPlan grid_plan(int field, int dim, int64_t M, const int64_t* shape,
                            const gsc_opts* o, const double* const* haxes) {
  const gsc::FactoredPlan p =
      gsc::plan_factored(dim, M, shape, o->precision, field_comps(field, dim), grid_tile(o));
  if (dim != 2 || p.ncomp != 1 || haxes == nullptr) return p;
  const int r = p.trans ? 1 : 0;
  double org[2], step[2];
  if (!uniform_axis(haxes[r], shape[r], &org[0], &step[0]) ||
      !uniform_axis(haxes[1 - r], shape[1 - r], &org[1], &step[1]))
    return p;
  const gsc::FactoredPlan q = gsc::plan_factored_product(
      p, gsc::gemm_tile_cols(gsc::precision_splits(o->precision), o->accum), org, step);
  return q.bytes <= p.bytes ? q : p;
}

// `work` starts with the mode table in every case.  haxes: the axes on the
// host when the caller has them (host-array entry points), else null.
int grid_launch(int field, int dim, int64_t M, const double* sf, const double* cov,
                const double* z1, const double* z2, const double* const* axes,
                const int64_t* shape, void* out, void* work, const gsc_opts* o, int path,
                hipStream_t s, const double* const* haxes) {
  if (path == GSC_PATH_FACTORED) {
    // the factored tables take the amplitude A (sf A for the Fourier method)
    // from amp[0]; the projector of the incompressible field is applied to the
    // last-axis table
    const int kind = field == GSC_FIELD_FOURIER ? gsc::KIND_FOURIER : gsc::KIND_SCALAR;
    GSC_HIP(gsc::launch_prep(kind, dim, M, cov, z1, z2, sf, work, s));
    const gsc::FactoredPlan p = grid_plan(field, dim, M, shape, o, haxes);
    GSC_HIP(gsc::launch_factored(p, M, axes, work, out, o->out_dtype, o->scale, o->accum, s));
    return GSC_OK;
  }
  if (field != GSC_FIELD_INCOMPR) {
    GSC_HIP(gsc::launch_prep(field_kind(field), dim, M, cov, z1, z2, sf, work, s));
    GSC_HIP(gsc::launch_direct_grid(dim, M, shape, axes, work, out, o->out_dtype, o->scale,
                                    resolve_variant(o, M), s));
    return GSC_OK;
  }
  const int64_t N = grid_points(dim, shape);
  const size_t b_pos = align256(sizeof(double) * (size_t)dim * (size_t)N);
  double* pos = static_cast<double*>(work);
  void* dwork = static_cast<char*>(work) + b_pos;
  GSC_HIP(gsc::launch_expand_grid(dim, shape, axes, pos, s));
  GSC_HIP(launch_prep_direct(gsc::KIND_INCOMPR, dim, M, N, cov, z1, z2, nullptr, dwork, pos, N,
                             out, N, o, s));
  return GSC_OK;
}

int validate_field(int field, int dim, int64_t M, const double* sf) {
  if (field < GSC_FIELD_SCALAR || field > GSC_FIELD_FOURIER)
    return fail(GSC_ERR_INVALID, "unknown field kind %d", field);
  if (field == GSC_FIELD_INCOMPR && dim != 2 && dim != 3)
    return fail(GSC_ERR_DIM, "Only two- and three-dimensional problems are supported.");
  if (field == GSC_FIELD_FOURIER && M > 0 && !sf)
    return fail(GSC_ERR_INVALID, "null spectrum_factor array");
  return GSC_OK;
}

int validate_precision(const gsc_opts* o) {
  if (o->precision < GSC_PREC_AUTO || o->precision > GSC_PREC_BF16X1)
    return fail(GSC_ERR_INVALID, "unknown precision %d", o->precision);
  if (o->path < GSC_PATH_AUTO || o->path > GSC_PATH_FACTORED)
    return fail(GSC_ERR_INVALID, "unknown path %d", o->path);
  if (o->accum < GSC_ACCUM_F64 || o->accum > GSC_ACCUM_F32)
    return fail(GSC_ERR_INVALID, "unknown accumulation mode %d", o->accum);
  return GSC_OK;
}

// ---- kriging -------------------------------------------------------------------

int validate_krige(int64_t n, int64_t N, const void* mat, const void* vecs, const void* cond,
                   const void* field) {
  if (n < 0 || N < 0) return fail(GSC_ERR_INVALID, "negative n or N");
  if (n > ((int64_t)1 << 24)) return fail(GSC_ERR_INVALID, "n=%lld is too large", (long long)n);
  if (N > ((int64_t)1 << 40)) return fail(GSC_ERR_INVALID, "N=%lld is too large", (long long)N);
  if (n > 0 && (!mat || !cond)) return fail(GSC_ERR_INVALID, "null krige_mat or cond");
  if (N > 0 && !field) return fail(GSC_ERR_INVALID, "null field");
  if (n > 0 && N > 0 && !vecs) return fail(GSC_ERR_INVALID, "null krig_vecs");
  return GSC_OK;
}

// Host form of the two kriging calculators (error == null: field only).
int host_krige(int64_t n, int64_t N, const double* mat, int64_t m0, int64_t m1,
               const double* vecs, int64_t v0, int64_t v1, const double* cond, int64_t cs,
               double* field, int64_t fs, double* error, int64_t es, const gsc_opts* opts_in) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts_in, &tmp);
  int rc = validate_krige(n, N, mat, vecs, cond, field);
  if (rc) return rc;
  if ((rc = validate_opts(o))) return rc;
  if (N == 0) return GSC_OK;
  try {
    std::vector<double> hm((size_t)std::max<int64_t>(n * n, 1)), hc((size_t)std::max<int64_t>(n, 1));
    std::vector<double> hv;
    const double* vsrc = vecs;
    if (n > 0) {
      gather2(mat, n, n, m0, m1, hm.data(), 1);
      gather1(cond, n, cs, hc.data());
      if (!(v1 == 1 && v0 == N)) {
        hv.resize((size_t)(n * N));
        gather2(vecs, n, N, v0, v1, hv.data(), o->num_threads);
        vsrc = hv.data();
      }
    }
    int dev = 0;
    if ((rc = resolve_device(o, &dev))) return rc;
    DeviceCtx& c = g_dev[dev];
    std::lock_guard<std::mutex> lock(c.mu);
    GSC_HIP(hipSetDevice(dev));
    if ((rc = check_arch(dev))) return rc;
    if (!c.stream) GSC_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    const int nout = error ? 2 : 1;
    const size_t b_work = align256(gsc::krige_work_bytes(n, N, error != nullptr));
    const size_t b_mat = align256(sizeof(double) * hm.size());
    const size_t b_cond = align256(sizeof(double) * hc.size());
    const size_t b_vecs = align256(sizeof(double) * (size_t)std::max<int64_t>(n * N, 1));
    const size_t b_out = align256(sizeof(double) * nout * N);
    const size_t need = b_work + b_mat + b_cond + b_vecs + b_out;
    if (need > c.cap) {
      if (c.buf) GSC_HIP(hipFree(c.buf));
      c.buf = nullptr;
      c.cap = 0;
      if (hipMalloc(&c.buf, need) != hipSuccess)
        return fail(GSC_ERR_ALLOC, "hipMalloc of %zu bytes failed", need);
      c.cap = need;
    }
    char* base = static_cast<char*>(c.buf);
    void* d_work = base;
    double* d_mat = reinterpret_cast<double*>(base + b_work);
    double* d_cond = reinterpret_cast<double*>(base + b_work + b_mat);
    double* d_vecs = reinterpret_cast<double*>(base + b_work + b_mat + b_cond);
    double* d_out = reinterpret_cast<double*>(base + b_work + b_mat + b_cond + b_vecs);
    if (n > 0) {
      GSC_HIP(hipMemcpyAsync(d_mat, hm.data(), sizeof(double) * n * n, hipMemcpyHostToDevice,
                             c.stream));
      GSC_HIP(hipMemcpyAsync(d_cond, hc.data(), sizeof(double) * n, hipMemcpyHostToDevice,
                             c.stream));
      GSC_HIP(hipMemcpyAsync(d_vecs, vsrc, sizeof(double) * n * N, hipMemcpyHostToDevice,
                             c.stream));
    }
    GSC_HIP(gsc::launch_krige(n, N, d_mat, d_vecs, N, d_cond, d_out, 1, error ? d_out + N : nullptr,
                              1, d_work, c.stream));
    std::vector<double> res((size_t)(nout * N));
    GSC_HIP(hipMemcpyAsync(res.data(), d_out, sizeof(double) * nout * N, hipMemcpyDeviceToHost,
                           c.stream));
    GSC_HIP(hipStreamSynchronize(c.stream));
    for (int64_t j = 0; j < N; ++j) field[j * fs] = res[(size_t)j];
    if (error)
      for (int64_t j = 0; j < N; ++j) error[j * es] = res[(size_t)(N + j)];
  } catch (const std::bad_alloc&) {
    return fail(GSC_ERR_ALLOC, "host allocation failed");
  } catch (...) {
    return fail(GSC_ERR_INVALID, "unexpected host exception");
  }
  return GSC_OK;
}

}  // namespace

extern "C" {

int gsc_krige(int64_t n, int64_t N, const double* mat, int64_t mat_s0, int64_t mat_s1,
              const double* vecs, int64_t vecs_s0, int64_t vecs_s1, const double* cond,
              int64_t cond_s, double* field, int64_t field_s, double* error, int64_t error_s,
              const gsc_opts* opts) {
  return host_krige(n, N, mat, mat_s0, mat_s1, vecs, vecs_s0, vecs_s1, cond, cond_s, field,
                    field_s, error, error_s, opts);
}

size_t gsc_krige_workspace_bytes(int64_t n, int64_t N, int with_error) {
  if (n < 0 || N < 0) return 0;
  return gsc::krige_work_bytes(n, N, with_error != 0);
}

int gsc_krige_dev(int64_t n, int64_t N, const double* mat, const double* vecs, int64_t vecs_ld,
                  const double* cond, double* field, double* error, void* work,
                  size_t work_bytes, const gsc_opts* opts, void* stream) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  int rc = validate_krige(n, N, mat, vecs, cond, field);
  if (rc) return rc;
  if ((rc = validate_opts(o))) return rc;
  if (n > 0 && N > 0 && vecs_ld < N)
    return fail(GSC_ERR_INVALID, "vecs_ld (%lld) < N (%lld)", (long long)vecs_ld, (long long)N);
  const size_t need = gsc::krige_work_bytes(n, N, error != nullptr);
  if (!work || work_bytes < need)
    return fail(GSC_ERR_WORKSPACE, "workspace of %zu bytes is below the %zu required", work_bytes,
                need);
  if (N == 0) return GSC_OK;
  GSC_HIP(gsc::launch_krige(n, N, mat, vecs, vecs_ld, cond, field, 1, error, 1, work,
                            static_cast<hipStream_t>(stream)));
  return GSC_OK;
}

void gsc_opts_init(gsc_opts* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->out_dtype = GSC_F64;
  o->path = GSC_PATH_AUTO;
  o->precision = GSC_PREC_AUTO;
  o->device = -1;
  o->num_threads = 0;
  o->accum = GSC_ACCUM_F32X2;
  o->direct_variant = GSC_DIRECT_AUTO;
  o->scale = 1.0;
  o->tile = GSC_TILE_AUTO;
}

const char* gsc_version(void) { return "1.1.0-mi355x.2"; }

const char* gsc_last_error(void) { return g_last_error.c_str(); }

int gsc_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

size_t gsc_workspace_bytes(int dim, int64_t M, int64_t N, const gsc_opts* opts) {
  if (dim < 1 || M < 0 || N < 0) return 0;
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  // the scalar and Fourier forms need the same space; incompr never more
  return direct_work_bytes(gsc::KIND_SCALAR, dim, M, N, o);
}

int gsc_summate(int dim, int64_t M, int64_t N, const double* cov, int64_t cov_sd, int64_t cov_sm,
                const double* z1, int64_t z1_s, const double* z2, int64_t z2_s, const double* pos,
                int64_t pos_sd, int64_t pos_sp, void* out, int64_t out_s, const gsc_opts* opts) {
  return host_summate(gsc::KIND_SCALAR, dim, M, N, nullptr, 0, cov, cov_sd, cov_sm, z1, z1_s, z2,
                      z2_s, pos, pos_sd, pos_sp, out, 0, out_s, opts);
}

int gsc_summate_incompr(int dim, int64_t M, int64_t N, const double* cov, int64_t cov_sd,
                        int64_t cov_sm, const double* z1, int64_t z1_s, const double* z2,
                        int64_t z2_s, const double* pos, int64_t pos_sd, int64_t pos_sp, void* out,
                        int64_t out_sd, int64_t out_sp, const gsc_opts* opts) {
  return host_summate(gsc::KIND_INCOMPR, dim, M, N, nullptr, 0, cov, cov_sd, cov_sm, z1, z1_s,
                      z2, z2_s, pos, pos_sd, pos_sp, out, out_sd, out_sp, opts);
}

int gsc_summate_fourier(int dim, int64_t M, int64_t N, const double* sf, int64_t sf_s,
                        const double* modes, int64_t md_sd, int64_t md_sm, const double* z1,
                        int64_t z1_s, const double* z2, int64_t z2_s, const double* pos,
                        int64_t pos_sd, int64_t pos_sp, void* out, int64_t out_s,
                        const gsc_opts* opts) {
  return host_summate(gsc::KIND_FOURIER, dim, M, N, sf, sf_s, modes, md_sd, md_sm, z1, z1_s, z2,
                      z2_s, pos, pos_sd, pos_sp, out, 0, out_s, opts);
}

int gsc_summate_dev(int dim, int64_t M, int64_t N, const double* cov, const double* z1,
                    const double* z2, const double* pos, int64_t pos_ld, void* out, void* work,
                    size_t work_bytes, const gsc_opts* opts, void* stream) {
  return dev_common(gsc::KIND_SCALAR, dim, M, N, nullptr, cov, z1, z2, pos, pos_ld, out, 0, work,
                    work_bytes, opts, stream);
}

int gsc_summate_dev_noalloc(int dim, int64_t M, int64_t N, const double* cov, const double* z1,
                            const double* z2, const double* pos, int64_t pos_ld, void* out,
                            const gsc_opts* opts, void* stream) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  int rc = validate_common(dim, M, N, cov, z1, z2, pos, out);
  if (rc || (rc = validate_opts(o))) return rc;
  const size_t need = direct_work_bytes(gsc::KIND_SCALAR, dim, M, N, o);
  if (need > (size_t)GSC_NOALLOC_SCRATCH_BYTES)
    return fail(GSC_ERR_WORKSPACE,
                "this call needs %zu bytes of device scratch, above the %zu bytes the "
                "no-allocation entry point holds (use gsc_summate_dev with a workspace)",
                need, (size_t)GSC_NOALLOC_SCRATCH_BYTES);
  if (N == 0) return GSC_OK;
  int dev = 0;
  if ((rc = resolve_device(o, &dev)) || (rc = check_arch(dev))) return rc;
  void* work = nullptr;
  if ((rc = scratch_address(dev, &work))) return rc;
  return dev_common(gsc::KIND_SCALAR, dim, M, N, nullptr, cov, z1, z2, pos, pos_ld, out, 0, work,
                    (size_t)GSC_NOALLOC_SCRATCH_BYTES, o, stream);
}

int gsc_direct_phase_times(int field, int dim, int64_t M, int64_t N, const double* sf,
                           const double* cov, const double* z1, const double* z2,
                           const double* pos, int64_t pos_ld, void* out, int64_t out_ld,
                           void* work, size_t work_bytes, const gsc_opts* opts, void* stream,
                           int repeats, double* ms) {
  if (field < GSC_FIELD_SCALAR || field > GSC_FIELD_FOURIER)
    return fail(GSC_ERR_INVALID, "unknown field kind %d", field);
  if (repeats < 1 || repeats > 10000 || !ms)
    return fail(GSC_ERR_INVALID, "repeats must be in 1..10000 with a result array");
  const int kind = field_kind(field);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // events: start, split and end of every call
  std::vector<hipEvent_t> ev((size_t)(3 * repeats), nullptr);
  hipError_t e = hipSuccess;
  for (auto& x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  int rc = GSC_OK;
  for (int r = 0; r < repeats && e == hipSuccess && rc == GSC_OK; ++r) {
    e = hipEventRecord(ev[(size_t)(3 * r)], s);
    if (e == hipSuccess)
      rc = dev_common(kind, dim, M, N, sf, cov, z1, z2, pos, pos_ld, out, out_ld, work,
                      work_bytes, opts, stream, ev[(size_t)(3 * r + 1)]);
    if (e == hipSuccess && rc == GSC_OK) e = hipEventRecord(ev[(size_t)(3 * r + 2)], s);
  }
  double acc[3] = {0.0, 0.0, 0.0};
  if (e == hipSuccess && rc == GSC_OK) e = hipEventSynchronize(ev.back());
  for (int r = 0; r < repeats && e == hipSuccess && rc == GSC_OK; ++r) {
    float a = 0.0f, b = 0.0f;
    e = hipEventElapsedTime(&a, ev[(size_t)(3 * r)], ev[(size_t)(3 * r + 1)]);
    if (e == hipSuccess) e = hipEventElapsedTime(&b, ev[(size_t)(3 * r + 1)], ev[(size_t)(3 * r + 2)]);
    acc[0] += a;
    acc[1] += b;
    acc[2] += a + b;
  }
  for (auto& x : ev)
    if (x) (void)hipEventDestroy(x);
  if (rc != GSC_OK) return rc;
  if (e != hipSuccess) return fail(GSC_ERR_HIP, "phase timing failed: %s", hipGetErrorString(e));
  for (int i = 0; i < 3; ++i) ms[i] = acc[i] / repeats;
  return GSC_OK;
}

int gsc_summate_incompr_dev(int dim, int64_t M, int64_t N, const double* cov, const double* z1,
                            const double* z2, const double* pos, int64_t pos_ld, void* out,
                            int64_t out_ld, void* work, size_t work_bytes, const gsc_opts* opts,
                            void* stream) {
  return dev_common(gsc::KIND_INCOMPR, dim, M, N, nullptr, cov, z1, z2, pos, pos_ld, out, out_ld,
                    work, work_bytes, opts, stream);
}

int gsc_summate_fourier_dev(int dim, int64_t M, int64_t N, const double* sf, const double* modes,
                            const double* z1, const double* z2, const double* pos, int64_t pos_ld,
                            void* out, void* work, size_t work_bytes, const gsc_opts* opts,
                            void* stream) {
  return dev_common(gsc::KIND_FOURIER, dim, M, N, sf, modes, z1, z2, pos, pos_ld, out, 0, work,
                    work_bytes, opts, stream);
}

int gsc_grid_path(int dim, int64_t M, const int64_t* shape, const gsc_opts* opts) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  if (validate_grid(dim, M, shape, shape) || validate_precision(o)) return -1;
  const int path = choose_grid_path(dim, M, shape, o);
  if (path == GSC_PATH_FACTORED && dim < 2) return -1;
  return path;
}

size_t gsc_grid_field_workspace_bytes(int field, int dim, int64_t M, const int64_t* shape,
                                      const gsc_opts* opts) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  if (validate_grid(dim, M, shape, shape) || validate_precision(o) ||
      field < GSC_FIELD_SCALAR || field > GSC_FIELD_FOURIER ||
      (field == GSC_FIELD_INCOMPR && dim != 2 && dim != 3))
    return 0;
  return grid_work_bytes(field, dim, M, shape, o, choose_grid_path(dim, M, shape, o));
}

size_t gsc_grid_workspace_bytes(int dim, int64_t M, const int64_t* shape, const gsc_opts* opts) {
  return gsc_grid_field_workspace_bytes(GSC_FIELD_SCALAR, dim, M, shape, opts);
}

int gsc_summate_grid_field(int field, int dim, int64_t M, const double* sf, int64_t sf_s,
                           const double* cov, int64_t cov_sd, int64_t cov_sm, const double* z1,
                           int64_t z1_s, const double* z2, int64_t z2_s,
                           const double* const* axes, const int64_t* axis_s,
                           const int64_t* shape, void* out, const gsc_opts* opts) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  int rc = validate_grid(dim, M, shape, axes);
  if (rc || (rc = validate_opts(o)) || (rc = validate_precision(o)) ||
      (rc = validate_field(field, dim, M, sf)))
    return rc;
  if (M > 0 && (!cov || !z1 || !z2)) return fail(GSC_ERR_INVALID, "null mode array");
  const int path = choose_grid_path(dim, M, shape, o);
  if (path == GSC_PATH_FACTORED && dim < 2)
    return fail(GSC_ERR_INVALID, "the factored path needs a 2-D or 3-D grid");
  const int64_t N = grid_points(dim, shape);
  if (N == 0) return GSC_OK;
  if (!out) return fail(GSC_ERR_INVALID, "null output");
  for (int i = 0; i < dim; ++i)
    if (!axes[i]) return fail(GSC_ERR_INVALID, "null axis %d", i);
  const size_t elt = o->out_dtype == GSC_F32 ? 4 : 8;
  const int ncomp = field_comps(field, dim);
  try {
    const int64_t Mm = std::max<int64_t>(M, 1);
    std::vector<double> modes((size_t)((dim + 3) * Mm), 0.0);
    if (M > 0) {
      gather2(cov, dim, M, cov_sd, cov_sm, modes.data(), 1);
      gather1(z1, M, z1_s, modes.data() + dim * M);
      gather1(z2, M, z2_s, modes.data() + (dim + 1) * M);
      if (field == GSC_FIELD_FOURIER) gather1(sf, M, sf_s, modes.data() + (dim + 2) * M);
    }
    int64_t ax_total = 0;
    for (int i = 0; i < dim; ++i) ax_total += shape[i];
    std::vector<double> ax((size_t)std::max<int64_t>(ax_total, 1));
    int64_t off = 0;
    for (int i = 0; i < dim; ++i) {
      gather1(axes[i], shape[i], axis_s ? axis_s[i] : 1, ax.data() + off);
      off += shape[i];
    }
    int dev = 0;
    if ((rc = resolve_device(o, &dev))) return rc;
    DeviceCtx& c = g_dev[dev];
    std::lock_guard<std::mutex> lock(c.mu);
    GSC_HIP(hipSetDevice(dev));
    if ((rc = check_arch(dev))) return rc;
    if (!c.stream) GSC_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    const size_t b_work = align256(grid_work_bytes(field, dim, M, shape, o, path));
    const size_t b_modes = align256(sizeof(double) * modes.size());
    const size_t b_ax = align256(sizeof(double) * ax.size());
    const size_t b_out = align256(elt * (size_t)ncomp * (size_t)N);
    const size_t need = b_work + b_modes + b_ax + b_out;
    if (need > c.cap) {
      if (c.buf) GSC_HIP(hipFree(c.buf));
      c.buf = nullptr;
      c.cap = 0;
      if (hipMalloc(&c.buf, need) != hipSuccess)
        return fail(GSC_ERR_ALLOC, "hipMalloc of %zu bytes failed", need);
      c.cap = need;
    }
    char* base = static_cast<char*>(c.buf);
    void* d_work = base;
    double* d_modes = reinterpret_cast<double*>(base + b_work);
    double* d_ax = reinterpret_cast<double*>(base + b_work + b_modes);
    void* d_out = base + b_work + b_modes + b_ax;
    GSC_HIP(hipMemcpyAsync(d_modes, modes.data(), sizeof(double) * modes.size(),
                           hipMemcpyHostToDevice, c.stream));
    GSC_HIP(hipMemcpyAsync(d_ax, ax.data(), sizeof(double) * ax.size(), hipMemcpyHostToDevice,
                           c.stream));
    const double* d_axes[3] = {nullptr, nullptr, nullptr};
    const double* h_axes[3] = {nullptr, nullptr, nullptr};
    off = 0;
    for (int i = 0; i < dim; ++i) {
      d_axes[i] = d_ax + off;
      h_axes[i] = ax.data() + off;
      off += shape[i];
    }
    if ((rc = grid_launch(field, dim, M, d_modes + (dim + 2) * M, d_modes, d_modes + dim * M,
                          d_modes + (dim + 1) * M, d_axes, shape, d_out, d_work, o, path,
                          c.stream, h_axes)))
      return rc;
    GSC_HIP(hipMemcpyAsync(out, d_out, elt * (size_t)ncomp * (size_t)N, hipMemcpyDeviceToHost,
                           c.stream));
    GSC_HIP(hipStreamSynchronize(c.stream));
  } catch (const std::bad_alloc&) {
    return fail(GSC_ERR_ALLOC, "host allocation failed");
  } catch (...) {
    return fail(GSC_ERR_INVALID, "unexpected host exception");
  }
  return GSC_OK;
}

int gsc_summate_grid(int dim, int64_t M, const double* cov, int64_t cov_sd, int64_t cov_sm,
                     const double* z1, int64_t z1_s, const double* z2, int64_t z2_s,
                     const double* const* axes, const int64_t* axis_s, const int64_t* shape,
                     void* out, const gsc_opts* opts) {
  return gsc_summate_grid_field(GSC_FIELD_SCALAR, dim, M, nullptr, 0, cov, cov_sd, cov_sm, z1,
                                z1_s, z2, z2_s, axes, axis_s, shape, out, opts);
}

int gsc_summate_grid_field_dev(int field, int dim, int64_t M, const double* sf,
                               const double* cov, const double* z1, const double* z2,
                               const double* const* axes, const int64_t* shape, void* out,
                               void* work, size_t work_bytes, const gsc_opts* opts,
                               void* stream) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  int rc = validate_grid(dim, M, shape, axes);
  if (rc || (rc = validate_opts(o)) || (rc = validate_precision(o)) ||
      (rc = validate_field(field, dim, M, sf)))
    return rc;
  if (M > 0 && (!cov || !z1 || !z2)) return fail(GSC_ERR_INVALID, "null mode array");
  const int path = choose_grid_path(dim, M, shape, o);
  if (path == GSC_PATH_FACTORED && dim < 2)
    return fail(GSC_ERR_INVALID, "the factored path needs a 2-D or 3-D grid");
  const size_t need = grid_work_bytes(field, dim, M, shape, o, path);
  if (!work || work_bytes < need)
    return fail(GSC_ERR_WORKSPACE, "workspace of %zu bytes is below the %zu required", work_bytes,
                need);
  if (grid_points(dim, shape) == 0) return GSC_OK;
  if (!out) return fail(GSC_ERR_INVALID, "null output");
  for (int i = 0; i < dim; ++i)
    if (!axes[i]) return fail(GSC_ERR_INVALID, "null axis %d", i);
  return grid_launch(field, dim, M, sf, cov, z1, z2, axes, shape, out, work, o, path,
                     static_cast<hipStream_t>(stream), nullptr);
}

int gsc_tune_grid_time(int field, int dim, int64_t M, const int64_t* shape, const gsc_opts* opts,
                       int repeats, double* ms) {
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  static const double kOwnFactors = 1.0;  // the run supplies its own spectrum factors
  int rc = validate_grid(dim, M, shape, shape);
  if (rc || (rc = validate_opts(o)) || (rc = validate_precision(o)) ||
      (rc = validate_field(field, dim, M, &kOwnFactors)))
    return rc;
  if (!ms) return fail(GSC_ERR_INVALID, "null result pointer");
  if (repeats < 1 || repeats > 10000) return fail(GSC_ERR_INVALID, "repeats must be in 1..10000");
  const int path = choose_grid_path(dim, M, shape, o);
  if (path == GSC_PATH_FACTORED && dim < 2)
    return fail(GSC_ERR_INVALID, "the factored path needs a 2-D or 3-D grid");
  const int64_t N = grid_points(dim, shape);
  *ms = 0.0;
  if (N == 0) return GSC_OK;
  int dev = 0;
  if ((rc = resolve_device(o, &dev))) return rc;
  GSC_HIP(hipSetDevice(dev));
  if ((rc = check_arch(dev))) return rc;
  const size_t elt = o->out_dtype == GSC_F32 ? 4 : 8;
  const int64_t Mm = std::max<int64_t>(M, 1);
  int64_t ax_total = 0;
  for (int i = 0; i < dim; ++i) ax_total += shape[i];
  const size_t b_modes = align256(sizeof(double) * (size_t)(dim + 3) * (size_t)Mm);
  const size_t b_ax = align256(sizeof(double) * (size_t)std::max<int64_t>(ax_total, 1));
  const size_t b_out = align256(elt * (size_t)field_comps(field, dim) * (size_t)N);
  const size_t b_work = align256(grid_work_bytes(field, dim, M, shape, o, path));
  char* mem = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (hipMalloc(&mem, b_modes + b_ax + b_out + b_work) != hipSuccess)
    return fail(GSC_ERR_ALLOC, "hipMalloc of %zu bytes failed", b_modes + b_ax + b_out + b_work);
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
  double* modes = reinterpret_cast<double*>(mem);
  double* ax = reinterpret_cast<double*>(mem + b_modes);
  // unit-spaced axes (kept on the host for the plan choice) and Gaussian
  // modes drawn on the device (length scale an eighth of the largest extent);
  // Fourier factors of one
  std::vector<double> hax((size_t)std::max<int64_t>(ax_total, 1));
  if (e == hipSuccess) {
    std::vector<double> ones((size_t)Mm, 1.0);
    int64_t off = 0, ext = 1;
    for (int i = 0; i < dim; ++i) {
      for (int64_t k = 0; k < shape[i]; ++k) hax[(size_t)(off + k)] = (double)k;
      off += shape[i];
      ext = std::max(ext, shape[i]);
    }
    e = hipMemcpy(ax, hax.data(), sizeof(double) * hax.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(modes + (dim + 2) * Mm, ones.data(), sizeof(double) * ones.size(),
                    hipMemcpyHostToDevice);
    if (e == hipSuccess && M > 0) {
      gsc::SampleArgs a{};
      a.model = GSC_MODEL_GAUSSIAN;
      a.dim = dim;
      a.count = 1;
      a.m_total = M;
      a.len_scale1 = (double)ext / 8.0;
      a.seed1 = 19031977;
      a.cov = modes;
      a.z1 = modes + dim * M;
      a.z2 = modes + (dim + 1) * M;
      e = gsc::launch_sample_modes(a, st);
    }
  }
  if (e == hipSuccess) {
    const double* axes[3] = {nullptr, nullptr, nullptr};
    int64_t off = 0;
    for (int i = 0; i < dim; ++i) {
      axes[i] = ax + off;
      off += shape[i];
    }
    void* out = mem + b_modes + b_ax;
    void* work = mem + b_modes + b_ax + b_out;
    // the host copy of the unit axes: the plan host-array calls would take
    const double* haxes[3] = {nullptr, nullptr, nullptr};
    off = 0;
    for (int i = 0; i < dim; ++i) {
      haxes[i] = hax.data() + off;
      off += shape[i];
    }
    for (int r = 0; r <= repeats && rc == GSC_OK; ++r) {
      if (r == 1 && (e = hipEventRecord(ev[0], st)) != hipSuccess) break;
      rc = grid_launch(field, dim, M, modes + (dim + 2) * Mm, modes, modes + dim * M,
                       modes + (dim + 1) * M, axes, shape, out, work, o, path, st, haxes);
    }
    if (rc == GSC_OK && e == hipSuccess) e = hipEventRecord(ev[1], st);
    if (rc == GSC_OK && e == hipSuccess) e = hipEventSynchronize(ev[1]);
    float t = 0.0f;
    if (rc == GSC_OK && e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
    if (rc == GSC_OK && e == hipSuccess) *ms = (double)t / repeats;
  }
  if (st) (void)hipStreamSynchronize(st);
  for (int i = 0; i < 2; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (st) (void)hipStreamDestroy(st);
  (void)hipFree(mem);
  if (rc != GSC_OK) return rc;
  if (e != hipSuccess) return fail(GSC_ERR_HIP, "tuning run failed: %s", hipGetErrorString(e));
  return GSC_OK;
}

int gsc_summate_grid_dev(int dim, int64_t M, const double* cov, const double* z1,
                         const double* z2, const double* const* axes, const int64_t* shape,
                         void* out, void* work, size_t work_bytes, const gsc_opts* opts,
                         void* stream) {
  return gsc_summate_grid_field_dev(GSC_FIELD_SCALAR, dim, M, nullptr, cov, z1, z2, axes, shape,
                                    out, work, work_bytes, opts, stream);
}

}  // extern "C"

// ---- ensembles ----------------------------------------------------------------

struct gsc_ensemble {
  int device = 0;
  void* mem = nullptr;
  size_t mem_bytes = 0;
  void* mode_mem = nullptr;   // seeded plans: modes, mode offsets, scales, seeds
  uint64_t* d_seeds = nullptr;
  int count = 0;
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t n_fact = 0, n_dir = 0;
  double point_modes = 0.0;
};

namespace {

void ensemble_release(gsc_ensemble* p) {
  if (!p) return;
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->cap) (void)hipStreamDestroy(p->cap);

  if (p->mem) (void)hipFree(p->mem);
  if (p->mode_mem) (void)hipFree(p->mode_mem);
  delete p;
}

// sample: null, or the mode sampler that fills cov / z1 / z2 at the start
// of every replay (seeded plans).  tiles: per-realisation gsc_tile of the
// factored ones, or null (opts->tile for all).
int ensemble_build(gsc_ensemble* plan, int count, int dim, const int64_t* shapes,
                   const int64_t* m_counts, const int32_t* paths, const int32_t* tiles,
                   const double* scales, const double* cov, const double* z1, const double* z2,
                   const double* axes, void* out, const gsc_opts* o,
                   const gsc::SampleArgs* sample) {
  const size_t elt = o->out_dtype == GSC_F32 ? 4 : 8;
  std::vector<gsc::EnsReal> rs((size_t)count);
  std::vector<int> fact, dir;
  std::vector<int> tile_of((size_t)count, GSC_TILE_256);
  std::vector<size_t> tab_off((size_t)count), f_off((size_t)count), fbytes((size_t)count, 0);
  std::vector<gsc::FactoredPlan> plans((size_t)count);
  const int tile_cols = gsc::gemm_tile_cols(gsc::precision_splits(o->precision), o->accum);
  // 2-D: the axes on the host, to find the uniform ones (product form)
  std::vector<double> hax;
  if (dim == 2) {
    int64_t total = 0;
    for (int64_t i = 0; i < (int64_t)count * dim; ++i) total += shapes[i];
    hax.resize((size_t)std::max<int64_t>(total, 1));
    if (total > 0)
      GSC_HIP(hipMemcpy(hax.data(), axes, sizeof(double) * (size_t)total, hipMemcpyDeviceToHost));
  }
  // pass 1: geometry, offsets into the caller's arrays, path, workspace sizes
  int64_t m_off = 0, a_off = 0, o_off = 0;
  size_t ws = 0;
  int max_tiles = 0;
  for (int r = 0; r < count; ++r) {
    gsc::EnsReal& e = rs[(size_t)r];
    std::memset(&e, 0, sizeof(e));
    const int64_t* sh = shapes + (int64_t)r * dim;
    const int64_t M = m_counts[r];
    e.cov = cov + dim * m_off;
    e.z1 = z1 + m_off;
    e.z2 = z2 + m_off;
    e.ax0 = axes + a_off;
    e.ax1 = axes + a_off + sh[0];
    e.ax2 = dim == 3 ? axes + a_off + sh[0] + sh[1] : nullptr;
    e.out = static_cast<char*>(out) + (size_t)o_off * elt;
    e.M = M;
    e.ntiles = (int)((M + 15) / 16);
    e.dim = dim;
    e.s0 = sh[0];
    e.s1 = sh[1];
    e.s2 = dim == 3 ? sh[2] : 1;
    e.npts = grid_points(dim, sh);
    e.scale = scales ? scales[r] : o->scale;
    gsc_opts po = *o;
    po.path = paths ? paths[r] : GSC_PATH_AUTO;
    if (tiles) po.tile = tiles[r];
    e.path = choose_grid_path(dim, M, sh, &po);
    m_off += M;
    for (int i = 0; i < dim; ++i) a_off += sh[i];
    o_off += e.npts;
    max_tiles = std::max(max_tiles, e.ntiles);
    tab_off[(size_t)r] = ws;
    ws += align256(gsc::mode_table_bytes(dim, M));
    if (e.npts == 0) continue;
    if (e.path == GSC_PATH_FACTORED) {
      tile_of[(size_t)r] = grid_tile(&po);
      const gsc::FactoredPlan fp =
          gsc::plan_factored(dim, M, sh, o->precision, 1, tile_of[(size_t)r]);
      e.trans = fp.trans;
      plans[(size_t)r] = fp;
      e.mp = fp.mp;
      e.na = fp.na;
      e.nrow = fp.nrow;
      e.arows = fp.arows;
      e.nb_out = fp.nb_out;
      e.ncols = fp.ncols;
      e.ncp = fp.ncp;
      e.row_tiles = fp.row_tiles;
      e.col_tiles = (fp.ncols + tile_cols - 1) / tile_cols;
      // S, A, C as laid out after the plan's own mode table, placed below
      // the single-call rule (grid_plan): the product form on uniform axes
      // when its tables are no larger
      gsc::FactoredPlan q = fp;
      if (dim == 2) {
        const double* hx[2] = {hax.data() + (e.ax0 - axes), hax.data() + (e.ax0 - axes) + sh[0]};
        gsc_opts po2 = *o;
        po2.tile = tile_of[(size_t)r];
        q = grid_plan(GSC_FIELD_SCALAR, dim, M, sh, &po2, hx);
      }
      if (q.prod) {
        // product form: S per tile, A and C one tile
        e.prod = 1;
        e.tcols = q.tcols;
        e.org[0] = q.org[0];
        e.org[1] = q.org[1];
        e.step[0] = q.step[0];
        e.step[1] = q.step[1];
        e.row_tiles = q.row_tiles;
        e.col_tiles = q.col_tiles;
        e.na = q.na;
        e.arows = q.arows;
        e.ncp = q.ncp;
        plans[(size_t)r].off_s = 0;  // offsets within the realisation's block
        plans[(size_t)r].off_a = q.off_a - q.off_s;
        plans[(size_t)r].off_c = q.off_c - q.off_s;
        fbytes[(size_t)r] = align256(q.bytes - q.off_s);
      } else {
        plans[(size_t)r].off_a -= fp.off_s;
        plans[(size_t)r].off_c -= fp.off_s;
        plans[(size_t)r].off_s = 0;
        fbytes[(size_t)r] = align256(fp.bytes - fp.off_s);
      }
      fact.push_back(r);
    } else {
      dir.push_back(r);
    }
    plan->point_modes += (double)e.npts * (double)M;
  }
  // table batches: a plan whose factored tables exceed the bound
  // (opts->table_bytes, default kEnsTableBytes) builds and consumes them in
  // batches of realisations that reuse one region of the workspace (the
  // batches run one after the other: tables, then GEMM).  The batch count
  // starts at total / bound and grows until a longest-first split of the
  // table sizes keeps every batch within the bound (a realisation larger than
  // the bound forms a batch of its own); the batches come out of about equal
  // size, so the region is about total / count.
  auto gemm_work = [&](int r) {
    const gsc::EnsReal& e = rs[(size_t)r];
    return (double)(e.row_tiles * e.col_tiles) * (double)e.mp;
  };
  size_t total_fb = 0;
  for (int r : fact) total_fb += fbytes[(size_t)r];
  const size_t cap = o->table_bytes > 0 ? (size_t)o->table_bytes : gsc::kEnsTableBytes;
  std::vector<std::vector<int>> bfact;
  if (total_fb <= cap) {
    bfact.push_back(fact);
  } else {
    std::vector<int> lst = fact;
    std::stable_sort(lst.begin(), lst.end(),
                     [&](int a, int b) { return fbytes[(size_t)a] > fbytes[(size_t)b]; });
    for (size_t n = std::min((total_fb + cap - 1) / cap, lst.size());; ++n) {
      bfact.assign(n, {});
      std::vector<size_t> used(n, 0);
      for (int r : lst) {
        const size_t bi = (size_t)(std::min_element(used.begin(), used.end()) - used.begin());
        used[bi] += fbytes[(size_t)r];
        bfact[bi].push_back(r);
      }
      bool fits = true;
      for (size_t bi = 0; bi < n; ++bi) fits = fits && (used[bi] <= cap || bfact[bi].size() == 1);
      if (fits || n >= lst.size()) break;
    }
  }
  const int nb = (int)bfact.size();
  size_t region_bytes = 0;
  for (auto& bl : bfact) {
    std::sort(bl.begin(), bl.end());
    size_t off = 0;
    for (int r : bl) off += fbytes[(size_t)r];
    region_bytes = std::max(region_bytes, off);
  }
  for (int bi = 0; bi < nb; ++bi) {
    size_t off = ws;
    for (int r : bfact[(size_t)bi]) {
      f_off[(size_t)r] = off;
      off += fbytes[(size_t)r];
    }
  }
  ws += region_bytes;
  // per batch: GEMM groups by tile shape (256-row tiles first) and operand
  // form (tables of the whole grid, then the product form); in each, longest
  // first (most modes, then most tiles), then a longest-processing-time split
  // into kXcdGroups groups of whole realisations (one per XCD), each keeping
  // that order; and the table tasks of every factored realisation, groups of
  // 64 rows of S (na, 3-D or product form), A (arows) and C (ncp)
  constexpr int NG = gsc::kEnsGemmGroups;
  const int tile_ids[2] = {GSC_TILE_256, GSC_TILE_128};
  struct Batch {
    std::vector<int> gemm_order[NG];
    std::vector<int64_t> tile_start[NG], group_start[NG];
    int ngemm = 0;
    std::vector<int> all_fact;
    std::vector<int64_t> task_start;
  };
  std::vector<Batch> bt((size_t)nb);
  for (int bi = 0; bi < nb; ++bi) {
    Batch& B = bt[(size_t)bi];
    for (int q = 0; q < NG; ++q) {
      std::vector<int> lst;
      for (int r : bfact[(size_t)bi])
        if (tile_of[(size_t)r] == tile_ids[q / 2] && rs[(size_t)r].prod == q % 2) lst.push_back(r);
      std::stable_sort(lst.begin(), lst.end(), [&](int a, int b) {
        const gsc::EnsReal &ea = rs[(size_t)a], &eb = rs[(size_t)b];
        if (ea.mp != eb.mp) return ea.mp > eb.mp;
        return ea.row_tiles * ea.col_tiles > eb.row_tiles * eb.col_tiles;
      });
      std::vector<std::vector<int>> groups(gsc::kXcdGroups);
      std::vector<double> load(gsc::kXcdGroups, 0.0);
      for (int r : lst) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        groups[(size_t)g].push_back(r);
        load[(size_t)g] += gemm_work(r);
      }
      std::vector<int>& ord = B.gemm_order[q];
      std::vector<int> glen;
      for (auto& g : groups) {
        ord.insert(ord.end(), g.begin(), g.end());
        glen.push_back((int)g.size());
      }
      B.tile_start[q].assign(ord.size() + 1, 0);
      for (size_t i = 0; i < ord.size(); ++i) {
        const gsc::EnsReal& e = rs[(size_t)ord[i]];
        B.tile_start[q][i + 1] = B.tile_start[q][i] + e.row_tiles * e.col_tiles;
      }
      B.group_start[q].assign(gsc::kXcdGroups + 1, 0);
      size_t k = 0;
      for (int g = 0; g < gsc::kXcdGroups; ++g) {
        k += (size_t)glen[(size_t)g];
        B.group_start[q][(size_t)g + 1] = B.tile_start[q][k];
      }
      if (!ord.empty()) B.ngemm = q + 1;
    }
    for (int q = 0; q < NG; ++q)
      B.all_fact.insert(B.all_fact.end(), B.gemm_order[q].begin(), B.gemm_order[q].end());
    B.task_start.assign(3 * B.all_fact.size() + 1, 0);
    for (size_t i = 0; i < B.all_fact.size(); ++i) {
      const gsc::EnsReal& e = rs[(size_t)B.all_fact[i]];
      B.task_start[3 * i + 1] = B.task_start[3 * i] + (dim == 3 || e.prod ? (e.na + 63) / 64 : 0);
      B.task_start[3 * i + 2] = B.task_start[3 * i + 1] + e.arows / 64;
      B.task_start[3 * i + 3] = B.task_start[3 * i + 2] + e.ncp / 64;
    }
  }
  std::stable_sort(dir.begin(), dir.end(), [&](int a, int b) {
    return rs[(size_t)a].M * rs[(size_t)a].npts > rs[(size_t)b].M * rs[(size_t)b].npts;
  });
  std::vector<int64_t> blk_start(dir.size() + 1, 0);
  for (size_t i = 0; i < dir.size(); ++i) {
    const gsc::EnsReal& e = rs[(size_t)dir[i]];
    blk_start[i + 1] = blk_start[i] + (e.npts + gsc::kEnsPoints - 1) / gsc::kEnsPoints;
  }
  plan->n_fact = (int64_t)fact.size();
  plan->n_dir = (int64_t)dir.size();
  // control block: descriptors, lists, prefix sums, counters (each piece
  // 256-byte aligned; offsets collected first, then one upload)
  std::vector<char> hctrl;
  // appends `bytes` from src (zeros when src is null) at a 256-byte boundary
  auto put = [&hctrl](const void* src, size_t bytes) {
    const size_t off = hctrl.size();
    hctrl.resize(off + align256(std::max<size_t>(bytes, 1)), 0);
    if (src != nullptr && bytes > 0) std::memcpy(hctrl.data() + off, src, bytes);
    return off;
  };
  const size_t c_desc = put(rs.data(), sizeof(gsc::EnsReal) * rs.size());
  const size_t c_dir = put(dir.data(), sizeof(int) * dir.size());
  const size_t c_bs = put(blk_start.data(), sizeof(int64_t) * blk_start.size());
  std::vector<size_t> c_fact((size_t)nb), c_tk((size_t)nb), c_go((size_t)nb * NG),
      c_ts((size_t)nb * NG), c_gs((size_t)nb * NG);
  for (int bi = 0; bi < nb; ++bi) {
    const Batch& B = bt[(size_t)bi];
    c_fact[(size_t)bi] = put(B.all_fact.data(), sizeof(int) * B.all_fact.size());
    c_tk[(size_t)bi] = put(B.task_start.data(), sizeof(int64_t) * B.task_start.size());
    for (int q = 0; q < NG; ++q) {
      const size_t i = (size_t)bi * NG + (size_t)q;
      c_go[i] = put(B.gemm_order[q].data(), sizeof(int) * B.gemm_order[q].size());
      c_ts[i] = put(B.tile_start[q].data(), sizeof(int64_t) * B.tile_start[q].size());
      c_gs[i] = put(B.group_start[q].data(), sizeof(int64_t) * B.group_start[q].size());
    }
  }
  const int ncounters = nb * gsc::kEnsBatchCounters + 1;
  const size_t c_cnt = put(nullptr, sizeof(int) * (size_t)ncounters);
  const size_t ctrl = hctrl.size();
  plan->mem_bytes = ctrl + ws;
  if (hipMalloc(&plan->mem, plan->mem_bytes) != hipSuccess) {
    plan->mem = nullptr;
    return fail(GSC_ERR_ALLOC, "hipMalloc of %zu bytes for the ensemble failed",
                plan->mem_bytes);
  }
  char* base = static_cast<char*>(plan->mem);
  char* work = base + ctrl;
  for (int r = 0; r < count; ++r) {
    gsc::EnsReal& e = rs[(size_t)r];
    const size_t rb = sizeof(gsc::ModeRec) * (size_t)(e.M > 0 ? e.M : 1);
    e.table = reinterpret_cast<gsc::ModeRec*>(work + tab_off[(size_t)r]);
    e.a_tab = reinterpret_cast<double*>(work + tab_off[(size_t)r] + align256(rb));
    e.amp_tab = reinterpret_cast<float*>(work + tab_off[(size_t)r] + align256(rb) +
                                         align256(sizeof(double) * 64 * (size_t)e.ntiles));
    if (e.path == GSC_PATH_FACTORED && e.npts > 0) {
      const gsc::FactoredPlan& fp = plans[(size_t)r];
      char* f = work + f_off[(size_t)r];
      e.S = reinterpret_cast<float2*>(f + fp.off_s);
      e.A = reinterpret_cast<uint32_t*>(f + fp.off_a);
      e.C = reinterpret_cast<float2*>(f + fp.off_c);
    }
  }
  std::memcpy(hctrl.data() + c_desc, rs.data(), sizeof(gsc::EnsReal) * rs.size());
  GSC_HIP(hipMemcpy(base, hctrl.data(), ctrl, hipMemcpyHostToDevice));

  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, plan->device);
  std::vector<gsc::EnsLaunch> els((size_t)nb);
  for (int bi = 0; bi < nb; ++bi) {
    const Batch& B = bt[(size_t)bi];
    gsc::EnsLaunch& el = els[(size_t)bi];
    el = gsc::EnsLaunch{};
    el.d = reinterpret_cast<const gsc::EnsReal*>(base + c_desc);
    el.count = count;
    el.fact = reinterpret_cast<const int*>(base + c_fact[(size_t)bi]);
    el.nfact = (int)B.all_fact.size();
    el.task_start = reinterpret_cast<const int64_t*>(base + c_tk[(size_t)bi]);
    el.n_table_tasks = B.task_start.back();
    el.dir = reinterpret_cast<const int*>(base + c_dir);
    el.ndir = (int)dir.size();
    el.blk_start = reinterpret_cast<const int64_t*>(base + c_bs);
    el.counters = reinterpret_cast<int*>(base + c_cnt);
    el.ncounters = ncounters;
    el.dir_counter = el.counters + nb * gsc::kEnsBatchCounters;
    el.dim = dim;
    el.max_tiles = max_tiles;
    el.ngemm = B.ngemm;
    for (int q = 0; q < NG; ++q) {
      const size_t i = (size_t)bi * NG + (size_t)q;
      gsc::EnsGemmGroup& g = el.gemm[q];
      g.fact = reinterpret_cast<const int*>(base + c_go[i]);
      g.nfact = (int)B.gemm_order[q].size();
      g.tile_start = reinterpret_cast<const int64_t*>(base + c_ts[i]);
      g.group_start = reinterpret_cast<const int64_t*>(base + c_gs[i]);
      g.counters = el.counters + (bi * NG + q) * gsc::kXcdGroups;
      g.brows = gsc::tile_block_rows(tile_ids[q / 2]);
      g.kr = dim == 3 || q % 2 == 1;
      // blocks per CU the GEMM is built for: 256-row tiles one (two with the
      // FP64 flush, which fits 256 registers), 128-row tiles two (four)
      const int64_t bpc_auto = (o->accum == GSC_ACCUM_F64 ? 2 : 1) * (q / 2 == 1 ? 2 : 1);
      const int64_t bpc = o->blocks_per_cu > 0 ? o->blocks_per_cu : bpc_auto;
      g.gemm_blocks =
          (int)std::max<int64_t>(1, std::min<int64_t>(B.tile_start[q].back(), bpc * cus));
    }
    el.direct_blocks =
        (int)std::max<int64_t>(1, std::min<int64_t>(blk_start.back(), 8 * (int64_t)cus));
  }

  GSC_HIP(hipStreamCreateWithFlags(&plan->cap, hipStreamNonBlocking));
  GSC_HIP(hipStreamBeginCapture(plan->cap, hipStreamCaptureModeThreadLocal));
  hipError_t le = sample ? gsc::launch_sample_modes(*sample, plan->cap) : hipSuccess;
  if (le == hipSuccess)
    le = gsc::launch_ensemble(els.data(), nb, o->precision, o->out_dtype, o->accum, plan->cap);
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(plan->cap, &g);
  plan->graph = g;
  if (le != hipSuccess) return fail(GSC_ERR_HIP, "ensemble launch capture: %s", hipGetErrorString(le));
  if (ce != hipSuccess) return fail(GSC_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ce));
  GSC_HIP(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
  return GSC_OK;
}

struct SeededModes {
  int model;
  const double* len_scales;  // host, count
  const uint64_t* seeds;     // host, count
};

int validate_sample(int model, int dim, int64_t M, double len_scale) {
  if (model != GSC_MODEL_GAUSSIAN && model != GSC_MODEL_EXPONENTIAL)
    return fail(GSC_ERR_INVALID, "unknown spectral model %d", model);
  if (dim < 1 || dim > 3) return fail(GSC_ERR_DIM, "mode sampling supports dim 1..3 (got %d)", dim);
  if (M < 0 || M > ((int64_t)1 << 40)) return fail(GSC_ERR_INVALID, "bad mode count");
  if (!(len_scale > 0.0) || !std::isfinite(len_scale))
    return fail(GSC_ERR_INVALID, "len_scale must be positive and finite");
  return GSC_OK;
}

// Shared body of the two creators; seeded: null (caller's mode arrays) or the
// sampler description (the plan owns the mode arrays and draws them in its
// graph).
int ensemble_create(gsc_ensemble** plan, int count, int dim, const int64_t* shapes,
                    const int64_t* m_counts, const int32_t* paths, const int32_t* tiles,
                    const double* scales,
                    const double* cov, const double* z1, const double* z2, const double* axes,
                    void* out, const gsc_opts* opts, const SeededModes* seeded) {
  if (!plan) return fail(GSC_ERR_INVALID, "null plan pointer");
  *plan = nullptr;
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  int rc;
  if ((rc = validate_opts(o)) || (rc = validate_precision(o))) return rc;
  if (dim != 2 && dim != 3) return fail(GSC_ERR_DIM, "ensembles support dim 2 or 3 (got %d)", dim);
  if (count < 1 || count > 20000)
    return fail(GSC_ERR_INVALID, "count must be in 1..20000 (got %d)", count);
  if (!shapes || !m_counts) return fail(GSC_ERR_INVALID, "null shapes or m_counts");
  int64_t m_total = 0, pts = 0;
  for (int r = 0; r < count; ++r) {
    if (m_counts[r] < 0 || m_counts[r] > (int64_t)0x7fffffff)
      return fail(GSC_ERR_INVALID, "realisation %d: bad mode count", r);
    for (int i = 0; i < dim; ++i)
      if (shapes[(int64_t)r * dim + i] < 0)
        return fail(GSC_ERR_INVALID, "realisation %d: negative axis length", r);
    if (paths && (paths[r] < GSC_PATH_AUTO || paths[r] > GSC_PATH_FACTORED))
      return fail(GSC_ERR_INVALID, "realisation %d: unknown path %d", r, paths[r]);
    if (tiles && (tiles[r] < GSC_TILE_AUTO || tiles[r] > GSC_TILE_128))
      return fail(GSC_ERR_INVALID, "realisation %d: unknown tile shape %d", r, tiles[r]);
    m_total += m_counts[r];
    pts += grid_points(dim, shapes + (int64_t)r * dim);
  }
  if (!seeded && m_total > 0 && (!cov || !z1 || !z2))
    return fail(GSC_ERR_INVALID, "null mode arrays");
  if (pts > 0 && (!axes || !out)) return fail(GSC_ERR_INVALID, "null axes or output");
  gsc_ensemble* p = nullptr;
  try {
    p = new gsc_ensemble();
    p->count = count;
    if ((rc = resolve_device(o, &p->device))) {
      delete p;
      return rc;
    }
    if (hipSetDevice(p->device) != hipSuccess) {
      delete p;
      return fail(GSC_ERR_HIP, "hipSetDevice failed");
    }
    if ((rc = check_arch(p->device))) {
      delete p;
      return rc;
    }
    gsc::SampleArgs sa{};
    if (seeded) {
      // plan-owned modes (dim + 2 doubles per mode), then m_off, len_scale, seeds
      const size_t b_modes = align256(sizeof(double) * (size_t)(dim + 2) * (size_t)(m_total + 1));
      const size_t b_off = align256(sizeof(int64_t) * (size_t)(count + 1));
      const size_t b_len = align256(sizeof(double) * (size_t)count);
      const size_t b_seed = align256(sizeof(uint64_t) * (size_t)count);
      if (hipMalloc(&p->mode_mem, b_modes + b_off + b_len + b_seed) != hipSuccess) {
        p->mode_mem = nullptr;
        rc = fail(GSC_ERR_ALLOC, "hipMalloc for the ensemble modes failed");
      } else {
        char* mb = static_cast<char*>(p->mode_mem);
        double* d_cov = reinterpret_cast<double*>(mb);
        int64_t* d_off = reinterpret_cast<int64_t*>(mb + b_modes);
        double* d_len = reinterpret_cast<double*>(mb + b_modes + b_off);
        p->d_seeds = reinterpret_cast<uint64_t*>(mb + b_modes + b_off + b_len);
        std::vector<int64_t> off((size_t)count + 1, 0);
        for (int r = 0; r < count; ++r) off[(size_t)r + 1] = off[(size_t)r] + m_counts[r];
        hipError_t e = hipMemcpy(d_off, off.data(), sizeof(int64_t) * off.size(),
                                 hipMemcpyHostToDevice);
        if (e == hipSuccess)
          e = hipMemcpy(d_len, seeded->len_scales, sizeof(double) * (size_t)count,
                        hipMemcpyHostToDevice);
        if (e == hipSuccess)
          e = hipMemcpy(p->d_seeds, seeded->seeds, sizeof(uint64_t) * (size_t)count,
                        hipMemcpyHostToDevice);
        if (e != hipSuccess) {
          rc = fail(GSC_ERR_HIP, "ensemble mode upload: %s", hipGetErrorString(e));
        } else {
          sa.model = seeded->model;
          sa.dim = dim;
          sa.count = count;
          sa.m_off = d_off;
          sa.len_scale = d_len;
          sa.seeds = p->d_seeds;
          sa.m_total = m_total;
          sa.stream_id = 0;
          sa.cov = d_cov;
          sa.z1 = d_cov + dim * m_total;
          sa.z2 = d_cov + (dim + 1) * m_total;
          cov = sa.cov;
          z1 = sa.z1;
          z2 = sa.z2;
        }
      }
    }
    if (rc == GSC_OK)
      rc = ensemble_build(p, count, dim, shapes, m_counts, paths, tiles, scales, cov, z1, z2,
                          axes, out,
                          o, seeded ? &sa : nullptr);
  } catch (const std::bad_alloc&) {
    rc = fail(GSC_ERR_ALLOC, "host allocation failed");
  }
  if (rc != GSC_OK) {
    ensemble_release(p);
    return rc;
  }
  *plan = p;
  return GSC_OK;
}

}  // namespace

extern "C" {

int gsc_ensemble_create(gsc_ensemble** plan, int count, int dim, const int64_t* shapes,
                        const int64_t* m_counts, const int32_t* paths, const double* scales,
                        const double* cov, const double* z1, const double* z2,
                        const double* axes, void* out, const gsc_opts* opts) {
  return ensemble_create(plan, count, dim, shapes, m_counts, paths, nullptr, scales, cov, z1, z2,
                         axes, out, opts, nullptr);
}

int gsc_ensemble_create_tuned(gsc_ensemble** plan, int count, int dim, const int64_t* shapes,
                              const int64_t* m_counts, const int32_t* paths, const int32_t* tiles,
                              const double* scales, const double* cov, const double* z1,
                              const double* z2, int model, const double* len_scales,
                              const uint64_t* seeds, const double* axes, void* out,
                              const gsc_opts* opts) {
  if (!seeds)
    return ensemble_create(plan, count, dim, shapes, m_counts, paths, tiles, scales, cov, z1, z2,
                           axes, out, opts, nullptr);
  if (model != GSC_MODEL_GAUSSIAN && model != GSC_MODEL_EXPONENTIAL)
    return fail(GSC_ERR_INVALID, "unknown spectral model %d", model);
  if (!len_scales) return fail(GSC_ERR_INVALID, "null len_scales");
  if (count < 1 || count > 20000 || !m_counts)
    return fail(GSC_ERR_INVALID, "count must be in 1..20000 with m_counts given");
  for (int r = 0; r < count; ++r)
    if (!(len_scales[r] > 0.0) || !std::isfinite(len_scales[r]))
      return fail(GSC_ERR_INVALID, "realisation %d: len_scale must be positive and finite", r);
  SeededModes sm{model, len_scales, seeds};
  return ensemble_create(plan, count, dim, shapes, m_counts, paths, tiles, scales, nullptr,
                         nullptr, nullptr, axes, out, opts, &sm);
}

int gsc_ensemble_create_seeded(gsc_ensemble** plan, int count, int dim, const int64_t* shapes,
                               const int64_t* m_counts, const int32_t* paths,
                               const double* scales, int model, const double* len_scales,
                               const uint64_t* seeds, const double* axes, void* out,
                               const gsc_opts* opts) {
  if (model != GSC_MODEL_GAUSSIAN && model != GSC_MODEL_EXPONENTIAL)
    return fail(GSC_ERR_INVALID, "unknown spectral model %d", model);
  if (!len_scales || !seeds) return fail(GSC_ERR_INVALID, "null len_scales or seeds");
  if (count < 1 || count > 20000 || !m_counts)
    return fail(GSC_ERR_INVALID, "count must be in 1..20000 with m_counts given");
  for (int r = 0; r < count; ++r)
    if (!(len_scales[r] > 0.0) || !std::isfinite(len_scales[r]))
      return fail(GSC_ERR_INVALID, "realisation %d: len_scale must be positive and finite", r);
  SeededModes sm{model, len_scales, seeds};
  return ensemble_create(plan, count, dim, shapes, m_counts, paths, nullptr, scales, nullptr,
                         nullptr, nullptr, axes, out, opts, &sm);
}

int gsc_ensemble_reseed(gsc_ensemble* plan, const uint64_t* seeds, void* stream) {
  if (!plan || !plan->exec) return fail(GSC_ERR_INVALID, "invalid ensemble plan");
  if (!plan->d_seeds) return fail(GSC_ERR_INVALID, "the plan was not created from seeds");
  if (!seeds) return fail(GSC_ERR_INVALID, "null seeds");
  GSC_HIP(hipMemcpyAsync(plan->d_seeds, seeds, sizeof(uint64_t) * (size_t)plan->count,
                         hipMemcpyHostToDevice, static_cast<hipStream_t>(stream)));
  return GSC_OK;
}

int gsc_sample_modes_dev(int model, int dim, int64_t M, double len_scale, uint64_t seed,
                         uint64_t stream_id, double* cov, double* z1, double* z2,
                         const gsc_opts* opts, void* stream) {
  int rc = validate_sample(model, dim, M, len_scale);
  if (rc) return rc;
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  if ((rc = validate_opts(o))) return rc;
  if (M == 0) return GSC_OK;
  if (!cov || !z1 || !z2) return fail(GSC_ERR_INVALID, "null output arrays");
  gsc::SampleArgs a{};
  a.model = model;
  a.dim = dim;
  a.count = 1;
  a.m_total = M;
  a.len_scale1 = len_scale;
  a.seed1 = seed;
  a.stream_id = stream_id;
  a.cov = cov;
  a.z1 = z1;
  a.z2 = z2;
  GSC_HIP(gsc::launch_sample_modes(a, static_cast<hipStream_t>(stream)));
  return GSC_OK;
}

int gsc_sample_modes(int model, int dim, int64_t M, double len_scale, uint64_t seed,
                     uint64_t stream_id, double* cov, double* z1, double* z2,
                     const gsc_opts* opts) {
  int rc = validate_sample(model, dim, M, len_scale);
  if (rc) return rc;
  gsc_opts tmp;
  const gsc_opts* o = opts_or_default(opts, &tmp);
  if ((rc = validate_opts(o))) return rc;
  if (M == 0) return GSC_OK;
  if (!cov || !z1 || !z2) return fail(GSC_ERR_INVALID, "null output arrays");
  int dev = 0;
  if ((rc = resolve_device(o, &dev))) return rc;
  GSC_HIP(hipSetDevice(dev));
  if ((rc = check_arch(dev))) return rc;
  const size_t n = static_cast<size_t>(dim + 2) * static_cast<size_t>(M);
  double* d = nullptr;
  if (hipMalloc(&d, sizeof(double) * n) != hipSuccess)
    return fail(GSC_ERR_ALLOC, "hipMalloc of %zu bytes failed", sizeof(double) * n);
  gsc::SampleArgs a{};
  a.model = model;
  a.dim = dim;
  a.count = 1;
  a.m_total = M;
  a.len_scale1 = len_scale;
  a.seed1 = seed;
  a.stream_id = stream_id;
  a.cov = d;
  a.z1 = d + dim * M;
  a.z2 = d + (dim + 1) * M;
  hipError_t e = gsc::launch_sample_modes(a, nullptr);
  if (e == hipSuccess) e = hipMemcpy(cov, a.cov, sizeof(double) * dim * M, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(z1, a.z1, sizeof(double) * M, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(z2, a.z2, sizeof(double) * M, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(GSC_ERR_HIP, "mode sampling failed: %s", hipGetErrorString(e));
  return GSC_OK;
}

int gsc_ensemble_run(gsc_ensemble* plan, void* stream) {
  if (!plan || !plan->exec) return fail(GSC_ERR_INVALID, "invalid ensemble plan");
  GSC_HIP(hipGraphLaunch(plan->exec, static_cast<hipStream_t>(stream)));
  return GSC_OK;
}

int gsc_ensemble_time(gsc_ensemble* plan, int repeats, double* ms) {
  if (!plan || !plan->exec) return fail(GSC_ERR_INVALID, "invalid ensemble plan");
  if (repeats < 1 || repeats > 10000 || !ms)
    return fail(GSC_ERR_INVALID, "repeats must be in 1..10000 with a result pointer");
  hipStream_t st = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipError_t e = hipSetDevice(plan->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) e = hipEventCreate(&ev[i]);
  if (e == hipSuccess) e = hipGraphLaunch(plan->exec, st);  // warm-up
  if (e == hipSuccess) e = hipEventRecord(ev[0], st);
  for (int r = 0; r < repeats && e == hipSuccess; ++r) e = hipGraphLaunch(plan->exec, st);
  if (e == hipSuccess) e = hipEventRecord(ev[1], st);
  if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
  float t = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, ev[0], ev[1]);
  if (st) (void)hipStreamSynchronize(st);
  for (int i = 0; i < 2; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) return fail(GSC_ERR_HIP, "ensemble timing failed: %s", hipGetErrorString(e));
  *ms = (double)t / repeats;
  return GSC_OK;
}

int gsc_ensemble_info(const gsc_ensemble* plan, int64_t* n_factored, int64_t* n_direct,
                      size_t* workspace_bytes, double* point_modes) {
  if (!plan) return fail(GSC_ERR_INVALID, "invalid ensemble plan");
  if (n_factored) *n_factored = plan->n_fact;
  if (n_direct) *n_direct = plan->n_dir;
  if (workspace_bytes) *workspace_bytes = plan->mem_bytes;
  if (point_modes) *point_modes = plan->point_modes;
  return GSC_OK;
}

void gsc_ensemble_destroy(gsc_ensemble* plan) { ensemble_release(plan); }

void* gsc_device_alloc(size_t bytes, int device) {
  if (device >= 0 && hipSetDevice(device) != hipSuccess) {
    fail(GSC_ERR_HIP, "hipSetDevice(%d) failed", device);
    return nullptr;
  }
  void* p = nullptr;
  if (hipMalloc(&p, bytes > 0 ? bytes : 1) != hipSuccess) {
    fail(GSC_ERR_ALLOC, "hipMalloc of %zu bytes failed", bytes);
    return nullptr;
  }
  return p;
}

void gsc_device_free(void* ptr) {
  if (ptr) (void)hipFree(ptr);
}

int gsc_memcpy_to_device(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GSC_OK;
  GSC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return GSC_OK;
}

int gsc_memcpy_to_host(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return GSC_OK;
  GSC_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return GSC_OK;
}

int gsc_synchronize(void) {
  GSC_HIP(hipDeviceSynchronize());
  return GSC_OK;
}

}  // extern "C"
