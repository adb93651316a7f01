// Internal declarations shared by the HIP kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace gsc {

// One spectral mode, prepared once per call in FP64 and read wave-uniformly
// (scalar loads) by the summation kernels.
//   kap[i] = k_i / (2 pi)          phase in revolutions: t = kap . x + th
//   th     = atan2(z1, z2) / (2 pi)
//   amp[c] = per-component amplitude of A sin(2 pi t), A = hypot(z1, z2):
//            scalar summator: amp[0] = A; Fourier: amp[0] = sf * A;
//            incompressible:  amp[c] = (delta_c0 - k_c k_0 / |k|^2) * A.
// z1 cos(phi) + z2 sin(phi) = A sin(phi + atan2(z1, z2)), so one sine per
// point-mode replaces the reference's cos + sin pair (field.rs:58).
struct alignas(16) ModeRec {
  double kap[3];
  double th;
  float amp[4];
};
static_assert(sizeof(ModeRec) == 48, "ModeRec layout");

// For dim <= 3 the table holds the ModeRec records followed by the tiles of
// the matrix-core kernels (see k_prep_small).  For 4 <= dim <= kMaxDim it is
// (dim + 2) doubles per mode: [th, amp, kap_0 .. kap_{dim-1}].
constexpr int kMaxDim = 16;

enum Kind { KIND_SCALAR = 0, KIND_INCOMPR = 1, KIND_FOURIER = 2 };

size_t mode_table_bytes(int dim, int64_t M);

// Builds the mode table from dense device arrays cov (dim, M) row-major,
// z1, z2 (M) and, for KIND_FOURIER, sf (M).
// The mode-table preparation as arguments of a kernel that does it in
// blocks of its own (dim <= 3): ntiles x 16 modes (prep_mode), table the
// mode table of mode_table_bytes(dim, M) bytes.
struct PrepArgs {
  int kind, dim;
  int64_t M;
  int ntiles;
  const double *cov, *z1, *z2, *sf;
  ModeRec* table;
  double* a_tab;
  float* amp_tab;
};
PrepArgs prep_args(int kind, int dim, int64_t M, const double* cov, const double* z1,
                   const double* z2, const double* sf, void* table);

hipError_t launch_prep(int kind, int dim, int64_t M, const double* cov, const double* z1,
                       const double* z2, const double* sf, void* table, hipStream_t stream);

// out[j] (N values of out_dtype) = scale * sum_m amp0_m sin(2 pi (kap_m . x_j + th_m)).
// variant 0: phases on the matrix cores (dim <= 3); 1: phases on the VALU.
hipError_t launch_direct_scalar(int dim, int64_t M, int64_t N, const void* table,
                                const double* pos, int64_t pos_ld, void* out, int out_dtype,
                                double scale, int variant, hipStream_t stream);

// out[c * out_ld + j] = scale * sum_m amp_{c,m} sin(2 pi (kap_m . x_j + th_m)), c < dim.
hipError_t launch_direct_incompr(int dim, int64_t M, int64_t N, const void* table,
                                 const double* pos, int64_t pos_ld, void* out, int64_t out_ld,
                                 int out_dtype, double scale, int variant, hipStream_t stream);

// Local-phase ("binned") direct kernel, dim 2 and 3 (direct_local.hip): a
// spatial counting sort of the points, then phases kap . (x - X_cell) + r0 on
// the matrix cores.  bin_work: binned_work_bytes(dim, M, N) bytes.
// vec = false: out[j] (the scalar and Fourier summators, amp[0] per mode);
// vec = true: out[c * out_ld + j] for c < dim (incompressible, amp[c]).
bool binned_supported(int dim, int64_t N);
size_t binned_work_bytes(int dim, int64_t M, int64_t N);
// split: null, or an event recorded on `stream` between the cell sort and the
// summation kernel (phase timing).  prep: null when `table` is prepared
// already, else the preparation, done by extra blocks of the first sort
// kernel (one launch less).
hipError_t launch_direct_binned(int dim, int64_t M, int64_t N, const void* table,
                                const double* pos, int64_t pos_ld, void* out, int64_t out_ld,
                                bool vec, int out_dtype, double scale, void* bin_work,
                                hipStream_t stream, hipEvent_t split = nullptr,
                                const PrepArgs* prep = nullptr);

// Seeded spectral-mode sampling (sampling.hip): Philox4x64-10 words of key
// (seed, stream_id), Box-Muller normals, k = g / l (model 0, Gaussian) or
// g / (l |w|) (model 1, exponential), z1, z2 ~ N(0, 1).
struct SampleArgs {
  int model;
  int dim;
  int count;               // realisations described by m_off / len_scale / seeds
  const int64_t* m_off;    // device, count + 1 prefix sums of the mode counts; null: one
                           // realisation of m_total modes with len_scale1, seed1
  const double* len_scale; // device, count
  const uint64_t* seeds;   // device, count
  int64_t m_total;
  double len_scale1;
  uint64_t seed1;
  uint64_t stream_id;
  double* cov;             // (dim, M_r) row-major blocks back to back
  double* z1;
  double* z2;
};
hipError_t launch_sample_modes(const SampleArgs& a, hipStream_t stream);

// Kriging calculators (krige.hip).  mat (n, n) and vecs (n, N) row-major
// device arrays (vecs row pitch vld), cond (n); field[j * fs]; error[j * es]
// or null for the field alone.  work: krige_work_bytes(n, N, error != null).
size_t krige_work_bytes(int64_t n, int64_t N, bool with_error);
hipError_t launch_krige(int64_t n, int64_t N, const double* mat, const double* vecs, int64_t vld,
                        const double* cond, double* field, int64_t fs, double* error, int64_t es,
                        void* work, hipStream_t stream);

}  // namespace gsc
