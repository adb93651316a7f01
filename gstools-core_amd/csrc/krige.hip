// Kriging calculators (src/krige.rs:24-118): with n conditions and N targets,
//   krig_fac_i(j) = sum_k mat[k][i] vecs[k][j]        (W = mat^T vecs)
//   field[j]      = sum_i cond[i] krig_fac_i(j)
//   error[j]      = sum_i vecs[i][j] krig_fac_i(j)
// (krige.rs:47-57: `m_row` walks the columns of krig_mat).
//
// field and error (calculator_field_krige_and_variance): W costs 2 n^2 N
// flops and is computed tile by tile on the FP64 matrix cores
// (v_mfma_f64_16x16x4f64); the two contractions with cond and vecs are fused
// into the tile epilogue, so W never reaches memory.  A block owns 128 rows i
// (four waves of 32: two 16-row MFMA tiles each, so every B operand read from
// LDS feeds two matrix instructions) and 128 targets j (eight 16-column MFMA
// tiles per wave); K runs over the n conditions in LDS-staged chunks of 16
// (the accumulators, 128 registers, and the next chunk's prefetch fit two
// waves per SIMD).  Each block writes its partial sums over its rows; a
// second kernel adds the partials of all row blocks in a fixed order
// (results are bitwise reproducible).
//
// field only (calculator_field_krige): the same sum reassociated as
// field[j] = sum_k vecs[k][j] w[k], w = mat cond (n^2 + 2 n N flops, bound
// by reading vecs once).  The reassociation changes rounding only.
#include "gsc_device.h"
#include "gsc_internal.h"

#include <hip/hip_runtime.h>

namespace gsc {
namespace {

typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kRS = 2;        // 16-row MFMA tiles per wave
constexpr int kKI = 64 * kRS; // rows i per block (four waves of 16 kRS)
constexpr int kKK = 16;       // conditions k per LDS chunk
constexpr int kTJ = 8;   // 16-column MFMA tiles per wave: kKJ = 16 kTJ targets per block
constexpr int kKJ = 16 * kTJ;
constexpr int kXcds = 8; // blocks are dispatched to the XCDs round-robin

// Block b -> (row block ib, target block jb): the row blocks of one target
// block run on one XCD, so its columns of vecs stay in that XCD's L2.
__device__ __forceinline__ bool krige_block(int nib, int njb, int& ib, int& jb) {
  const int b = blockIdx.x;
  const int xcd = b % kXcds, s = b / kXcds;
  jb = xcd + kXcds * (s / nib);
  ib = s % nib;
  return jb < njb;
}

// LDS of a block: the staged chunk of mat (kKK x kKI) and vecs (kKK x kKJ),
// reused after the K loop for the epilogue's partial sums (2 x 16 x kKJ).
struct KrigeSmem {
  union {
    struct {
      double m[kKK][kKI + 1];
      double v[kKK][kKJ + 1];
    } st;
    struct {
      double f[16][kKJ + 1];
      double e[16][kKJ + 1];
    } red;
  };
};

// MFMA operand / result layouts (v_mfma_f64_16x16x4f64, verified on gfx950):
//   A: lane l holds A[l % 16][l / 16], B: lane l holds B[l / 16][l % 16],
//   D: lane l, item it holds D[4 it + l / 16][l % 16].
// Here A[i][k] = mat[k][i] (16 rows i of one of the wave's kRS tiles), B[k][j] = vecs[k][j].
// Every block reads all of mat once per kKJ targets, so the target width
// sets mat's share of the traffic (n^2 8 / kKJ bytes per target).
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_krige_tiles(int64_t n, int64_t N, const double* __restrict__ mat,
              const double* __restrict__ vecs, int64_t vld, const double* __restrict__ cond,
              int nib, int njb, double* __restrict__ part_f, double* __restrict__ part_e) {
  __shared__ KrigeSmem sm;
  int ib, jb;
  if (!krige_block(nib, njb, ib, jb)) return;
  const int64_t i0 = static_cast<int64_t>(ib) * kKI, j0 = static_cast<int64_t>(jb) * kKJ;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, g = lane >> 4, jl = lane & 15;
  f64x4 acc[kRS][kTJ];
#pragma unroll
  for (int r = 0; r < kRS; ++r)
#pragma unroll
    for (int t = 0; t < kTJ; ++t) acc[r][t] = f64x4{0.0, 0.0, 0.0, 0.0};
  // this thread's share of a chunk: mat[k0 + tm_k + kStepM u][i0 + tm_i] and
  // vecs[k0 + tv_k + kStepV u][j0 + tv_j] (zero outside), i.e. element
  // e = threadIdx.x + 256 u of the row-major chunk tiles
  constexpr int kPerM = kKK * kKI / 256, kPerV = kKK * kKJ / 256;
  constexpr int kStepM = 256 / kKI, kStepV = 256 / kKJ;
  const int tm_k = threadIdx.x / kKI, tm_i = threadIdx.x % kKI;
  const int tv_k = threadIdx.x / kKJ, tv_j = threadIdx.x % kKJ;
  const bool m_ok = i0 + tm_i < n, v_ok = j0 + tv_j < N;
  const double* mp = mat + (m_ok ? i0 + tm_i : 0);
  const double* vp = vecs + (v_ok ? j0 + tv_j : 0);
  double rm[kPerM], rv[kPerV];
  // interior tiles and whole chunks need no bounds checks: one running row
  // offset per operand and uniform strides between this thread's rows
  const bool interior = i0 + kKI <= n && j0 + kKJ <= N;
  const int64_t m_step = kStepM * n, v_step = kStepV * vld;
  const int64_t m_row0 = tm_k * n, v_row0 = tv_k * vld;
  auto gload = [&](int64_t k0) {
    if (interior && k0 + kKK <= n) {
      const double* pm = mp + m_row0 + k0 * n;
      const double* pv = vp + v_row0 + k0 * vld;
#pragma unroll
      for (int u = 0; u < kPerM; ++u) rm[u] = pm[u * m_step];
#pragma unroll
      for (int u = 0; u < kPerV; ++u) rv[u] = pv[u * v_step];
      return;
    }
#pragma unroll
    for (int u = 0; u < kPerM; ++u) {
      const int64_t k = k0 + tm_k + kStepM * u;
      rm[u] = (m_ok && k < n) ? mp[k * n] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kPerV; ++u) {
      const int64_t k = k0 + tv_k + kStepV * u;
      rv[u] = (v_ok && k < n) ? vp[k * vld] : 0.0;
    }
  };
  gload(0);
  for (int64_t k0 = 0; k0 < n; k0 += kKK) {
#pragma unroll
    for (int u = 0; u < kPerM; ++u) sm.st.m[tm_k + kStepM * u][tm_i] = rm[u];
#pragma unroll
    for (int u = 0; u < kPerV; ++u) sm.st.v[tv_k + kStepV * u][tv_j] = rv[u];
    __syncthreads();
    // the next chunk's loads are in flight during this chunk's MFMAs
    if (k0 + kKK < n) gload(k0 + kKK);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int kk = 0; kk < kKK; kk += 4) {
      double a[kRS];
#pragma unroll
      for (int r = 0; r < kRS; ++r) a[r] = sm.st.m[kk + g][16 * (kRS * w + r) + jl];
#pragma unroll
      for (int t = 0; t < kTJ; ++t) {
        const double b = sm.st.v[kk + g][16 * t + jl];
#pragma unroll
        for (int r = 0; r < kRS; ++r)
          acc[r][t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[r], b, acc[r][t], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // epilogue: acc[r][t][it] = W[i0 + 16 (kRS w + r) + 4 it + g][j0 + 16 t + jl]
#pragma unroll
  for (int t = 0; t < kTJ; ++t) {
    const int64_t j = j0 + 16 * t + jl;
    double pf = 0.0, pe = 0.0;
#pragma unroll
    for (int r = 0; r < kRS; ++r)
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        const int64_t i = i0 + 16 * (kRS * w + r) + 4 * it + g;
        const bool ok = i < n && j < N;
        const double wv = acc[r][t][it];
        pf = __builtin_fma(ok ? cond[i] : 0.0, wv, pf);
        pe = __builtin_fma(ok ? vecs[i * vld + j] : 0.0, wv, pe);
      }
    sm.red.f[4 * w + g][16 * t + jl] = pf;
    sm.red.e[4 * w + g][16 * t + jl] = pe;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kKJ; c += 256) {
    double f = 0.0, e = 0.0;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      f += sm.red.f[r][c];
      e += sm.red.e[r][c];
    }
    const int64_t j = j0 + c;
    if (j < N) {
      part_f[static_cast<int64_t>(ib) * N + j] = f;
      part_e[static_cast<int64_t>(ib) * N + j] = e;
    }
  }
}

__global__ void __launch_bounds__(256)
k_krige_sum(int64_t N, int nib, const double* __restrict__ part_f,
            const double* __restrict__ part_e, double* __restrict__ field, int64_t fs,
            double* __restrict__ error, int64_t es) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (j >= N) return;
  double f = 0.0, e = 0.0;
  for (int ib = 0; ib < nib; ++ib) {
    f += part_f[static_cast<int64_t>(ib) * N + j];
    e += part_e[static_cast<int64_t>(ib) * N + j];
  }
  field[j * fs] = f;
  error[j * es] = e;
}

// w[k] = sum_i mat[k][i] cond[i]: one wave per row k, lanes over i, then a
// fixed-order tree over the lanes.
__global__ void __launch_bounds__(256)
k_krige_w(int64_t n, const double* __restrict__ mat, const double* __restrict__ cond,
          double* __restrict__ w) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (k >= n) return;
  double s = 0.0;
  for (int64_t i = lane; i < n; i += 64) s = __builtin_fma(mat[k * n + i], cond[i], s);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) w[k] = s;
}

// field[j] = sum_k vecs[k][j] w[k]: one thread per target, rows in order;
// kFieldRows rows' loads are issued before their FMAs, so that many loads
// per wave are in flight.
constexpr int kFieldRows = 8;
__global__ void __launch_bounds__(256)
k_krige_field(int64_t n, int64_t N, const double* __restrict__ vecs, int64_t vld,
              const double* __restrict__ w, double* __restrict__ field, int64_t fs) {
  const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (j >= N) return;
  const double* p = vecs + j;
  double f = 0.0;
  int64_t k = 0;
  for (; k + kFieldRows <= n; k += kFieldRows) {
    double v[kFieldRows];
#pragma unroll
    for (int u = 0; u < kFieldRows; ++u) v[u] = p[(k + u) * vld];
#pragma unroll
    for (int u = 0; u < kFieldRows; ++u) f = __builtin_fma(v[u], w[k + u], f);
  }
  for (; k < n; ++k) f = __builtin_fma(p[k * vld], w[k], f);
  field[j * fs] = f;
}

}  // namespace

size_t krige_work_bytes(int64_t n, int64_t N, bool with_error) {
  if (!with_error) return sizeof(double) * static_cast<size_t>(n > 0 ? n : 1);
  const int64_t nib = (n + kKI - 1) / kKI;
  return 2 * sizeof(double) * static_cast<size_t>(nib > 0 ? nib : 1) *
         static_cast<size_t>(N > 0 ? N : 1);
}

hipError_t launch_krige(int64_t n, int64_t N, const double* mat, const double* vecs, int64_t vld,
                        const double* cond, double* field, int64_t fs, double* error, int64_t es,
                        void* work, hipStream_t s) {
  if (N <= 0) return hipSuccess;
  if (!error) {
    double* w = static_cast<double*>(work);
    if (n > 0) k_krige_w<<<static_cast<unsigned>((n + 3) / 4), 256, 0, s>>>(n, mat, cond, w);
    k_krige_field<<<static_cast<unsigned>((N + 255) / 256), 256, 0, s>>>(n, N, vecs, vld, w,
                                                                          field, fs);
    return hipGetLastError();
  }
  const int nib = static_cast<int>((n + kKI - 1) / kKI);
  const int njb = static_cast<int>((N + kKJ - 1) / kKJ);
  double* part_f = static_cast<double*>(work);
  double* part_e = part_f + static_cast<int64_t>(nib) * N;
  if (nib > 0) {
    const int64_t blocks = static_cast<int64_t>(kXcds) * nib * ((njb + kXcds - 1) / kXcds);
    k_krige_tiles<<<static_cast<unsigned>(blocks), 256, 0, s>>>(n, N, mat, vecs, vld, cond, nib,
                                                                njb, part_f, part_e);
  }
  k_krige_sum<<<static_cast<unsigned>((N + 255) / 256), 256, 0, s>>>(N, nib, part_f, part_e,
                                                                      field, fs, error, es);
  return hipGetLastError();
}

}  // namespace gsc
