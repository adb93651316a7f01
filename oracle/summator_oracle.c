/*
 * CPU oracle for the randomization-method summators -- TEST INFRASTRUCTURE ONLY.
 *
 * This file restates, in plain C with IEEE-754 binary64 arithmetic, what the
 * reference Rust crate computes in src/field.rs.  It exists to check the HIP
 * product path (tests/, the smoke entry and the CPU-baseline leg of bench.py);
 * nothing in the shipped library links or calls it.
 *
 * Build with -ffp-contract=off and without -ffast-math: Rust never contracts
 * a*b+c into an FMA, and the reference tests compare bit for bit
 * (field.rs:363 uses assert_eq!), so the association written here is the
 * association the reference evaluates:
 *
 *   summator        field.rs:53-62   phase = ((0 + k0*x0) + k1*x1) + k2*x2
 *                                    (ndarray 1-D dot over strided columns is a
 *                                    sequential fold from 0.0), then
 *                                    z12 = z1*cos(phase) + z2*sin(phase) and a
 *                                    mode-ordered fold sum = sum + z12 from 0.0.
 *   summator_incompr field.rs:130-174 per mode k_2 = k[0] / dot(k,k)
 *                                    (ShortVec::dot, short_vec.rs:31-33, is the
 *                                    same sequential sum), then for every point
 *                                    sum[0] += (1 - k0*k_2)*z12 and
 *                                    sum[i] -= (k_i*k_2)*z12  (field.rs:148-151).
 *   summator_fourier field.rs:236-246 z12 = sf*(z1*cos + z2*sin).
 *
 * cos/sin are the platform libm (glibc), as in the reference (Rust's f64::cos
 * and f64::sin lower to the system libm).
 *
 * Parallel decomposition is OpenMP over points (the reference's rayon
 * par_map_collect over pos.columns(), field.rs:53).  For summator and
 * summator_fourier every point's value is independent of the split, so the
 * result is deterministic and equal to the reference's.  For
 * summator_incompr the reference folds modes in chunks of >= 100
 * (field.rs:134) and reduces chunk accumulators (field.rs:158-163); the
 * point-parallel form below sums modes strictly in index order, which is what
 * the reference does whenever M <= 100 (one chunk), and differs from it only
 * by reassociation otherwise.  gsc_oracle_summate_incompr_modechunk keeps the
 * reference's mode-outer loop order for CPU timing.
 *
 * All arrays are addressed through signed element strides so that the
 * oracle can read numpy views of any layout.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define AT2(p, sd, sx, i, j) ((p)[(int64_t)(i) * (sd) + (int64_t)(j) * (sx)])
#define AT1(p, s, i) ((p)[(int64_t)(i) * (s)])

static void set_threads(int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
}

/* field.rs:57 -- ndarray dot of two strided 1-D views: sequential fold from 0. */
static inline double phase_of(int dim, const double* k, int64_t ksd, int64_t km,
                              const double* x, int64_t xsd, int64_t xj) {
  double s = 0.0;
  for (int i = 0; i < dim; ++i) s = s + AT2(k, ksd, 1, i, km) * AT2(x, xsd, 1, i, xj);
  return s;
}

/*
 * summator (field.rs:37-65).  cov: (dim, M) with strides (cov_sd, cov_sm);
 * pos: (dim, N) with strides (pos_sd, pos_sp); out: N with stride out_s.
 */
void gsc_oracle_summate(int dim, int64_t M, int64_t N,
                        const double* cov, int64_t cov_sd, int64_t cov_sm,
                        const double* z1, int64_t s1, const double* z2, int64_t s2,
                        const double* pos, int64_t pos_sd, int64_t pos_sp,
                        double* out, int64_t out_s, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j) {
    double sum = 0.0;
    for (int64_t m = 0; m < M; ++m) {
      double phase = 0.0;
      for (int i = 0; i < dim; ++i)
        phase = phase + AT2(cov, cov_sd, cov_sm, i, m) * AT2(pos, pos_sd, pos_sp, i, j);
      double z12 = AT1(z1, s1, m) * cos(phase) + AT1(z2, s2, m) * sin(phase);
      sum = sum + z12;
    }
    AT1(out, out_s, j) = sum;
  }
}

/*
 * summator_fourier (field.rs:219-249): same fold with a per-mode weight
 * spectrum_factor applied as sf * (z1*cos + z2*sin) (field.rs:243).
 */
void gsc_oracle_summate_fourier(int dim, int64_t M, int64_t N,
                                const double* sf, int64_t ssf,
                                const double* modes, int64_t md_sd, int64_t md_sm,
                                const double* z1, int64_t s1, const double* z2, int64_t s2,
                                const double* pos, int64_t pos_sd, int64_t pos_sp,
                                double* out, int64_t out_s, int nthreads) {
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j) {
    double sum = 0.0;
    for (int64_t m = 0; m < M; ++m) {
      double phase = 0.0;
      for (int i = 0; i < dim; ++i)
        phase = phase + AT2(modes, md_sd, md_sm, i, m) * AT2(pos, pos_sd, pos_sp, i, j);
      double z12 = AT1(sf, ssf, m) * (AT1(z1, s1, m) * cos(phase) + AT1(z2, s2, m) * sin(phase));
      sum = sum + z12;
    }
    AT1(out, out_s, j) = sum;
  }
}

/* Per-mode projector factor k_2 = k0 / dot(k, k) (field.rs:138). */
static inline double proj_k2(int dim, const double* cov, int64_t sd, int64_t sm, int64_t m) {
  double kk = 0.0;
  for (int i = 0; i < dim; ++i) kk = kk + AT2(cov, sd, sm, i, m) * AT2(cov, sd, sm, i, m);
  return AT2(cov, sd, sm, 0, m) / kk;
}

/*
 * summator_incompr (field.rs:97-182), point-parallel, modes in index order.
 * variant 0: the reference's association (field.rs:138-151):
 *            acc0 += (1 - k0*k_2)*z12 ; acc_i -= (k_i*k_2)*z12.
 * variant 1: projector association p_i = delta_i0 - (k_i*k0)/|k|^2,
 *            acc_i += p_i*z12 (matches the golden values at field.rs:396-427
 *            bit for bit; kept so that tests can pin both).
 * out: (dim, N) with strides (out_sd, out_sp).  Returns -1 if dim not in {2,3}
 * (the reference panics at field.rs:180).
 */
int gsc_oracle_summate_incompr(int dim, int64_t M, int64_t N,
                               const double* cov, int64_t cov_sd, int64_t cov_sm,
                               const double* z1, int64_t s1, const double* z2, int64_t s2,
                               const double* pos, int64_t pos_sd, int64_t pos_sp,
                               double* out, int64_t out_sd, int64_t out_sp,
                               int variant, int nthreads) {
  if (dim != 2 && dim != 3) return -1;
  set_threads(nthreads);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j) {
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t m = 0; m < M; ++m) {
      double phase = 0.0;
      for (int i = 0; i < dim; ++i)
        phase = phase + AT2(cov, cov_sd, cov_sm, i, m) * AT2(pos, pos_sd, pos_sp, i, j);
      double z12 = AT1(z1, s1, m) * cos(phase) + AT1(z2, s2, m) * sin(phase);
      if (variant == 0) {
        double k_2 = proj_k2(dim, cov, cov_sd, cov_sm, m);
        acc[0] += (1.0 - AT2(cov, cov_sd, cov_sm, 0, m) * k_2) * z12;
        for (int i = 1; i < dim; ++i) acc[i] -= AT2(cov, cov_sd, cov_sm, i, m) * k_2 * z12;
      } else {
        double kk = 0.0;
        for (int i = 0; i < dim; ++i) kk = kk + AT2(cov, cov_sd, cov_sm, i, m) * AT2(cov, cov_sd, cov_sm, i, m);
        double k0 = AT2(cov, cov_sd, cov_sm, 0, m);
        for (int i = 0; i < dim; ++i) {
          double p = (i == 0 ? 1.0 : 0.0) - (AT2(cov, cov_sd, cov_sm, i, m) * k0) / kk;
          acc[i] += p * z12;
        }
      }
    }
    for (int i = 0; i < dim; ++i) AT2(out, out_sd, out_sp, i, j) = acc[i];
  }
  return 0;
}

/*
 * summator_incompr with the reference's loop order, for CPU timing only
 * (field.rs:130-163): modes are split into contiguous chunks of at least
 * min_len (the reference uses with_min_len(100)); each chunk owns a full
 * N x dim accumulator and streams it once per mode; chunk accumulators are
 * then summed element-wise.  out: (dim, N) with strides (out_sd, out_sp).
 * Returns -1 for dim not in {2,3}, -2 if the accumulators cannot be
 * allocated.
 */
int gsc_oracle_summate_incompr_modechunk(int dim, int64_t M, int64_t N,
                                         const double* cov, int64_t cov_sd, int64_t cov_sm,
                                         const double* z1, int64_t s1, const double* z2, int64_t s2,
                                         const double* pos, int64_t pos_sd, int64_t pos_sp,
                                         double* out, int64_t out_sd, int64_t out_sp,
                                         int64_t min_len, int nthreads) {
  if (dim != 2 && dim != 3) return -1;
  set_threads(nthreads);
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  if (min_len < 1) min_len = 1;
  int64_t nchunks = M / min_len;
  if (nchunks > nt) nchunks = nt;
  if (nchunks < 1) nchunks = 1;
  /* point copies in the reference's ShortVec layout (field.rs:120-123) */
  double* p = (double*)malloc(sizeof(double) * (size_t)(N * dim + 1));
  double* acc = (double*)calloc((size_t)(nchunks * N * dim + 1), sizeof(double));
  if (!p || !acc) { free(p); free(acc); return -2; }
  for (int64_t j = 0; j < N; ++j)
    for (int i = 0; i < dim; ++i) p[j * dim + i] = AT2(pos, pos_sd, pos_sp, i, j);
#pragma omp parallel for schedule(static, 1)
  for (int64_t c = 0; c < nchunks; ++c) {
    int64_t m0 = M * c / nchunks, m1 = M * (c + 1) / nchunks;
    double* a = acc + c * N * dim;
    for (int64_t m = m0; m < m1; ++m) {
      double k[3] = {0, 0, 0};
      for (int i = 0; i < dim; ++i) k[i] = AT2(cov, cov_sd, cov_sm, i, m);
      double kk = 0.0;
      for (int i = 0; i < dim; ++i) kk = kk + k[i] * k[i];
      double k_2 = k[0] / kk;
      double w1 = AT1(z1, s1, m), w2 = AT1(z2, s2, m);
      for (int64_t j = 0; j < N; ++j) {
        const double* x = p + j * dim;
        double phase = 0.0;
        for (int i = 0; i < dim; ++i) phase = phase + k[i] * x[i];
        double z12 = w1 * cos(phase) + w2 * sin(phase);
        a[j * dim] += (1.0 - k[0] * k_2) * z12;
        for (int i = 1; i < dim; ++i) a[j * dim + i] -= k[i] * k_2 * z12;
      }
    }
  }
  for (int64_t c = 1; c < nchunks; ++c)
    for (int64_t t = 0; t < N * dim; ++t) acc[t] += acc[c * N * dim + t];
  for (int64_t j = 0; j < N; ++j)
    for (int i = 0; i < dim; ++i) AT2(out, out_sd, out_sp, i, j) = acc[j * dim + i];
  free(p);
  free(acc);
  return 0;
}

/* Number of threads OpenMP would use (reported as `cores` by the baseline). */
int gsc_oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
