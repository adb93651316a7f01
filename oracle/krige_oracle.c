/*
 * CPU oracle of the kriging calculators -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 restatement of src/krige.rs:
 *   calculator_field_krige_and_variance (krige.rs:24-70): for every target j
 *   (column v_col of krig_vecs) and every condition i (in index order):
 *     krig_fac = krig_mat[:, i] . v_col          (ndarray dot, k in order)
 *     field += cond[i] * krig_fac;  error += v_col[i] * krig_fac
 *   calculator_field_krige (krige.rs:88-118): the field part alone.
 * The reference folds over i with rayon (fold + reduce, krige.rs:50-66), so
 * its order depends on the split; the sequential order used here reproduces
 * the golden values of krige.rs:188-245 exactly.  -ffp-contract=off keeps
 * multiply and add unfused, as Rust does.  OpenMP over targets.
 */
#include <stdint.h>

void gsc_oracle_krige(int64_t n, int64_t N, const double* mat, int64_t m0, int64_t m1,
                      const double* vecs, int64_t v0, int64_t v1, const double* cond,
                      int64_t cs, double* field, int64_t fs, double* error, int64_t es) {
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < N; ++j) {
    double acc_f = 0.0, acc_e = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      double krig_fac = 0.0;
      for (int64_t k = 0; k < n; ++k) krig_fac += mat[k * m0 + i * m1] * vecs[k * v0 + j * v1];
      acc_f += cond[i * cs] * krig_fac;
      acc_e += vecs[i * v0 + j * v1] * krig_fac;
    }
    field[j * fs] = acc_f;
    if (error) error[j * es] = acc_e;
  }
}
