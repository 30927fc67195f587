// Generic stage-3 matvec: the reference's contraction order (tensordot A<->v, then
// B<->T over (alpha, q)) for any P, Q, d, built from the FP64 MFMA GEMM and permute
// kernels.  Used for every shape the fused kernels are not instantiated for.
#include "cc_common.h"

extern "C" int cc_stage3_matvec_generic_work_size(int64_t n_alpha, int64_t P, int64_t Q, int64_t d, int64_t* out) {
  *out = 2 * n_alpha * P * Q * d + n_alpha * Q * Q + P * Q * d;
  return CC_OK;
}

extern "C" int cc_stage3_matvec_generic_f64(void* stream, int n_terms, const double* const* A_slabs,
                                            const double* const* B_slabs, const double* const* O_mats,
                                            int64_t n_alpha, int64_t P, int64_t Q, int64_t d,
                                            const double* v, double* y, double* work) {
  if (n_terms < 0 || n_alpha < 0 || P <= 0 || Q <= 0 || d <= 0) { cc_set_last_error("stage3: bad sizes"); return CC_ERR_ARG; }
  const int64_t n = P * Q * d;
  if (n_terms == 0 || n_alpha == 0) return cc_fill_f64(stream, n, 0.0, y);
  double* T = work;                       // [alpha][p'][q][s]
  double* Tp = T + n_alpha * P * Q * d;   // [p'][alpha][q][s]
  double* Bt = Tp + n_alpha * P * Q * d;  // [alpha][q][q']
  double* Vo = Bt + n_alpha * Q * Q;      // [p][q][s'] = sum_s O[s'][s] v[p][q][s]
  int rc;
  for (int t = 0; t < n_terms; ++t) {
    const double* V = v;
    if (O_mats[t] != nullptr) {
      // Vo[m][s'] = sum_s v[m][s] O[s'][s]
      rc = cc_gemm_f64(stream, 1, P * Q, d, d, 1.0, v, 0, d, 1, O_mats[t], 0, 1, d, 0.0, Vo, 0, d, 1);
      if (rc) return rc;
      V = Vo;
    }
    // T[(alpha,p')][(q,s)] = A[(alpha,p')][p] V[p][(q,s)]
    rc = cc_gemm_f64(stream, 1, n_alpha * P, Q * d, P, 1.0, A_slabs[t], 0, P, 1, V, 0, Q * d, 1, 0.0, T, 0, Q * d, 1);
    if (rc) return rc;
    // Tp[p'][alpha][q][s] = T[alpha][p'][q][s]
    {
      int64_t shape[3] = {P, n_alpha, Q * d};
      int64_t strides[3] = {Q * d, P * Q * d, 1};
      rc = cc_permute(stream, CC_F64, T, Tp, 3, shape, strides);
      if (rc) return rc;
    }
    // Bt[alpha][q][q'] = B[alpha][q'][q]
    {
      int64_t shape[3] = {n_alpha, Q, Q};
      int64_t strides[3] = {Q * Q, 1, Q};
      rc = cc_permute(stream, CC_F64, B_slabs[t], Bt, 3, shape, strides);
      if (rc) return rc;
    }
    // Y[p'][q'][s] (+)= sum_{(alpha,q)} Tp[p'][(alpha,q)][s] Bt[(alpha,q)][q'], batched over s
    rc = cc_gemm_f64(stream, d, P, Q, n_alpha * Q, 1.0, Tp, 1, n_alpha * Q * d, d, Bt, 0, Q, 1,
                     t == 0 ? 0.0 : 1.0, y, 1, Q * d, d);
    if (rc) return rc;
  }
  return CC_OK;
}
