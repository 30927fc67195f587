/*
 * carcassonne_hip.h -- C ABI of the MI355X (gfx950) implementation of
 * Carcassonne's variational-sweep hot path.
 *
 * Conventions
 *   - Every function returns an int status (CC_OK = 0, negative on error);
 *     cc_last_error() returns a thread-local message for the last failure.
 *   - Pointers are device pointers unless named host_*; sizes are element
 *     counts; tensors are dense row-major unless strides are passed.
 *   - All work is stream-ordered on the caller's hipStream_t (passed as void*).
 *     The caller owns every buffer; the library never frees caller memory.
 *   - Reductions run in a fixed order, so results do not depend on scheduling.
 *
 * Each entry point names the reference (gcross/Carcassonne) interface whose
 * arithmetic it carries out, as file:line in that repository.
 */
#ifndef CARCASSONNE_HIP_H
#define CARCASSONNE_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* status codes -- the Python layer maps them to the reference's exceptions */
#define CC_OK 0
#define CC_ERR_RANK (-1)          /* UnexpectedTensorRankError  carcassonne/utils.py:34-41 */
#define CC_ERR_DIM (-2)           /* DimensionMismatchError     carcassonne/utils.py:13-23 */
#define CC_ERR_NOT_CONVERGED (-3) /* RelaxFailed                carcassonne/utils.py:25-33 */
#define CC_ERR_HIP (-4)           /* HIP runtime failure */
#define CC_ERR_OOM (-5)           /* device allocation failure */
#define CC_ERR_ARG (-6)           /* ValueError */
#define CC_ERR_UNSUPPORTED (-7)   /* shape / dtype without a kernel */

/* element types */
#define CC_F64 0
#define CC_F32 1
#define CC_F16 2
#define CC_BF16 3

int cc_version(void);
const char* cc_last_error(void);

/* ---- Data protocol primitives (NDArrayData, carcassonne/data/__init__.py) ---- */

/* dst = src viewed with `shape` and element strides `src_strides`, written densely
 * (row-major).  NDArrayData.join / transpose   data/__init__.py:247-256, 347-349 */
int cc_permute(void* stream, int dtype, const void* src, void* dst, int rank,
               const int64_t* shape, const int64_t* src_strides);

/* C[b] = alpha * A[b] @ B[b] + beta * C[b] with arbitrary element strides
 * (A: b,m,k  B: b,k,n  C: b,m,n).  NDArrayData.contractWith (numpy.tensordot)
 * data/__init__.py:157-159; every generated contractor of utils.py:562-711 */
int cc_gemm_f64(void* stream, int64_t batch, int64_t M, int64_t N, int64_t K, double alpha,
                const double* A, int64_t sAb, int64_t sAm, int64_t sAk,
                const double* B, int64_t sBb, int64_t sBk, int64_t sBn, double beta,
                double* C, int64_t sCb, int64_t sCm, int64_t sCn);

/* y = alpha * x + beta * y.   NDArrayData.__add__/__iadd__/__mul__  data/__init__.py:93-139 */
int cc_axpby_f64(void* stream, int64_t n, double alpha, const double* x, double beta, double* y);
/* y = (alpha_scale * alpha_dev[0]) * x + beta * y with the coefficient read from device memory (solver steps
 * that must not synchronise with the host).  Same arithmetic as cc_axpby_f64. */
int cc_axpby_dev_f64(void* stream, int64_t n, const double* alpha_dev, double alpha_scale, const double* x, double beta,
                     double* y);

/* y[i] = value.   newZeros / newFilled / newTrivial  data/__init__.py:51-90 */
int cc_fill_f64(void* stream, int64_t n, double value, double* y);

/* y[i*(n+1)] = 1 on an n x n zeroed matrix.   newIdentity  data/__init__.py:59-62 */
int cc_set_identity_f64(void* stream, int64_t n, double* y);

/* out = x (op) y elementwise, op: 0 mul, 1 div, 2 add, 3 sub.
 * NDArrayData.__mul__/__truediv__/__add__/__sub__  data/__init__.py:93-139 */
int cc_binary_f64(void* stream, int op, int64_t n, const double* x, const double* y, double* out);

/* out = f(x) elementwise, op: 0 sqrt, 1 reciprocal where |x| > param (else x), 2 abs,
 * 3 x * param, 4 reciprocal.   NDArrayData.sqrt and the 1/s cut-off of normalizeAxis
 * data/__init__.py:287-293, 341-343 */
int cc_unary_f64(void* stream, int op, int64_t n, const double* x, double param, double* out);

/* out[i][j] = A[i][j] * s[i] (cols = 0) or * s[j] (cols = 1).  The broadcast V*S of
 * normalizeAxis  data/__init__.py:295-300 */
int cc_scale_rows_cols_f64(void* stream, int cols, int64_t m, int64_t n, const double* A, const double* s, double* out);

/* out[0] = sum_i x[i] * y[i] (fixed-order).   contractWith over all legs,
 * computeScalarUsingMultiplier  system/_2d.py:429-431; norm  data/__init__.py:260-262 */
int cc_dot_f64(void* stream, int64_t n, const double* x, const double* y, double* out);

/* out[0] = number of non-finite entries.   hasNaN  data/__init__.py:209-211 */
int cc_count_nonfinite_f64(void* stream, int64_t n, const double* x, int64_t* out);

/* dst = (dst_dtype) src.  Precision conversion of slabs and vectors (no reference counterpart). */
int cc_cast(void* stream, int64_t n, int src_dtype, const void* src, int dst_dtype, void* dst);

/* Thin SVD of a batch of m x n row-major matrices (m >= n, n <= 256) by one-sided
 * Jacobi: A = U diag(S) Vt with U m x n, S n (descending), Vt n x n.  `work` holds
 * cc_svd_work_size() doubles.
 * NDArrayData.svd / normalizeAxis / unitize  data/__init__.py:263-301,344-352; utils.py:879-882 */
int cc_svd_work_size(int64_t batch, int64_t m, int64_t n, int64_t* out);
int cc_svd_f64(void* stream, int64_t batch, int64_t m, int64_t n, const double* A,
               double* U, double* S, double* Vt, double* work);

/* Truncation basis of a fused environment bond (the state compression of compressCornerStateTowards,
 * system/_2d.py:191-228 with computeProductCompressor, compression.py:26-45, for a bond whose ket, bra and
 * operator legs are fused into one index).  C is the corner's Identity channel matricized as
 * (x, a, b) x far: the fused bond (x = the corner's old bond, a / b = the absorbed side's center ket / bra
 * legs) against the corner's far bond (far <= 256, x a^2 >= far).  The rows are symmetrized under the
 * a <-> b exchange (so the projected environment keeps its ket <-> bra symmetry); the singular vectors come
 * from the eigen-decomposition of the far x far Gram matrix (U = Cs V S^-1; columns whose singular value is
 * at most 1e-7 of the largest are zero) and U_out (x a^2 by
 * new_dim, orthonormal columns, signs fixed by cc_sign_normalize_columns_f64) receives the leading new_dim
 * left singular vectors, S_out (far) all
 * singular values, descending.  new_dim <= far: the corner's rank bound, at which the Identity-channel
 * product of corner and side is kept exactly.  `work` holds cc_fused_truncation_work_size doubles. */
int cc_fused_truncation_work_size(int64_t nx, int64_t na, int64_t far, int64_t* out);
/* Fixes the sign freedom of singular vectors: each column of the m x n row-major U is negated when its entry of
 * largest magnitude (the first such row) is negative.  Applied to the bases of cc_fused_truncation_f64 and to
 * the deterministic compressor start, so that the GPU and the oracle take the same gauge (the absorptions that
 * follow identify the two bonds of a side, so the gauge is not free there). */
int cc_sign_normalize_columns_f64(void* stream, int64_t m, int64_t n, double* U);
int cc_fused_truncation_f64(void* stream, int64_t nx, int64_t na, int64_t far, int64_t new_dim, const double* C,
                            double* U_out, double* S_out, double* work);

/* ---- stage 3: the center-site effective-Hamiltonian matvec ----------------------
 * Y = sum_t (A_{a(t)} (x) B_{b(t)} (x) O_{o(t)}) v, summed over alpha, with
 *   A_a[alpha][p'][p] (P x P slabs), B_b[alpha][q'][q] (Q x Q slabs), v[p][q][s] (P,Q,d),
 *   O_o[s'][s] (d x d, o = -1 means identity).
 * formNormalizationStage3 / formDenseStage3_multiply  tensors/_2d/dense.py:115-173,
 * summed as multiplyExpectation  tensors/_2d/sparse.py:129-133.
 * Generic path: any P, Q, d, FP64, via cc_gemm_f64-class kernels; `work` must hold
 * n_alpha*P*Q*d doubles. */
int cc_stage3_matvec_generic_work_size(int64_t n_alpha, int64_t P, int64_t Q, int64_t d, int64_t* out);
int cc_stage3_matvec_generic_f64(void* stream, int n_terms, const double* const* A_slabs,
                                 const double* const* B_slabs, const double* const* O_mats,
                                 int64_t n_alpha, int64_t P, int64_t Q, int64_t d,
                                 const double* v, double* y, double* work);

/* Fused FP64 MFMA stage-3 matvec for slab extents P = Q in {4, 9, 16, 25, 36} and d = 2
 * (D = 2..6).  Slabs are first packed in MFMA-operand order with cc_stage3_pack_f64
 * (cc_stage3_packed_size doubles per channel).  Terms (a, b, o) must be sorted by a;
 * o indexes host_ops (n_ops row-major d x d matrices) or is -1 for the identity.
 * `work` holds cc_stage3_fused_work_size doubles.  Same arithmetic as
 * cc_stage3_matvec_generic_f64 (tensors/_2d/dense.py:115-173, sparse.py:129-133). */
int cc_stage3_fused_supported(int64_t P, int64_t d);
int cc_stage3_packed_size(int64_t n_alpha, int64_t P, int64_t* out);
int cc_stage3_fused_work_size(int64_t n_alpha, int64_t P, int64_t d, int64_t* out);
int cc_stage3_pack_f64(void* stream, int64_t n_alpha, int64_t P, const double* slabs, double* packed);
/* The same packing read in place from a stage-2 tensor: slab[alpha][p'][p] with alpha = a0 * na1 + a1,
 * p' = r0 * d1 + r1, p = c0 * d1 + c1 (P = d0 * d1) at element offset a0 s[0] + a1 s[1] + r0 s[2] +
 * r1 s[3] + c0 s[4] + c1 s[5] of src (strides = s, six entries).  This replaces the join of the stage-2
 * tensor into (alpha, P, P) slabs (stage2_0.join((0,1),4,5,2,3), stage2_1.join((1,0),4,5,2,3);
 * tensors/_2d/dense.py:115-173) and the packing in one pass (P <= 64). */
int cc_stage3_pack_strided_f64(void* stream, int64_t na0, int64_t na1, int64_t d0, int64_t d1, const int64_t* strides,
                               const double* src, double* packed);
int cc_stage3_matvec_fused_f64(void* stream, int n_a, const double* const* A_packed, int n_b,
                               const double* const* B_packed, int n_terms, const int* term_a,
                               const int* term_b, const int* term_o, int n_ops, const double* host_ops,
                               int64_t n_alpha, int64_t P, int64_t d, const double* v, double* y,
                               double* work);

/* cc_stage3_matvec_fused_f64 without any allocation and without a workspace argument: no counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The workgroup partial sums live in a file-scope device array of
 * CC_NOALLOC_SCRATCH_DOUBLES doubles (8 MiB) that every GPU the library runs on holds until the process ends;
 * its address is looked up once per device on the first call made with that device current (the stream must
 * belong to the current device).  Problems whose partial sums exceed it (groups * P^2 * d doubles, groups =
 * min(1024, ceil(n_alpha / 16)); e.g. n_alpha > 6464 at P = 36) return CC_ERR_ARG.  All calls on one GPU share
 * the array: calls that overlap in time on one GPU (different streams, or different host threads) race. */
#define CC_NOALLOC_SCRATCH_DOUBLES 1048576
int cc_stage3_matvec_fused_f64_noalloc(void* stream, int n_a, const double* const* A_packed, int n_b,
                                       const double* const* B_packed, int n_terms, const int* term_a,
                                       const int* term_b, const int* term_o, int n_ops, const double* host_ops,
                                       int64_t n_alpha, int64_t P, int64_t d, const double* v, double* y);

/* The fused matvec with FP32 operands on v_mfma_f32_16x16x4f32 and FP32 accumulation (the workgroup partials
 * are summed in FP64): the FP32 form of the plain method (no reference counterpart; the arithmetic is that of
 * tensors/_2d/dense.py:115-173 summed as sparse.py:129-133).  Slabs are packed once per channel from the
 * stage-2 tensor in place (as cc_stage3_pack_strided_f64) into FP32 operand order, role 0 for A channels and
 * role 1 for B channels (the B slabs take the K order in which the stage-A result is the stage-B operand);
 * cc_stage3_packed_size_f32 gives the floats per channel.  cc_stage3_plan_create_f32 makes a plan that
 * cc_stage3_plan_apply, cc_lanczos_f64 and cc_stage3_plan_destroy take like an FP64 plan; x and y stay FP64,
 * `work` holds cc_stage3_fused_work_size doubles. */
int cc_stage3_packed_size_f32(int64_t n_alpha, int64_t P, int role, int64_t* out);
int cc_stage3_pack_strided_f32(void* stream, int role, int64_t na0, int64_t na1, int64_t d0, int64_t d1,
                               const int64_t* strides, const double* src, float* packed);
int cc_stage3_plan_create_f32(void** out, int n_a, const float* const* A_packed, int n_b,
                              const float* const* B_packed, int n_terms, const int* term_a, const int* term_b,
                              const int* term_o, int n_ops, const double* host_ops, int64_t n_alpha, int64_t P,
                              int64_t d, double* work);

/* Persistent fused-operator plan: validated and copied once, applied many times (the
 * operator of one solve).  `work` (cc_stage3_fused_work_size doubles) stays caller-owned.
 * The plan object holding them is freed by cc_stage3_plan_destroy. */
int cc_stage3_plan_create(void** out, int n_a, const double* const* A_packed, int n_b,
                          const double* const* B_packed, int n_terms, const int* term_a,
                          const int* term_b, const int* term_o, int n_ops, const double* host_ops,
                          int64_t n_alpha, int64_t P, int64_t d, double* work);
int cc_stage3_plan_apply(void* stream, const void* plan, const double* x, double* y);
int cc_stage3_plan_destroy(void* plan);

/* Timing of the stage-3 kernels: between begin and end every launch of the FP64 fused kernel (kind 0)
 * and of the low-precision kernel (kind 1) is bracketed by HIP events on its stream; end fills
 * total_ms[k] and launches[k] (the launches timed) for k < n_kinds.  begin_sampled(every) times only
 * launches 0, every, 2 every, ... of each kind, so that the event records (each idles the GPU for a few
 * microseconds) cost a fraction of the timed work.  Single-threaded use only.  (Measurement support;
 * no reference counterpart.) */
int cc_profile_begin(void);
int cc_profile_begin_sampled(int every);
/* Reads n_bytes (a multiple of 16384) of src with direct global->LDS loads, the load path of the FP16
 * stage-3 kernel, over `blocks` workgroups, writing one float per workgroup to sink (blocks floats): a kernel
 * of known fetched bytes that calibrates the FETCH_SIZE hardware counter for that path.  (Measurement
 * support; no reference counterpart.) */
int cc_calibration_lds_read(void* stream, int64_t n_bytes, const void* src, float* sink, int blocks);
/* Matrix-core throughput check: `blocks` workgroups of four waves, each wave issuing `iters` rounds of eight
 * independent MFMAs (kind 0: v_mfma_f32_16x16x32_f16, 16384 FLOP; kind 1: v_mfma_f64_16x16x4f64, 2048 FLOP;
 * kind 2: v_mfma_f32_16x16x16_f16, 8192 FLOP),
 * writing blocks x 256 floats to sink: confirms the published dense peaks on the machine the figures are
 * priced against.  (Measurement support; no reference counterpart.) */
int cc_calibration_mfma(void* stream, int kind, int iters, int blocks, float* sink);
int cc_profile_end(int n_kinds, double* total_ms, int64_t* launches);

/* ---- dense linear algebra of the solve --------------------------------------------
 * In-place blocked Cholesky A = L L^T (lower; upper zeroed); info_dev[0] = 0 or the first
 * non-positive pivot (1-based); `work` holds cc_linalg_work_size doubles.  The LU of the normalisation matrix in relaxOver
 * (utils.py:814-817) for the symmetric positive definite case. */
int cc_linalg_work_size(int64_t n, int64_t* out);
int cc_cholesky_f64(void* stream, int64_t n, double* A, int* info_dev, double* work);
/* X = L^-1 for lower-triangular L (n x n); `work` holds cc_linalg_work_size doubles. */
int cc_trtri_lower_f64(void* stream, int64_t n, const double* L, double* X, double* work);
/* A = L L^T in place (upper zeroed) and X = L^-1 in one recursive pass; info_dev as cc_cholesky_f64;
 * `work` holds cc_linalg_work_size doubles.  The normalization factor of the whitened eigenproblem
 * (the reference inverts N through relaxOver's LU, utils.py:814-817). */
int cc_cholesky_inverse_f64(void* stream, int64_t n, double* A, double* X, int* info_dev, double* work);
/* B = A^T (n x n). */
int cc_transpose_f64(void* stream, int64_t n, const double* A, double* B);
/* Y (n x nrhs) = M (n x n) X restricted to tri = 0 full, 1 lower, 2 upper; nrhs <= 4. */
int cc_skinny_f64(void* stream, int64_t n, int tri, const double* M, int nrhs, const double* X, double* Y);

/* Lanczos with full reorthogonalisation for the lowest eigenpair of the pencil (H, N),
 * N = (L L^T) (x) I_d, on the whitened operator W H W^T (W = L^-1).  H is a stage-3 plan.
 * Stops when the Ritz residual <= tol_residual*|theta|, when the Ritz value moved by at most
 * tol_energy*|theta| since the previous check (every check_every steps), or at max_iter.
 * x_out is N-normalised and written in stream order (the call does not wait for it); host_result =
 * {theta, residual estimate, iterations}.  check_every = 0 runs
 * exactly max_iter steps (fewer if the Krylov space completes) without any host read or
 * synchronisation, so the call can be captured in a HIP graph; host_result (may be null) then receives
 * NaN for theta and the residual, which stay in device memory.
 * Replaces relaxOver (utils.py:805-878) as the residual-controlled solver of
 * minimizeExpectation (system/_2d.py:489-497). */
int cc_lanczos_work_size(int64_t n, int max_iter, int64_t* out);
int cc_lanczos_f64(void* stream, const void* plan, int64_t n_env, int64_t d, const double* L,
                   const double* Lt, const double* W, const double* Wt, const double* x0,
                   double tol_residual, double tol_energy, int max_iter, int check_every,
                   double* work, double* x_out, double* host_result);

/* ---- precision form: FP16 / BF16 stage-3 matvec with FP32 accumulation ---------------
 * (no reference counterpart; the same arithmetic as formDenseStage3 at reduced precision)
 * Low-precision slabs are stored (n_alpha, P, P) with the per-alpha stride rounded up to a multiple
 * of 8 elements (zero padding); cc_stage3_lp_slab_size gives the element count of one channel.
 * cc_stage3_lp_convert: out = slab / max|slab| rounded to dtype (slab is the dense FP64 (n_alpha, P, P)
 * channel); scale_out (513 doubles of scratch) receives max|slab| in scale_out[0].
 * cc_stage3_lp_plan_create: term_coef[t] is the d x d matrix scale_a * scale_b * O_t (identity for a
 * normalization-type term); split 1 carries the vector and the intermediate as hi + lo pairs (only
 * the slabs are rounded), split 2 only the vector (the intermediate is rounded once), split 0 neither;
 * the plan is uploaded on `stream` without a host wait; channel_scales (device, n_a + n_b doubles:
 * A channels then B channels, e.g. the scale_out[0] of each conversion) may be null -- otherwise
 * term_coef holds the unscaled O_t and the scales are applied on the device; `work` holds cc_stage3_lp_work_size floats and receives the device copy of
 * the plan.  The terms are cut into chunks whose slabs fit the kernel's LDS buffers;
 * cc_stage3_lp_plan_chunks reports how many.  Apply computes y (FP64) = H_lp x (FP64). */
int cc_stage3_lp_slab_size(int64_t n_alpha, int64_t P, int64_t* out);
int cc_stage3_lp_convert(void* stream, int dtype, int64_t n_alpha, int64_t P, const double* slab, double* scale_out,
                         void* out);
/* cc_stage3_lp_convert reading the channel in place through the slab map of cc_stage3_pack_strided_f64
 * (P <= 64).  src_elems: the element count of src when the map covers every element of src exactly once
 * (the maximum is then taken over src contiguously), else 0. */
int cc_stage3_lp_convert_strided(void* stream, int dtype, int64_t na0, int64_t na1, int64_t d0, int64_t d1,
                                 const int64_t* strides, int64_t src_elems, const double* src, double* scale_out,
                                 void* out);
int cc_stage3_lp_work_size(int64_t n_alpha, int64_t P, int64_t d, int groups, int64_t* out);
int cc_stage3_lp_plan_create(void** out, int dtype, int n_a, const void* const* A, int n_b, const void* const* B,
                             int n_terms, const int* term_a, const int* term_b, const float* term_coef,
                             int64_t n_alpha, int64_t P, int64_t d, int groups, int split, float* work,
                             void* stream, const double* channel_scales);
int cc_stage3_lp_plan_chunks(const void* plan, int* out);
/* out = 1 when every term's coefficient matrix is diagonal or anti-diagonal (d = 2) and the plan runs the
 * kernel's diagonal-form combine, 0 when some term has a full coefficient matrix. */
int cc_stage3_lp_plan_diagonal_form(const void* plan, int* out);
/* Recomputes the plan's device coefficients from the unscaled term coefficients given at creation and
 * new channel scales (same layout as at creation), on `stream` with one kernel and no host work: after
 * the plan's slab buffers are converted again from a new environment of the same structure (the
 * reference rebuilds the stage-3 multipliers for every solve, tensors/_2d/sparse.py:59-71), this makes the
 * plan apply the new operator; a kernel-only sequence, so it can be captured in a HIP graph. */
int cc_stage3_lp_plan_rescale(void* stream, void* plan, const double* channel_scales);
int cc_stage3_lp_plan_apply(void* stream, const void* plan, const double* x, double* y);
int cc_stage3_lp_plan_destroy(void* plan);

/* Lanczos (as cc_lanczos_f64) on a low-precision plan. */
int cc_lanczos_lp(void* stream, const void* lp_plan, int64_t n_env, int64_t d, const double* L, const double* Lt,
                  const double* W, const double* Wt, const double* x0, double tol_residual, double tol_energy,
                  int max_iter, int check_every, double* work, double* x_out, double* host_result);

/* Jacobi-Davidson refinement pieces (FP64):
 * out[j*stride_out] = U[j] . v for j < k;  out = sum_j c[j] U[j];  lowest eigenpair of the k x k
 * pencil (Hs, Ns), k <= 32, out[0] = value, vec = N-normalised coefficients;  the correction
 * t = W^T P y with y from `iters` CG steps on P (W H_lp W^T - rho) P y = -P W r (P = I - u u^T,
 * u = L^T x / |L^T x|, rho = rho_dev[0] read on the device); `work` holds cc_jd_work_size doubles. */
int cc_basis_dots_f64(void* stream, int k, int64_t n, const double* U, const double* v, double* out, int stride_out);
int cc_basis_combine_f64(void* stream, int k, int64_t n, const double* U, const double* c, double* out);
int cc_small_geneig_f64(void* stream, int k, const double* Hs, const double* Ns, double* out, double* vec);
int cc_jd_work_size(int64_t n, int64_t* out);
int cc_jd_correction_lp(void* stream, const void* lp_plan, int64_t n_env, int64_t d, const double* Lt,
                        const double* W, const double* Wt, const double* rho_dev, const double* x, const double* r,
                        int iters, double* work, double* t_out);
/* y = x / |x| (one workgroup, fixed-order sum; the Lanczos normalisation). */
int cc_normalize_f64(void* stream, int64_t n, const double* x, double* y);

/* ---- learned warm start: a 3-layer MLP proposing the eigensolver's start vector ----------------
 * (no reference counterpart: the reference always starts relaxOver from the current center,
 * system/_2d.py:491-492; the plain solve stays the arbiter of every proposal)
 * Network x (K0) -> relu(W1 x + b1) (H) -> relu(W2 . + b2) (H) -> W3 . + b3 (NO); parameters in one FP32
 * buffer [W1 (H x K0), b1, W2 (H x H), b2, W3 (NO x H), b3] (cc_mlp_param_count floats) with an FP16
 * mirror P16 of the same layout.  K0 % 16 == 0; H, NO multiples of 64; X16 is B x K0 FP16.
 * cc_mlp_features: one FP16 feature row from n_spec spectra of k singular values (each divided by its
 *   largest value) followed by h, zero-padded to K0.
 * cc_mlp_forward: FP16 operands, FP32 accumulation on matrix cores; A16 (2 B H halves) and A32 (2 B H
 *   floats) receive the hidden layers, Y (B x NO floats) the output.
 * cc_mlp_backward: loss[b] = 1 - (y_b . v_b)^2 / (|y_b|^2 |v_b|^2) against target v (B x NO floats) and the
 *   gradient of the batch mean into G (parameter layout); work holds cc_mlp_work_size floats.
 * cc_mlp_adam: one Adam step on P32 with moments M, V (step counts from 1); refreshes P16. */
int cc_mlp_param_count(int64_t K0, int64_t H, int64_t NO, int64_t* out);
int cc_mlp_work_size(int64_t B, int64_t H, int64_t NO, int64_t* out);
int cc_mlp_features(void* stream, int n_spec, int k, const double* S, double h, int64_t K0, void* out16);
int cc_mlp_forward(void* stream, int64_t B, int64_t K0, int64_t H, int64_t NO, const void* P16, const float* P32,
                   const void* X16, void* A16, float* A32, float* Y);
int cc_mlp_backward(void* stream, int64_t B, int64_t K0, int64_t H, int64_t NO, const void* P16, const void* X16,
                    const void* A16, const float* A32, const float* Y, const float* target, float* loss, float* G,
                    float* work);
int cc_mlp_adam(void* stream, int64_t n, float* P32, const float* G, float* M, float* V, void* P16, double lr,
                double beta1, double beta2, double eps, int64_t step);

#ifdef __cplusplus
}
#endif
#endif
