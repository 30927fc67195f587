// Lanczos for the lowest eigenpair of the pencil (H, N), N = (L L^T) (x) I_d, on the whitened
// operator W H W^T (W = L^-1 (x) I_d), with full (twice-applied) reorthogonalisation.  All
// vectors and recurrence scalars stay in device memory; the host loop only reads two doubles
// (Ritz value and residual estimate) every `check_every` steps.
#include "cc_common.h"

#include <cmath>
#include <vector>

namespace {

// sum_i a[i] b[i] over the 256 threads of a workgroup in a fixed order: each thread loads its DOT_U
// pairs at once (clamped index, masked product), then a tree over the threads.
constexpr int DOT_U = 12;
__device__ inline double block_dot256(int64_t n, const double* __restrict__ a, const double* __restrict__ b, double* red) {
  double acc = 0.0;
  for (int64_t base = threadIdx.x; base < n; base += 256 * DOT_U) {
    double x[DOT_U], y[DOT_U];
#pragma unroll
    for (int u = 0; u < DOT_U; ++u) {
      const int64_t i = base + 256 * u < n ? base + 256 * u : base;
      x[u] = a[i];
      y[u] = b[i];
    }
#pragma unroll
    for (int u = 0; u < DOT_U; ++u)
      if (base + 256 * u < n) acc += x[u] * y[u];
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  return red[0];
}

// c[b] = Q_b . w for b < nq; a block b == nq (when launched) gives c[nq] = w . w
__global__ void __launch_bounds__(256) lz_coeff_kernel(int64_t n, const double* __restrict__ Q, const double* __restrict__ w,
                                                       double* __restrict__ c, int nq) {
  __shared__ double red[256];
  const double* q = (int)blockIdx.x == nq ? w : Q + (int64_t)blockIdx.x * n;
  const double t = block_dot256(n, q, w, red);
  if (threadIdx.x == 0) c[blockIdx.x] = t;
}

// w[i] -= sum_j c[j] Q[j][i]   (or z[i] = sum_j c[j] Q[j][i] when `assign`): block = 16 outputs x
// 16 basis slices, slice sums combined in a fixed order.
// w -= Q^T c (assign: w = Q^T c) over the first m basis vectors; record (when given) receives c[m - 1]
__global__ void __launch_bounds__(256) lz_combine_kernel(int m, int64_t n, const double* __restrict__ Q,
                                                         const double* __restrict__ c, double* __restrict__ w, int assign,
                                                         double* __restrict__ record = nullptr) {
  __shared__ double red[16][16];
  if (record != nullptr && blockIdx.x == 0 && threadIdx.x == 0) *record = c[m - 1];
  const int e = threadIdx.x & 15, slice = threadIdx.x >> 4;
  const int64_t i = (int64_t)blockIdx.x * 16 + e;
  double acc = 0.0;
  if (i < n)
    for (int j = slice; j < m; j += 16) acc += c[j] * Q[(int64_t)j * n + i];
  red[slice][e] = acc;
  __syncthreads();
  if (slice == 0 && i < n) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += red[k][e];
    w[i] = assign ? t : w[i] - t;
  }
}

// The second Gram-Schmidt pass and the normalisation in one kernel: q = (w - Q c) / beta with
// beta^2 = w.w - c.c (c[m] holds w.w from lz_coeff_kernel; exact for orthonormal Q, and free of
// cancellation because the second pass's coefficients are at rounding level), beta written by block 0.
__global__ void __launch_bounds__(256) lz_combine_normalize_kernel(int m, int64_t n, const double* __restrict__ Q,
                                                                   const double* __restrict__ c, const double* __restrict__ w,
                                                                   double* __restrict__ q, double* __restrict__ beta) {
  __shared__ double red[16][16];
  const int e = threadIdx.x & 15, slice = threadIdx.x >> 4;
  const int64_t i = (int64_t)blockIdx.x * 16 + e;
  double acc = 0.0;
  if (i < n)
    for (int j = slice; j < m; j += 16) acc += c[j] * Q[(int64_t)j * n + i];
  red[slice][e] = acc;
  double cc = 0.0;
  for (int j = 0; j < m; ++j) cc += c[j] * c[j];
  const double b2 = c[m] - cc, b = b2 > 0.0 ? sqrt(b2) : 0.0, inv = b > 0.0 ? 1.0 / b : 0.0;
  __syncthreads();
  if (slice == 0 && i < n) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += red[k][e];
    q[i] = (w[i] - t) * inv;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) *beta = b;
}

// w -= alpha q_j + beta q_{j-1}  (alpha, beta read from device memory)

// beta = ||w||, q = w / beta (single 1024-thread block, fixed-order tree)
__global__ void __launch_bounds__(1024) lz_normalize_kernel(int64_t n, const double* __restrict__ w, double* __restrict__ q,
                                                            double* __restrict__ beta) {
  __shared__ double red[1024];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) acc += w[i] * w[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  const double b = sqrt(red[0]);
  const double inv = b > 0.0 ? 1.0 / b : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) q[i] = w[i] * inv;
  if (threadIdx.x == 0 && beta) *beta = b;
}

constexpr int LZ_LDS_M = 640;

// True when T - x I is positive definite, i.e. x lies below the lowest eigenvalue of the tridiagonal
// (a, b): every leading principal minor is positive.  The minors follow p_{i+1} = (a_i - x) p_i -
// b_{i-1}^2 p_{i-1} (one dependent FMA per step, no division); every fourth step both are rescaled by
// the same power of two, which changes no sign.  b2[i] = b[i]^2.
__device__ bool below_spectrum(int m, const double* a, const double* b2, double x) {
  double pm = 1.0, p = a[0] - x;
  bool pd = p > 0.0;
#pragma unroll 8
  for (int i = 1; i < m; ++i) {  // unrolled: the loads of a and b2 run ahead of the recurrence
    const double pn = (a[i] - x) * p - b2[i - 1] * pm;
    pm = p;
    p = pn;
    pd = pd && (p > 0.0);
    if ((i & 3) == 0) {
      int e;
      (void)frexp(p, &e);
      p = ldexp(p, -e);
      pm = ldexp(pm, -e);
    }
  }
  return pd;
}

// Lowest eigenpair of the m x m tridiagonal (a, b): multisection over one wave, then inverse
// iteration (LU with partial pivoting) on lane 0.  out[0] = theta, out[1] = |b[m-1] * s[m-1]|.
__global__ void __launch_bounds__(64) lz_tridiag_kernel(int m, const double* __restrict__ ga, const double* __restrict__ gb,
                                                        double* __restrict__ out, double* __restrict__ s,
                                                        double* __restrict__ scratch, int want_vector) {
  const int lane = threadIdx.x;
  // stage the tridiagonal in LDS (every lane reads all of it in the Sturm sequences)
  __shared__ double lds_ab[3 * LZ_LDS_M];
  const bool in_lds = m <= LZ_LDS_M;
  if (in_lds) {
    for (int i = lane; i < m; i += 64) {
      lds_ab[i] = ga[i];
      lds_ab[LZ_LDS_M + i] = gb[i];
      lds_ab[2 * LZ_LDS_M + i] = gb[i] * gb[i];
    }
    __syncthreads();
  } else {
    for (int i = lane; i < m; i += 64) scratch[i] = gb[i] * gb[i];   // b^2 past the LDS capacity
    __syncthreads();
  }
  const double* a = in_lds ? lds_ab : ga;
  const double* b = in_lds ? lds_ab + LZ_LDS_M : gb;
  const double* b2 = in_lds ? lds_ab + 2 * LZ_LDS_M : scratch;
  double lo = INFINITY, hi = -INFINITY, tnorm = 0.0;
  for (int i = lane; i < m; i += 64) {
    double r = (i > 0 ? fabs(b[i - 1]) : 0.0) + (i < m - 1 ? fabs(b[i]) : 0.0);
    lo = fmin(lo, a[i] - r);
    hi = fmax(hi, a[i] + r);
    tnorm = fmax(tnorm, fabs(a[i]) + r);
  }
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    tnorm = fmax(tnorm, __shfl_xor(tnorm, o, 64));
  }
  const double eps = 2.220446049250313e-16;
  const double pivmin = fmax(tnorm * 1e-300, 1e-300) + 4.9e-324;
  lo -= 2.0 * eps * tnorm + pivmin;
  hi += 2.0 * eps * tnorm + pivmin;
  // interlacing: the lowest Ritz value never increases with m, so the previous one bounds it
  if (!isnan(out[0])) hi = fmin(hi, out[0] + 4.0 * eps * tnorm + pivmin);
  for (int round = 0; round < 20; ++round) {
    double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    unsigned long long mask = __ballot(!below_spectrum(m, a, b2, x));
    if (mask == 0ULL) {
      lo = lo + (hi - lo) * 64.0 / 65.0;
    } else {
      int first = __ffsll((long long)mask) - 1;
      double nhi = lo + (hi - lo) * (double)(first + 1) / 65.0;
      double nlo = first == 0 ? lo : lo + (hi - lo) * (double)first / 65.0;
      lo = nlo;
      hi = nhi;
    }
    if (hi - lo <= 2.0 * eps * fmax(fabs(lo), fabs(hi)) + pivmin) break;
  }
  const double theta = 0.5 * (lo + hi);
  if (!want_vector) {
    if (lane == 0) {
      out[0] = theta;
      out[1] = NAN;
    }
    return;
  }
  // One step of inverse iteration on T - theta I from a fixed start vector: theta is the eigenvalue to
  // a couple of ulps, so one solve already gives the eigenvector.  Lane 0 runs the LU with partial
  // pivoting (the tridiagonal form of LAPACK gttrf) fused with the forward substitution, carrying the
  // running row in registers and storing reciprocal pivots; the back substitution is then a short FMA
  // chain, and the whole wave normalises.  Scratch lives in LDS when it fits, else in global memory.
  __shared__ double lds_scr[4 * LZ_LDS_M];
  double* base = (m <= LZ_LDS_M) ? lds_scr : scratch;
  double* rd = base;            // 1 / pivot of row i
  double* du = base + m;        // first superdiagonal of U
  double* du2 = base + 2 * m;   // second superdiagonal of U (from row swaps)
  double* sv = (m <= LZ_LDS_M) ? base + 3 * m : s;
  __syncthreads();              // every lane is done reading b^2 from the scratch before it is reused
  if (lane == 0) {
    const double tiny = eps * fmax(tnorm, 1e-300);
    auto start = [](int i) { return 1.0 + 0.01 * (double)((i * 7919) % 101) / 101.0; };
    double cd = a[0] - theta, cdu = m > 1 ? b[0] : 0.0, cs = start(0);
#pragma unroll 4
    for (int i = 0; i < m - 1; ++i) {
      const double dn = a[i + 1] - theta, dli = b[i], dun = (i + 1 < m - 1) ? b[i + 1] : 0.0, ns = start(i + 1);
      if (fabs(cd) >= fabs(dli)) {
        if (cd == 0.0) cd = tiny;
        const double f = dli / cd;
        rd[i] = 1.0 / cd;
        du[i] = cdu;
        du2[i] = 0.0;
        sv[i] = cs;
        cd = dn - f * cdu;
        cdu = dun;
        cs = ns - f * cs;
      } else {
        const double f = cd / dli;
        rd[i] = 1.0 / dli;
        du[i] = dn;
        du2[i] = dun;
        sv[i] = ns;
        cd = cdu - f * dn;
        cdu = -f * dun;
        cs = cs - f * ns;
      }
    }
    if (cd == 0.0) cd = tiny;
    rd[m - 1] = 1.0 / cd;
    double x1 = cs * rd[m - 1], x2 = 0.0;
    sv[m - 1] = x1;
#pragma unroll 4
    for (int i = m - 2; i >= 0; --i) {
      const double x = (sv[i] - du[i] * x1 - du2[i] * x2) * rd[i];
      sv[i] = x;
      x2 = x1;
      x1 = x;
    }
  }
  __syncthreads();
  double big = 0.0;
  for (int i = lane; i < m; i += 64) big = fmax(big, fabs(sv[i]));
  for (int o = 32; o > 0; o >>= 1) big = fmax(big, __shfl_xor(big, o, 64));
  const double ib = big > 0.0 ? 1.0 / big : 1.0;
  double nrm = 0.0;
  for (int i = lane; i < m; i += 64) { const double v = sv[i] * ib; nrm += v * v; }
  for (int o = 32; o > 0; o >>= 1) nrm += __shfl_xor(nrm, o, 64);
  const double sc = ib / sqrt(nrm);
  for (int i = lane; i < m; i += 64) s[i] = sv[i] * sc;
  __syncthreads();
  if (lane == 0) {
    out[0] = theta;
    out[1] = fabs(b[m - 1] * s[m - 1]);
  }
}

__global__ void set_nan_kernel(double* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = NAN;
}

struct Whitening {
  int64_t n_env, d;
  const double *L, *Lt, *W, *Wt;
};

typedef int (*ApplyFn)(hipStream_t, const void*, const double*, double*);

int op_whitened(hipStream_t st, ApplyFn apply, const void* plan, const Whitening& wh, const double* q, double* t1, double* t2,
                double* w) {
  int rc = cc_skinny_f64(st, wh.n_env, 2, wh.Wt, (int)wh.d, q, t1);  // t1 = W^T q
  if (rc) return rc;
  rc = apply(st, plan, t1, t2);                                      // t2 = H t1
  if (rc) return rc;
  return cc_skinny_f64(st, wh.n_env, 1, wh.W, (int)wh.d, t2, w);     // w = W t2
}
}  // namespace

extern "C" int cc_lanczos_work_size(int64_t n, int max_iter, int64_t* out) {
  *out = (int64_t)(max_iter + 1) * n + 3 * n + 12 * (int64_t)(max_iter + 2) + 16;
  return CC_OK;
}

static int lanczos_driver(hipStream_t st, ApplyFn apply, const void* plan, int64_t n_env, int64_t d, const double* L,
                          const double* Lt, const double* W, const double* Wt, const double* x0, double tol_residual,
                          double tol_energy, int max_iter, int check_every, double* work, double* x_out,
                          double* host_result);

extern "C" int cc_lanczos_f64(void* stream, const void* plan, int64_t n_env, int64_t d, const double* L, const double* Lt,
                              const double* W, const double* Wt, const double* x0, double tol_residual,
                              double tol_energy, int max_iter, int check_every, double* work, double* x_out,
                              double* host_result /* theta, residual, iterations */) {
  if (!plan || n_env * d <= 0 || max_iter < 1 || check_every < 0) { cc_set_last_error("lanczos: bad arguments"); return CC_ERR_ARG; }
  if (cc_internal_stage3_n(plan) != n_env * d) { cc_set_last_error("lanczos: operator size does not match the whitening"); return CC_ERR_DIM; }
  return lanczos_driver(cc_stream(stream), cc_internal_stage3_apply, plan, n_env, d, L, Lt, W, Wt, x0, tol_residual,
                        tol_energy, max_iter, check_every, work, x_out, host_result);
}

extern "C" int cc_lanczos_lp(void* stream, const void* lp_plan, int64_t n_env, int64_t d, const double* L, const double* Lt,
                             const double* W, const double* Wt, const double* x0, double tol_residual, double tol_energy,
                             int max_iter, int check_every, double* work, double* x_out, double* host_result) {
  if (!lp_plan || n_env * d <= 0 || max_iter < 1 || check_every < 0) { cc_set_last_error("lanczos: bad arguments"); return CC_ERR_ARG; }
  if (cc_internal_lp_n(lp_plan) != n_env * d) { cc_set_last_error("lanczos: operator size does not match the whitening"); return CC_ERR_DIM; }
  return lanczos_driver(cc_stream(stream), cc_internal_lp_apply, lp_plan, n_env, d, L, Lt, W, Wt, x0, tol_residual,
                        tol_energy, max_iter, check_every, work, x_out, host_result);
}

static int lanczos_driver(hipStream_t st, ApplyFn apply, const void* plan, int64_t n_env, int64_t d, const double* L,
                          const double* Lt, const double* W, const double* Wt, const double* x0, double tol_residual,
                          double tol_energy, int max_iter, int check_every, double* work, double* x_out,
                          double* host_result) {
  const int64_t n = n_env * d;
  Whitening wh{n_env, d, L, Lt, W, Wt};
  double* Q = work;
  double* w = Q + (int64_t)(max_iter + 1) * n;
  double* t1 = w + n;
  double* t2 = t1 + n;
  double* alpha = t2 + n;
  double* beta = alpha + (max_iter + 2);
  double* c = beta + (max_iter + 2);
  double* s = c + (max_iter + 2);
  double* scratch = s + (max_iter + 2);  // 5 * (max_iter + 2)
  double* res = scratch + 5 * (max_iter + 2);
  int rc;

  // q0 = L^T x0 / ||L^T x0||
  rc = cc_skinny_f64(st, n_env, 2, Lt, (int)d, x0, w);
  if (rc) return rc;
  lz_normalize_kernel<<<1, 1024, 0, st>>>(n, w, Q, nullptr);
  CC_LAUNCH_CHECK();

  const bool fixed = check_every == 0;  // exactly max_iter steps, no host reads (capturable in a graph)
  set_nan_kernel<<<1, 64, 0, st>>>(res, 2);
  CC_LAUNCH_CHECK();
  double prev_theta = NAN, theta = NAN, resid = NAN;
  int m = 0, checked_m = -1;
  for (int j = 0; j < max_iter; ++j) {
    double* qj = Q + (int64_t)j * n;
    rc = op_whitened(st, apply, plan, wh, qj, t1, t2, w);
    if (rc) return rc;
    // full reorthogonalisation, two classical Gram-Schmidt passes against q_0..q_j; the first pass's
    // coefficient on q_j is alpha_j = q_j . A q_j (the three-term recurrence is contained in it)
    lz_coeff_kernel<<<j + 1, 256, 0, st>>>(n, Q, w, c, -1);
    CC_LAUNCH_CHECK();
    lz_combine_kernel<<<(unsigned)((n + 15) / 16), 256, 0, st>>>(j + 1, n, Q, c, w, 0, alpha + j);
    CC_LAUNCH_CHECK();
    // second pass with the normalisation folded in (w . w from one more coefficient block)
    lz_coeff_kernel<<<j + 2, 256, 0, st>>>(n, Q, w, c, j + 1);
    CC_LAUNCH_CHECK();
    lz_combine_normalize_kernel<<<(unsigned)((n + 15) / 16), 256, 0, st>>>(j + 1, n, Q, c, w, Q + (int64_t)(j + 1) * n,
                                                                           beta + j);
    CC_LAUNCH_CHECK();
    m = j + 1;
    if (m == n) break;
    if (!fixed && (m % check_every == 0 || m == max_iter)) {
      // Ritz value, Ritz vector and residual estimate in one pass (one read of two doubles); the
      // residual counts once the value has settled
      lz_tridiag_kernel<<<1, 64, 0, st>>>(m, alpha, beta, res, s, scratch, 1);
      CC_LAUNCH_CHECK();
      double h[2];
      CC_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st));
      CC_HIP(hipStreamSynchronize(st));
      theta = h[0];
      resid = h[1];
      checked_m = m;
      const bool settle_check = !std::isnan(prev_theta) && prev_theta - theta <= 1e3 * tol_energy * fabs(theta);
      bool done = m == n || m == max_iter ||
                  (settle_check && (resid <= tol_residual * fabs(theta) ||
                                    (prev_theta - theta <= tol_energy * fabs(theta))));
      prev_theta = theta;
      if (done) break;
    }
  }
  // the final Ritz vector of T_m (already in s when the last check was at this m)
  if (checked_m != m) {
    lz_tridiag_kernel<<<1, 64, 0, st>>>(m, alpha, beta, res, s, scratch, 1);
    CC_LAUNCH_CHECK();
    if (!fixed) {
      double h[2];
      CC_HIP(hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, st));
      CC_HIP(hipStreamSynchronize(st));
      theta = h[0];
      resid = h[1];
    }
  }
  // Ritz vector z = Q s, x = W^T z   (N-normalised because ||z|| = 1)
  lz_combine_kernel<<<(unsigned)((n + 15) / 16), 256, 0, st>>>(m, n, Q, s, t2, 1);
  CC_LAUNCH_CHECK();
  rc = cc_skinny_f64(st, n_env, 2, Wt, (int)d, t2, x_out);  // stream-ordered: no wait for it here
  if (rc) return rc;
  if (host_result) {
    host_result[0] = theta;   // NaN in the fixed mode: the Ritz value and residual stay on the device
    host_result[1] = resid;
    host_result[2] = (double)m;
  }
  return CC_OK;
}

// ---------------------------------------------------------------- Jacobi-Davidson pieces
namespace {
// out[j] = sum_i U[j][i] v[i] for j < k  (one block per basis vector, fixed-order tree)
__global__ void __launch_bounds__(256) jd_dots_kernel(int64_t n, const double* __restrict__ U, const double* __restrict__ v,
                                                      double* __restrict__ out, int stride_out) {
  __shared__ double red[256];
  const double t = block_dot256(n, U + (int64_t)blockIdx.x * n, v, red);
  if (threadIdx.x == 0) out[(int64_t)blockIdx.x * stride_out] = t;
}

// y = a*x + b*y with a, b read from device scalars (a = sa[0]*fa, b = sb ? sb[0]*fb : fb)
__global__ void cg_axpby_kernel(int64_t n, const double* __restrict__ sa, double fa, const double* __restrict__ x,
                                const double* __restrict__ sb, double fb, double* __restrict__ y) {
  const double a = (sa ? sa[0] : 1.0) * fa, b = (sb ? sb[0] : 1.0) * fb;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

// ratio[0] = num[0] / den[0] (0 when den == 0)

// lowest eigenpair of the k x k pencil (Hs, Ns), k <= 32, single thread: Cholesky of Ns, cyclic Jacobi
// on L^-1 Hs L^-T, back-transform.  out[0] = value, vec[0..k) N-normalised coefficients.
// Sum of v over the 1024 threads of the workgroup in a fixed order (wave shuffles, then the 16 wave
// sums in wave order); every thread gets the result.
__device__ inline double cg_block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < 16; ++k) t += red[k];
  return t;
}

// One CG step of the projected correction equation after ap = (W Hlp W^T) p, in one 1024-thread
// workgroup:  ap -= rho p;  ap -= u (u . ap);  alpha = rr / (p . ap);  y += alpha p;  res -= alpha ap;
// rr' = res . res;  p = res + (rr' / rr) p;  rr = rr'.   Fixed-order reductions.
__global__ void __launch_bounds__(1024) cg_step_kernel(int64_t n, const double* __restrict__ rho_dev, const double* __restrict__ u,
                                                       double* __restrict__ ap, double* __restrict__ p,
                                                       double* __restrict__ y, double* __restrict__ res,
                                                       double* __restrict__ rr) {
  __shared__ double red[16];
  const double rho = rho_dev[0];
  double du = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) {
    const double a = ap[i] - rho * p[i];
    ap[i] = a;
    du += u[i] * a;
  }
  const double c = cg_block_sum(du, red);
  double pap = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) {
    const double a = ap[i] - c * u[i];
    ap[i] = a;
    pap += p[i] * a;
  }
  const double pap_sum = cg_block_sum(pap, red);
  // a converged (zero) residual or search direction ends the iteration with y unchanged, not a NaN
  const double alpha = (pap_sum != 0.0 && rr[0] != 0.0) ? rr[0] / pap_sum : 0.0;
  double rn = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) {
    y[i] += alpha * p[i];
    const double r = res[i] - alpha * ap[i];
    res[i] = r;
    rn += r * r;
  }
  const double rr_new = cg_block_sum(rn, red);
  const double beta = rr[0] != 0.0 ? rr_new / rr[0] : 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) p[i] = res[i] + beta * p[i];
  __syncthreads();
  if (threadIdx.x == 0) rr[0] = rr_new;
}

// The same step with the vectors in registers (n <= 1024 R): one load round, three fixed-order sums,
// one store round (ap is scratch for the next step's operator product and is not written back).
template <int R>
__global__ void __launch_bounds__(1024) cg_step_reg_kernel(int64_t n, const double* __restrict__ rho_dev, const double* __restrict__ u,
                                                           const double* __restrict__ ap, double* __restrict__ p,
                                                           double* __restrict__ y, double* __restrict__ res,
                                                           double* __restrict__ rr) {
  __shared__ double red[16];
  const double rho = rho_dev[0], rr0 = rr[0];
  double a[R], pv[R], uv[R], yv[R], rv[R];
#pragma unroll
  for (int k = 0; k < R; ++k) {  // clamped index: every load in bounds, the values masked in the sums
    const int64_t i = threadIdx.x + 1024 * k, ic = i < n ? i : 0;
    a[k] = ap[ic];
    pv[k] = p[ic];
    uv[k] = u[ic];
    yv[k] = y[ic];
    rv[k] = res[ic];
  }
  double du = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    a[k] -= rho * pv[k];
    if (threadIdx.x + 1024 * k < n) du += uv[k] * a[k];
  }
  const double c = cg_block_sum(du, red);
  double pap = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    a[k] -= c * uv[k];
    if (threadIdx.x + 1024 * k < n) pap += pv[k] * a[k];
  }
  const double pap_sum = cg_block_sum(pap, red);
  // a converged (zero) residual or search direction ends the iteration with y unchanged, not a NaN
  const double alpha = (pap_sum != 0.0 && rr0 != 0.0) ? rr0 / pap_sum : 0.0;
  double rn = 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    yv[k] += alpha * pv[k];
    rv[k] -= alpha * a[k];
    if (threadIdx.x + 1024 * k < n) rn += rv[k] * rv[k];
  }
  const double rr_new = cg_block_sum(rn, red);
  const double beta = rr0 != 0.0 ? rr_new / rr0 : 0.0;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    const int64_t i = threadIdx.x + 1024 * k;
    if (i < n) {
      p[i] = rv[k] + beta * pv[k];
      y[i] = yv[k];
      res[i] = rv[k];
    }
  }
  if (threadIdx.x == 0) rr[0] = rr_new;
}

__global__ void small_geneig_kernel(int k, const double* __restrict__ Hs, const double* __restrict__ Ns,
                                    double* __restrict__ out, double* __restrict__ vec) {
  // the four k x k work matrices live in LDS (no private scratch: concurrent launches on several
  // streams would each need a large scratch allocation)
  __shared__ double Lm[32][32], C[32][32], V[32][32], Li[32][32];
  if (threadIdx.x != 0) return;
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) { Lm[i][j] = 0.0; Li[i][j] = 0.0; V[i][j] = (i == j) ? 1.0 : 0.0; }
  for (int j = 0; j < k; ++j) {
    double s = Ns[j * k + j];
    for (int c = 0; c < j; ++c) s -= Lm[j][c] * Lm[j][c];
    Lm[j][j] = sqrt(fmax(s, 1e-300));
    for (int i = j + 1; i < k; ++i) {
      double t = Ns[i * k + j];
      for (int c = 0; c < j; ++c) t -= Lm[i][c] * Lm[j][c];
      Lm[i][j] = t / Lm[j][j];
    }
  }
  for (int c = 0; c < k; ++c)
    for (int i = c; i < k; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int t = c; t < i; ++t) v -= Lm[i][t] * Li[t][c];
      Li[i][c] = v / Lm[i][i];
    }
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      double acc = 0.0;
      for (int a = 0; a <= i; ++a)
        for (int b = 0; b <= j; ++b) acc += Li[i][a] * Hs[a * k + b] * Li[j][b];
      C[i][j] = acc;
    }
  for (int sweep = 0; sweep < 50; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < k; ++i)
      for (int j = i + 1; j < k; ++j) off += C[i][j] * C[i][j];
    if (off < 1e-30) break;
    for (int p = 0; p < k; ++p)
      for (int q = p + 1; q < k; ++q) {
        if (C[p][q] == 0.0) continue;
        double th = (C[q][q] - C[p][p]) / (2.0 * C[p][q]);
        double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int r = 0; r < k; ++r) {
          double a = C[r][p], b = C[r][q];
          C[r][p] = c * a - s * b;
          C[r][q] = s * a + c * b;
        }
        for (int r = 0; r < k; ++r) {
          double a = C[p][r], b = C[q][r];
          C[p][r] = c * a - s * b;
          C[q][r] = s * a + c * b;
        }
        for (int r = 0; r < k; ++r) {
          double a = V[r][p], b = V[r][q];
          V[r][p] = c * a - s * b;
          V[r][q] = s * a + c * b;
        }
      }
  }
  int best = 0;
  for (int i = 1; i < k; ++i)
    if (C[i][i] < C[best][best]) best = i;
  out[0] = C[best][best];
  for (int i = 0; i < k; ++i) {  // coefficients = L^-T v
    double acc = 0.0;
    for (int j = i; j < k; ++j) acc += Li[j][i] * V[j][best];
    vec[i] = acc;
  }
}
}  // namespace

extern "C" int cc_basis_dots_f64(void* stream, int k, int64_t n, const double* U, const double* v, double* out, int stride_out) {
  if (k <= 0) return CC_OK;
  jd_dots_kernel<<<k, 256, 0, cc_stream(stream)>>>(n, U, v, out, stride_out);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_basis_combine_f64(void* stream, int k, int64_t n, const double* U, const double* c, double* out) {
  lz_combine_kernel<<<(unsigned)((n + 15) / 16), 256, 0, cc_stream(stream)>>>(k, n, U, c, out, 1);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_small_geneig_f64(void* stream, int k, const double* Hs, const double* Ns, double* out, double* vec) {
  if (k < 1 || k > 32) { cc_set_last_error("small_geneig: 1 <= k <= 32"); return CC_ERR_ARG; }
  small_geneig_kernel<<<1, 64, 0, cc_stream(stream)>>>(k, Hs, Ns, out, vec);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_jd_work_size(int64_t n, int64_t* out) {
  *out = 10 * n + 16;
  return CC_OK;
}

// Approximate solution of the projected correction equation
//   (I - u u^T)(W Hlp W^T - rho)(I - u u^T) y = -(I - u u^T) W r,  u = L^T x / |L^T x|,  t = W^T (I - u u^T) y
// by `iters` CG steps on the low-precision operator (all scalars on the device).
extern "C" int cc_jd_correction_lp(void* stream, const void* lp_plan, int64_t n_env, int64_t d, const double* Lt,
                                   const double* W, const double* Wt, const double* rho_dev, const double* x,
                                   const double* r, int iters, double* work, double* t_out) {
  hipStream_t st = cc_stream(stream);
  const int64_t n = n_env * d;
  if (cc_internal_lp_n(lp_plan) != n) { cc_set_last_error("jd: operator size mismatch"); return CC_ERR_DIM; }
  double* u = work;
  double* y = u + n;
  double* res = y + n;
  double* p = res + n;
  double* ap = p + n;
  double* t1 = ap + n;
  double* t2 = t1 + n;
  double* tmp = t2 + n;
  double* sc = tmp + 2 * n;  // scalars: 0 uu, 1 rr, 2 rr_new, 3 pAp, 4 alpha, 5 beta, 6 dot
  const unsigned gb = cc_blocks(n, 256);
  int rc;
  Whitening wh{n_env, d, nullptr, Lt, W, Wt};
  // u = normalized L^T x
  rc = cc_skinny_f64(st, n_env, 2, Lt, (int)d, x, tmp);
  if (rc) return rc;
  lz_normalize_kernel<<<1, 1024, 0, st>>>(n, tmp, u, nullptr);
  CC_LAUNCH_CHECK();
  auto project = [&](double* v) -> int {  // v -= u (u . v)
    rc = cc_dot_f64(st, n, u, v, sc + 6);
    if (rc) return rc;
    cg_axpby_kernel<<<gb, 256, 0, st>>>(n, sc + 6, -1.0, u, nullptr, 1.0, v);
    CC_LAUNCH_CHECK();
    return CC_OK;
  };
  // res = -P W r ; p = res ; y = 0
  rc = cc_skinny_f64(st, n_env, 1, W, (int)d, r, res);
  if (rc) return rc;
  cg_axpby_kernel<<<gb, 256, 0, st>>>(n, nullptr, -1.0, res, nullptr, 0.0, res);
  CC_LAUNCH_CHECK();
  if ((rc = project(res))) return rc;
  rc = cc_fill_f64(st, n, 0.0, y);
  if (rc) return rc;
  cg_axpby_kernel<<<gb, 256, 0, st>>>(n, nullptr, 1.0, res, nullptr, 0.0, p);
  CC_LAUNCH_CHECK();
  rc = cc_dot_f64(st, n, res, res, sc + 1);
  if (rc) return rc;
  for (int k = 0; k < iters; ++k) {
    // ap = P (W H W^T - rho) p   (p is kept in the projected space)
    rc = op_whitened(st, cc_internal_lp_apply, lp_plan, wh, p, t1, t2, ap);
    if (rc) return rc;
    if (n <= 4096) cg_step_reg_kernel<4><<<1, 1024, 0, st>>>(n, rho_dev, u, ap, p, y, res, sc + 1);
    else cg_step_kernel<<<1, 1024, 0, st>>>(n, rho_dev, u, ap, p, y, res, sc + 1);
    CC_LAUNCH_CHECK();
  }
  if ((rc = project(y))) return rc;
  return cc_skinny_f64(st, n_env, 2, Wt, (int)d, y, t_out);                // t = W^T P y
}

extern "C" int cc_normalize_f64(void* stream, int64_t n, const double* x, double* y) {
  if (n <= 0) return CC_OK;
  lz_normalize_kernel<<<1, 1024, 0, cc_stream(stream)>>>(n, x, y, nullptr);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
