// Thin SVD by one-sided Jacobi (Hestenes) with the round-robin parallel ordering:
// one 1024-thread workgroup per matrix, one wave per disjoint column pair per round.
// The columns of A and V live in LDS when (m + n) n <= 8192 doubles (64 KiB, every Gram matrix of a
// bond up to 64 and the small normalizations), otherwise in global memory (`work`), so m is unbounded;
// n <= 256.  Both placements do the same arithmetic in the same order.
#include "cc_common.h"

namespace {
constexpr int SVD_THREADS = 1024;
constexpr int SVD_MAX_N = 256;
constexpr double SVD_ZERO = 0x1p-960;  // singular values below this count as zero (W / s loses precision)

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int SVD_LDS_DOUBLES = 8192;

// Player at position e (0 <= e < np) of round r of the round-robin tournament: position 0 keeps player 0,
// the others rotate by one position per round (after np - 1 rounds every pair has met once).
__device__ inline int tournament_player(int e, int r, int np) {
  if (e == 0) return 0;
  const int k = np - 1;
  return 1 + ((e - 1 - r) % k + k) % k;
}

template <bool IN_LDS>
__global__ void __launch_bounds__(SVD_THREADS) svd_jacobi_kernel(int64_t m, int n, const double* __restrict__ A,
                                                                 double* __restrict__ U, double* __restrict__ S,
                                                                 double* __restrict__ Vt, double* __restrict__ work) {
  const int64_t b = blockIdx.x;
  A += b * m * n;
  U += b * m * n;
  S += b * n;
  Vt += b * n * n;
  __shared__ double cols[IN_LDS ? SVD_LDS_DOUBLES : 1];
  double* W = IN_LDS ? cols : work + b * (m * n + (int64_t)n * n);  // W[j][i] = column j of A (n x m)
  double* V = W + m * n;                                             // V[j][k] = column j of V (n x n)
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = SVD_THREADS / 64;
  __shared__ int rotated;
  __shared__ double snorm[SVD_MAX_N];
  __shared__ int perm[SVD_MAX_N];

  for (int64_t e = tid; e < m * n; e += SVD_THREADS) {
    int64_t i = e / n;
    int j = (int)(e % n);
    W[(int64_t)j * m + i] = A[e];
  }
  for (int e = tid; e < n * n; e += SVD_THREADS) V[e] = (e / n == e % n) ? 1.0 : 0.0;
  const int np = (n + 1) & ~1;  // even count; index n (when odd) is a dummy
  __syncthreads();

  for (int sweep = 0; sweep < 60; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      for (int pr = wave; pr < np / 2; pr += nwaves) {
        int p = tournament_player(pr, round, np), q = tournament_player(np - 1 - pr, round, np);
        if (p >= n || q >= n) continue;
        if (p > q) { int t = p; p = q; q = t; }
        double* wp = W + (int64_t)p * m;
        double* wq = W + (int64_t)q * m;
        double al = 0.0, be = 0.0, ga = 0.0;
        for (int64_t i = lane; i < m; i += 64) {
          double x = wp[i], y = wq[i];
          al += x * x;
          be += y * y;
          ga += x * y;
        }
        al = wave_sum(al);
        be = wave_sum(be);
        ga = wave_sum(ga);
        if (ga != 0.0 && fabs(ga) > 1e-15 * sqrt(al * be)) {
          double zeta = (be - al) / (2.0 * ga);
          double t = (zeta >= 0.0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1.0 + zeta * zeta));
          double c = 1.0 / sqrt(1.0 + t * t), s = c * t;
          for (int64_t i = lane; i < m; i += 64) {
            double x = wp[i], y = wq[i];
            wp[i] = c * x - s * y;
            wq[i] = s * x + c * y;
          }
          double* vp = V + (int64_t)p * n;
          double* vq = V + (int64_t)q * n;
          for (int k = lane; k < n; k += 64) {
            double x = vp[k], y = vq[k];
            vp[k] = c * x - s * y;
            vq[k] = s * x + c * y;
          }
          if (lane == 0) rotated = 1;
        }
      }
      __syncthreads();
    }
    if (!rotated) break;
  }

  // singular values = column norms, sorted descending (stable by index)
  for (int j = wave; j < n; j += nwaves) {
    double s2 = 0.0;
    for (int64_t i = lane; i < m; i += 64) { double x = W[(int64_t)j * m + i]; s2 += x * x; }
    s2 = wave_sum(s2);
    if (lane == 0) snorm[j] = sqrt(s2);
  }
  __syncthreads();
  for (int j = tid; j < n; j += SVD_THREADS) {
    int r = 0;
    for (int k = 0; k < n; ++k) r += (snorm[k] > snorm[j]) || (snorm[k] == snorm[j] && k < j);
    perm[r] = j;
  }
  __syncthreads();
  for (int r = tid; r < n; r += SVD_THREADS) S[r] = snorm[perm[r]];
  for (int64_t e = tid; e < m * n; e += SVD_THREADS) {
    int64_t i = e / n;
    int r = (int)(e % n);
    int j = perm[r];
    double s = snorm[j];
    U[e] = s >= SVD_ZERO ? W[(int64_t)j * m + i] / s : 0.0;
  }
  for (int e = tid; e < n * n; e += SVD_THREADS) {
    int r = e / n, k = e % n;
    Vt[e] = V[(int64_t)perm[r] * n + k];
  }

  // Columns of zero singular value carry no direction of A; give them an orthonormal completion so U
  // always has orthonormal columns (LAPACK's thin U does, and polar factors U Vt must be isometries).
  // Column r takes the unit vector e_k least covered by the accepted columns (smallest row norm, lowest
  // k on ties, so |P_perp e_k|^2 >= (m - r) / m), orthogonalised twice and normalised.
  __shared__ int first_zero;
  __shared__ double coef[SVD_MAX_N];
  __shared__ double red_v[SVD_THREADS / 64];
  __shared__ int64_t red_i[SVD_THREADS / 64];
  if (tid == 0) {
    int f = n;
    while (f > 0 && !(snorm[perm[f - 1]] >= SVD_ZERO)) --f;
    first_zero = f;
  }
  __syncthreads();
  if (first_zero == n) return;
  double* rn = W;  // m row norms of U's accepted columns (W is no longer needed)
  for (int64_t i = tid; i < m; i += SVD_THREADS) {
    double a = 0.0;
    for (int c = 0; c < first_zero; ++c) { double u = U[i * n + c]; a += u * u; }
    rn[i] = a;
  }
  __syncthreads();
  for (int r = first_zero; r < n; ++r) {
    // k = argmin over rows of rn (lowest index on ties)
    double bv = INFINITY;
    int64_t bi = m;
    for (int64_t i = tid; i < m; i += SVD_THREADS)
      if (rn[i] < bv) { bv = rn[i]; bi = i; }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      double ov = __shfl_xor(bv, o, 64);
      int64_t oi = __shfl_xor(bi, o, 64);
      if (ov < bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { red_v[wave] = bv; red_i[wave] = bi; }
    __syncthreads();
    if (tid == 0) {
      for (int w = 1; w < nwaves; ++w)
        if (red_v[w] < red_v[0] || (red_v[w] == red_v[0] && red_i[w] < red_i[0])) { red_v[0] = red_v[w]; red_i[0] = red_i[w]; }
    }
    __syncthreads();
    const int64_t k = red_i[0] < m ? red_i[0] : 0;  // NaN row norms leave no candidate
    for (int c = tid; c < r; c += SVD_THREADS) coef[c] = U[k * n + c];
    __syncthreads();
    for (int64_t i = tid; i < m; i += SVD_THREADS) {
      double v = (i == k) ? 1.0 : 0.0;
      for (int c = 0; c < r; ++c) v -= U[i * n + c] * coef[c];
      U[i * n + r] = v;
    }
    __syncthreads();
    // second Gram-Schmidt pass: coef[c] = <U_c, v>, one wave per column
    for (int c = wave; c < r; c += nwaves) {
      double d = 0.0;
      for (int64_t i = lane; i < m; i += 64) d += U[i * n + c] * U[i * n + r];
      d = wave_sum(d);
      if (lane == 0) coef[c] = d;
    }
    __syncthreads();
    double nn = 0.0;
    for (int64_t i = tid; i < m; i += SVD_THREADS) {
      double v = U[i * n + r];
      for (int c = 0; c < r; ++c) v -= U[i * n + c] * coef[c];
      U[i * n + r] = v;
      nn += v * v;
    }
    nn = wave_sum(nn);
    if (lane == 0) red_v[wave] = nn;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int w = 0; w < nwaves; ++w) t += red_v[w];
      red_v[0] = 1.0 / sqrt(t);
    }
    __syncthreads();
    const double sc = red_v[0];
    for (int64_t i = tid; i < m; i += SVD_THREADS) {
      double v = U[i * n + r] * sc;
      U[i * n + r] = v;
      rn[i] += v * v;
    }
    __syncthreads();
  }
}
// Eigen-decomposition of a symmetric n x n matrix (n <= 64) by the parallel cyclic Jacobi method, matrix and
// eigenvectors in LDS: each round of the round-robin ordering rotates n/2 disjoint (p, q) pairs, G -> J^T G J
// and V -> V J, in one phase over the 2 x 2 blocks that pairs of pairs cut out of G.  A pair is rotated while
// |g_pq| > 1e-15 sqrt(|g_pp g_qq|) and |g_pq| > 1e-15 tr|G| (off-diagonal entries below the second bound move
// the eigenvalues by less than the rounding of G itself, and rotating them would not terminate for
// rank-deficient G).  Outputs: lam descending (stable by index) and Vt[r][i] = component i of eigenvector r.
constexpr int SYM_THREADS = 1024;
constexpr int SYM_MAX_N = 64;
constexpr int SYM_TASKS = SYM_MAX_N * (SYM_MAX_N / 2) / SYM_THREADS;  // (entry pair) tasks per thread and phase
static_assert(SYM_TASKS * SYM_THREADS == SYM_MAX_N * (SYM_MAX_N / 2), "tasks must cover n * n/2 exactly");

template <int NC>  // NC = n known at compile time (index arithmetic by shifts and constant divisors), 0 = any n
__global__ void __launch_bounds__(SYM_THREADS) sym_jacobi_kernel(int n_rt, const double* __restrict__ A,
                                                                 double* __restrict__ lam, double* __restrict__ Vt) {
  const int n = NC ? NC : n_rt;
  __shared__ double G[SYM_MAX_N * SYM_MAX_N];
  __shared__ double V[SYM_MAX_N * SYM_MAX_N];
  __shared__ double rot_c[SYM_MAX_N / 2], rot_s[SYM_MAX_N / 2];
  __shared__ int rot_p[SYM_MAX_N / 2], rot_q[SYM_MAX_N / 2], rot_on[SYM_MAX_N / 2];
  __shared__ int perm[SYM_MAX_N];
  __shared__ double trace_abs;
  __shared__ int rotated;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += SYM_THREADS) {
    G[e] = A[e];
    V[e] = (e / n == e % n) ? 1.0 : 0.0;
  }
  __syncthreads();
  if (tid == 0) {
    double t = 0.0;
    for (int i = 0; i < n; ++i) t += fabs(G[i * n + i]);
    trace_abs = t;
  }
  __syncthreads();
  const double floor_abs = 1e-15 * trace_abs;
  const int np = (n + 1) & ~1, half = np / 2;
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (tid == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      if (tid < half) {
        int p = tournament_player(tid, round, np), q = tournament_player(np - 1 - tid, round, np);
        if (p > q) { const int t = p; p = q; q = t; }
        int on = 0;
        if (q < n) {
          const double a = G[p * n + p], b = G[q * n + q], g = G[p * n + q];
          if (fabs(g) > 1e-15 * sqrt(fabs(a * b)) && fabs(g) > floor_abs) {
            const double theta = (b - a) / (2.0 * g);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0);
            rot_c[tid] = c;
            rot_s[tid] = t * c;
            on = 1;
            rotated = 1;
          }
        }
        rot_p[tid] = p;
        rot_q[tid] = q;
        rot_on[tid] = on;
      }
      __syncthreads();
      // One phase: thread t < half^2 takes the 2 x 2 block of G on the rows of pair t / half and the columns
      // of pair t % half and forms J1^T (B J2) in that order (the column rotation, then the row rotation,
      // each entry through the same operations as a columns-then-rows sweep of the whole matrix); the V
      // columns of every pair (V J) are spread over all threads.  Blocks and V entries are disjoint, so
      // one barrier ends the round.  A pair that does not rotate acts as the identity.
      if (tid < half * half) {
        const int k1 = tid / half, k2 = tid - (tid / half) * half;
        const int a0 = rot_p[k1], a1 = rot_q[k1], b0 = rot_p[k2], b1 = rot_q[k2];
        const bool on1 = rot_on[k1], on2 = rot_on[k2];
        if (on1 || on2) {
          const double c1 = on1 ? rot_c[k1] : 1.0, s1 = on1 ? rot_s[k1] : 0.0;
          const double c2 = on2 ? rot_c[k2] : 1.0, s2 = on2 ? rot_s[k2] : 0.0;
          const bool r0 = a0 < n, r1 = a1 < n, q0 = b0 < n, q1 = b1 < n;   // indices past n are padding
          const double x00 = (r0 && q0) ? G[a0 * n + b0] : 0.0, x01 = (r0 && q1) ? G[a0 * n + b1] : 0.0;
          const double x10 = (r1 && q0) ? G[a1 * n + b0] : 0.0, x11 = (r1 && q1) ? G[a1 * n + b1] : 0.0;
          // columns (B J2)
          const double y00 = c2 * x00 - s2 * x01, y01 = s2 * x00 + c2 * x01;
          const double y10 = c2 * x10 - s2 * x11, y11 = s2 * x10 + c2 * x11;
          // rows (J1^T (B J2))
          if (r0 && q0) G[a0 * n + b0] = c1 * y00 - s1 * y10;
          if (r0 && q1) G[a0 * n + b1] = c1 * y01 - s1 * y11;
          if (r1 && q0) G[a1 * n + b0] = s1 * y00 + c1 * y10;
          if (r1 && q1) G[a1 * n + b1] = s1 * y01 + c1 * y11;
        }
      }
      {  // columns p, q of V
#pragma unroll
        for (int it = 0; it < SYM_TASKS; ++it) {
          const int e = tid + it * SYM_THREADS, k = e % half, i = e / half;
          if (e < n * half && rot_on[k]) {
            const int ip = i * n + rot_p[k], iq = i * n + rot_q[k];
            const double c = rot_c[k], sn = rot_s[k], vx = V[ip], vy = V[iq];
            V[ip] = c * vx - sn * vy;
            V[iq] = sn * vx + c * vy;
          }
        }
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  for (int j = tid; j < n; j += SYM_THREADS) {
    const double dj = G[j * n + j];
    int r = 0;
    for (int k = 0; k < n; ++k) {
      const double dk = G[k * n + k];
      r += (dk > dj) || (dk == dj && k < j);
    }
    perm[r] = j;
  }
  __syncthreads();
  for (int r = tid; r < n; r += SYM_THREADS) lam[r] = G[perm[r] * n + perm[r]];
  for (int e = tid; e < n * n; e += SYM_THREADS) {
    const int r = e / n, i = e % n;
    Vt[e] = V[i * n + perm[r]];
  }
}

void launch_svd(hipStream_t st, unsigned batch, int64_t m, int n, const double* A, double* U, double* S, double* Vt,
                double* work) {
  if ((m + n) * n <= SVD_LDS_DOUBLES)
    svd_jacobi_kernel<true><<<batch, SVD_THREADS, 0, st>>>(m, n, A, U, S, Vt, work);
  else
    svd_jacobi_kernel<false><<<batch, SVD_THREADS, 0, st>>>(m, n, A, U, S, Vt, work);
}
}  // namespace

extern "C" int cc_svd_work_size(int64_t batch, int64_t m, int64_t n, int64_t* out) {
  *out = batch * (m * n + n * n);
  return CC_OK;
}

extern "C" int cc_svd_f64(void* stream, int64_t batch, int64_t m, int64_t n, const double* A, double* U, double* S,
                          double* Vt, double* work) {
  if (batch < 0 || m < 0 || n < 0) { cc_set_last_error("cc_svd_f64: negative size"); return CC_ERR_ARG; }
  if (n > SVD_MAX_N) { cc_set_last_error("cc_svd_f64: more than 256 columns"); return CC_ERR_UNSUPPORTED; }
  if (m < n) { cc_set_last_error("cc_svd_f64: needs m >= n"); return CC_ERR_ARG; }
  if (batch == 0 || n == 0) return CC_OK;
  launch_svd(cc_stream(stream), (unsigned)batch, m, (int)n, A, U, S, Vt, work);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// ---- truncation basis of a fused environment bond -----------------------------------------------------
namespace {
// Cs[(x, a, b)][j] = (C[(x, a, b)][j] + C[(x, b, a)][j]) / 2: the part of the bond matricization that is
// symmetric under the ket <-> bra exchange of the absorbed center legs (a, b).
__global__ void fused_symmetrize_kernel(int64_t nx, int64_t na, int64_t far, const double* __restrict__ C,
                                        double* __restrict__ Cs) {
  const int64_t total = nx * na * na * far;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = i % far, r = i / far, b = r % na, a = (r / na) % na, x = r / (na * na);
    Cs[i] = 0.5 * (C[i] + C[((x * na + b) * na + a) * far + j]);
  }
}
// One workgroup per column j of U (m x n row-major): the column is negated when its entry of largest
// magnitude (the first such row) is negative, so that SVD bases computed by different routines agree.
__global__ void sign_normalize_columns_kernel(int64_t m, int64_t n, double* __restrict__ U) {
  __shared__ double best_v[256];
  __shared__ int64_t best_i[256];
  const int64_t j = blockIdx.x;
  double bv = -1.0;
  int64_t bi = m;
  for (int64_t i = threadIdx.x; i < m; i += blockDim.x) {
    const double a = fabs(U[i * n + j]);
    if (a > bv) { bv = a; bi = i; }
  }
  best_v[threadIdx.x] = bv;
  best_i[threadIdx.x] = bi;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double v2 = best_v[threadIdx.x + o];
      const int64_t i2 = best_i[threadIdx.x + o];
      if (v2 > best_v[threadIdx.x] || (v2 == best_v[threadIdx.x] && i2 < best_i[threadIdx.x])) {
        best_v[threadIdx.x] = v2;
        best_i[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  const int64_t ib = best_i[0];
  if (ib < m && U[ib * n + j] < 0.0)
    for (int64_t i = threadIdx.x; i < m; i += blockDim.x) U[i * n + j] = -U[i * n + j];
}
}  // namespace

extern "C" int cc_sign_normalize_columns_f64(void* stream, int64_t m, int64_t n, double* U) {
  if (m < 0 || n < 0) { cc_set_last_error("sign normalize: negative size"); return CC_ERR_ARG; }
  if (m == 0 || n == 0) return CC_OK;
  sign_normalize_columns_kernel<<<(unsigned)n, 256, 0, cc_stream(stream)>>>(m, n, U);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

namespace {
// W[i][j] = V[i][j] / s_j with V = Vt^T and s_j = sqrt(lambda_j) the singular values of the matrix whose Gram
// eigenvalues lambda (descending) and vectors Vt came from the SVD of the Gram matrix; columns whose s_j is at
// most 1e-7 s_0 (below what the Gram matrix resolves) are zero.  S_out receives every s_j.
__global__ void fused_gram_scale_kernel(int64_t far, int64_t k, const double* __restrict__ lam, const double* __restrict__ Vt,
                                        double* __restrict__ W, double* __restrict__ S_out) {
  const double s0 = sqrt(fmax(lam[0], 0.0));
  for (int64_t e = threadIdx.x; e < far * k; e += blockDim.x) {
    const int64_t i = e / k, j = e % k;
    const double sj = sqrt(fmax(lam[j], 0.0));
    W[e] = (sj > 1e-7 * s0 && sj > 0.0) ? Vt[j * far + i] / sj : 0.0;
  }
  for (int64_t j = threadIdx.x; j < far; j += blockDim.x) S_out[j] = sqrt(fmax(lam[j], 0.0));
}
}  // namespace

// The Gram product G = Cs^T Cs has a far x far output and a sum over m: for m >= 2048 (and far <= 128) it
// is cut into fused_gram_slices(m) slices summed by one batched product, and the slices are added in
// order by a second product with a vector of ones (a single product would occupy a handful of
// workgroups for the whole sum).
inline int64_t fused_gram_slices(int64_t m, int64_t far) { return (m >= 2048 && far <= 128) ? (m / 64 < 64 ? m / 64 : 64) : 1; }

extern "C" int cc_fused_truncation_work_size(int64_t nx, int64_t na, int64_t far, int64_t* out) {
  const int64_t m = nx * na * na;
  int64_t svd = 0;
  cc_svd_work_size(1, far, far, &svd);
  const int64_t S = fused_gram_slices(m, far);
  *out = svd + m * far + 4 * far * far + far + (S > 1 ? (S + 1) * far * far + (S + 1) : 0);
  return CC_OK;
}

// The leading left singular vectors of the symmetrized C (m x far, m >= far) from its Gram matrix:
// G = Cs^T Cs = V diag(lambda) V^T (two-sided Jacobi on the far x far G for far <= 64, else one-sided Jacobi),
// U = Cs V diag(1/sqrt(lambda)).
extern "C" int cc_fused_truncation_f64(void* stream, int64_t nx, int64_t na, int64_t far, int64_t new_dim,
                                       const double* C, double* U_out, double* S_out, double* work) {
  const int64_t m = nx * na * na;
  if (nx < 1 || na < 1 || far < 1 || new_dim < 1) { cc_set_last_error("fused truncation: bad sizes"); return CC_ERR_ARG; }
  if (far > SVD_MAX_N) { cc_set_last_error("fused truncation: far bond above 256"); return CC_ERR_UNSUPPORTED; }
  if (m < far) { cc_set_last_error("fused truncation: the fused bond must be at least the far bond"); return CC_ERR_ARG; }
  if (new_dim > far) { cc_set_last_error("fused truncation: new dimension above the rank bound (the far bond)"); return CC_ERR_ARG; }
  hipStream_t st = cc_stream(stream);
  int64_t svd_work = 0;
  cc_svd_work_size(1, far, far, &svd_work);
  double* Cs = work + svd_work;
  double* G = Cs + m * far;
  double* Ug = G + far * far;
  double* Vt = Ug + far * far;
  double* W = Vt + far * far;
  double* lam = W + far * far;
  const unsigned blocks = (unsigned)((m * far + 255) / 256 < 65535 ? (m * far + 255) / 256 : 65535);
  fused_symmetrize_kernel<<<blocks, 256, 0, st>>>(nx, na, far, C, Cs);
  CC_LAUNCH_CHECK();
  int rc;
  const int64_t S = fused_gram_slices(m, far);
  if (S > 1) {
    double* parts = lam + far;                       // (S + 1) slices of far x far
    double* ones = parts + (S + 1) * far * far;      // S + 1 ones
    const int64_t kc = m / S, rest = m - S * kc, np = S + (rest ? 1 : 0);
    rc = cc_gemm_f64(stream, S, far, far, kc, 1.0, Cs, kc * far, 1, far, Cs, kc * far, far, 1, 0.0, parts, far * far, far, 1);
    if (rc) return rc;
    if (rest) {
      rc = cc_gemm_f64(stream, 1, far, far, rest, 1.0, Cs + S * kc * far, 0, 1, far, Cs + S * kc * far, 0, far, 1, 0.0,
                       parts + S * far * far, 0, far, 1);
      if (rc) return rc;
    }
    rc = cc_fill_f64(stream, np, 1.0, ones);
    if (rc) return rc;
    rc = cc_gemm_f64(stream, 1, 1, far * far, np, 1.0, ones, 0, np, 1, parts, 0, far * far, 1, 0.0, G, 0, far * far, 1);
  } else {
    rc = cc_gemm_f64(stream, 1, far, far, m, 1.0, Cs, 0, 1, far, Cs, 0, far, 1, 0.0, G, 0, far, 1);
  }
  if (rc) return rc;
  if (far == SYM_MAX_N)
    sym_jacobi_kernel<SYM_MAX_N><<<1, SYM_THREADS, 0, st>>>((int)far, G, lam, Vt);
  else if (far < SYM_MAX_N)
    sym_jacobi_kernel<0><<<1, SYM_THREADS, 0, st>>>((int)far, G, lam, Vt);
  else
    launch_svd(st, 1, far, (int)far, G, Ug, lam, Vt, work);
  CC_LAUNCH_CHECK();
  fused_gram_scale_kernel<<<1, 256, 0, st>>>(far, new_dim, lam, Vt, W, S_out);
  CC_LAUNCH_CHECK();
  rc = cc_gemm_f64(stream, 1, m, new_dim, far, 1.0, Cs, 0, far, 1, W, 0, new_dim, 1, 0.0, U_out, 0, new_dim, 1);
  if (rc) return rc;
  return cc_sign_normalize_columns_f64(stream, m, new_dim, U_out);
}
