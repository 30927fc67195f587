// Dense FP64 linear algebra for the solve: blocked Cholesky, lower-triangular inverse,
// skinny products with the whitening factor, Lanczos reorthogonalisation, and the
// lowest eigenpair of a symmetric tridiagonal matrix.
#include "cc_common.h"

namespace {
constexpr int CB = 64;  // Cholesky block size

// Unblocked Cholesky of the kb x kb diagonal block at (k0, k0) in LDS; writes L (lower) back.
__global__ void __launch_bounds__(256) chol_diag_kernel(int64_t n, int64_t k0, int kb, double* __restrict__ A,
                                                        int* __restrict__ info) {
  __shared__ double a[CB][CB + 1];
  for (int e = threadIdx.x; e < kb * kb; e += 256) {
    int i = e / kb, j = e % kb;
    a[i][j] = (j <= i) ? A[(k0 + i) * n + k0 + j] : 0.0;
  }
  __syncthreads();
  for (int j = 0; j < kb; ++j) {
    __syncthreads();
    double d = a[j][j];
    if (!(d > 0.0)) {
      if (threadIdx.x == 0 && *info == 0) *info = (int)(k0 + j + 1);
      d = 1e-300;
    }
    double s = sqrt(d);
    __syncthreads();
    if (threadIdx.x == 0) a[j][j] = s;
    for (int i = j + 1 + threadIdx.x; i < kb; i += 256) a[i][j] /= s;
    __syncthreads();
    for (int e = threadIdx.x; e < (kb - j - 1) * (kb - j - 1); e += 256) {
      int i = j + 1 + e / (kb - j - 1), c = j + 1 + e % (kb - j - 1);
      if (c <= i) a[i][c] -= a[i][j] * a[c][j];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kb * kb; e += 256) {
    int i = e / kb, j = e % kb;
    A[(k0 + i) * n + k0 + j] = (j <= i) ? a[i][j] : 0.0;
  }
}

// Panel: rows i >= k0+kb, columns [k0, k0+kb): solve X L_kk^T = A_panel, one thread per row.
__global__ void __launch_bounds__(256) chol_panel_kernel(int64_t n, int64_t k0, int kb, double* __restrict__ A) {
  __shared__ double l[CB][CB + 1];
  for (int e = threadIdx.x; e < kb * kb; e += 256) {
    int i = e / kb, j = e % kb;
    l[i][j] = A[(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  int64_t row = k0 + kb + (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (row >= n) return;
  double x[CB];
  for (int j = 0; j < kb; ++j) {
    double v = A[row * n + k0 + j];
    for (int c = 0; c < j; ++c) v -= x[c] * l[j][c];
    x[j] = v / l[j][j];
  }
  for (int j = 0; j < kb; ++j) A[row * n + k0 + j] = x[j];
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// Zero the strict upper triangle.
__global__ void zero_upper_kernel(int64_t n, double* __restrict__ A) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x)
    if (e % n > e / n) A[e] = 0.0;
}

// X = L^-1 by forward substitution on the identity; workgroup g owns columns [32g, 32g+32).
// 256 threads = 32 columns x 8 partial sums per column (fixed-order reduction).
__global__ void __launch_bounds__(256) trtri_lower_kernel(int64_t n, const double* __restrict__ L, double* __restrict__ X) {
  const int cl = threadIdx.x & 31, part = threadIdx.x >> 5;
  const int64_t c = (int64_t)blockIdx.x * 32 + cl;
  const int64_t c0 = (int64_t)blockIdx.x * 32;
  __shared__ double red[8][32];
  // zero the rows above the block and the block's own upper part
  for (int64_t i = 0; i < n; ++i) {
    double acc = 0.0;
    if (c < n && i >= c0) {
      for (int64_t k = c0 + part; k < i; k += 8) acc += L[i * n + k] * X[k * n + c];
    }
    red[part][cl] = acc;
    __syncthreads();
    if (part == 0 && c < n) {
      double s = 0.0;
      for (int p = 0; p < 8; ++p) s += red[p][cl];
      double rhs = (i == c) ? 1.0 : 0.0;
      X[i * n + c] = (i < c) ? 0.0 : (rhs - s) / L[i * n + i];
    }
    __syncthreads();
  }
}

// Y[i][s] = sum_k M[i][k] X[k][s] for k in [kbeg(i), kend(i)) with M row-major n x n,
// tri: 0 full, 1 lower (k <= i), 2 upper (k >= i).  One wave per output row.
template <int NR>
__global__ void __launch_bounds__(256) skinny_kernel(int64_t n, int tri, const double* __restrict__ M,
                                                     const double* __restrict__ X, double* __restrict__ Y) {
  const int lane = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= n) return;
  int64_t kb = (tri == 2) ? row : 0, ke = (tri == 1) ? row + 1 : n;
  double acc[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) acc[s] = 0.0;
  const double* mr = M + row * n;
  for (int64_t k = kb + lane; k < ke; k += 64) {
    double m = mr[k];
#pragma unroll
    for (int s = 0; s < NR; ++s) acc[s] += m * X[k * NR + s];
  }
#pragma unroll
  for (int s = 0; s < NR; ++s) {
    double v = acc[s];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    acc[s] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int s = 0; s < NR; ++s) Y[row * NR + s] = acc[s];
  }
}

__global__ void transpose_kernel(int64_t n, const double* __restrict__ A, double* __restrict__ B) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = e / n, j = e % n;
    B[j * n + i] = A[e];
  }
}
}  // namespace

extern "C" int cc_cholesky_f64(void* stream, int64_t n, double* A, int* info_dev) {
  if (n <= 0) { cc_set_last_error("cholesky: n must be positive"); return CC_ERR_ARG; }
  hipStream_t st = cc_stream(stream);
  set_int_kernel<<<1, 1, 0, st>>>(info_dev, 0);
  CC_LAUNCH_CHECK();
  for (int64_t k0 = 0; k0 < n; k0 += CB) {
    int kb = (int)((n - k0) < CB ? (n - k0) : CB);
    chol_diag_kernel<<<1, 256, 0, st>>>(n, k0, kb, A, info_dev);
    CC_LAUNCH_CHECK();
    int64_t rest = n - k0 - kb;
    if (rest > 0) {
      chol_panel_kernel<<<(unsigned)((rest + 255) / 256), 256, 0, st>>>(n, k0, kb, A);
      CC_LAUNCH_CHECK();
      const double* Lp = A + (k0 + kb) * n + k0;
      double* T = A + (k0 + kb) * n + (k0 + kb);
      int rc = cc_gemm_f64(stream, 1, rest, rest, kb, -1.0, Lp, 0, n, 1, Lp, 0, 1, n, 1.0, T, 0, n, 1);
      if (rc) return rc;
    }
  }
  zero_upper_kernel<<<cc_blocks(n * n, 256), 256, 0, st>>>(n, A);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_trtri_lower_f64(void* stream, int64_t n, const double* L, double* X) {
  if (n <= 0) { cc_set_last_error("trtri: n must be positive"); return CC_ERR_ARG; }
  trtri_lower_kernel<<<(unsigned)((n + 31) / 32), 256, 0, cc_stream(stream)>>>(n, L, X);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_transpose_f64(void* stream, int64_t n, const double* A, double* B) {
  if (n <= 0) return CC_OK;
  transpose_kernel<<<cc_blocks(n * n, 256), 256, 0, cc_stream(stream)>>>(n, A, B);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_skinny_f64(void* stream, int64_t n, int tri, const double* M, int nrhs, const double* X, double* Y) {
  if (n <= 0) return CC_OK;
  unsigned g = (unsigned)((n + 3) / 4);
  hipStream_t st = cc_stream(stream);
  switch (nrhs) {
    case 1: skinny_kernel<1><<<g, 256, 0, st>>>(n, tri, M, X, Y); break;
    case 2: skinny_kernel<2><<<g, 256, 0, st>>>(n, tri, M, X, Y); break;
    case 3: skinny_kernel<3><<<g, 256, 0, st>>>(n, tri, M, X, Y); break;
    case 4: skinny_kernel<4><<<g, 256, 0, st>>>(n, tri, M, X, Y); break;
    default: cc_set_last_error("skinny: nrhs must be 1..4"); return CC_ERR_UNSUPPORTED;
  }
  CC_LAUNCH_CHECK();
  return CC_OK;
}
