// Dense FP64 linear algebra for the solve: blocked Cholesky, lower-triangular inverse,
// skinny products with the whitening factor, Lanczos reorthogonalisation, and the
// lowest eigenpair of a symmetric tridiagonal matrix.
#include "cc_common.h"

namespace {
constexpr int CB = 64;  // Cholesky block size

// Cholesky of the kb x kb diagonal block at (k0, k0) in LDS on one 256-thread workgroup (kb <= CB; the
// block is padded to CB with the identity).  Step k updates the trailing lower triangle with the
// unscaled column k and the pivot (a[i][j] -= a[i][k] a[j][k] / a[k][k]), one barrier per step; the
// columns are scaled at the end.  The factor is then inverted in place column by column from the last
// (lower-triangular TRTI2 order: column j of L^-1 below the diagonal is -L^-1[j+1:, j+1:] L[j+1:, j] /
// L[j][j]), four threads per row sharing each dot product.  Writes L back to A and L^-1 to dinv
// (kb x kb row-major); info receives the first non-positive pivot (1-based, global index).
__global__ void __launch_bounds__(256) chol_diag_kernel(int64_t n, int64_t k0, int kb, double* __restrict__ A,
                                                       double* __restrict__ dinv, int* __restrict__ info) {
  __shared__ double a[CB][CB + 1];
  __shared__ double piv[CB];
  __shared__ double col[CB];
  const int t = threadIdx.x;
  for (int e = t; e < CB * CB; e += 256) {
    const int i = e / CB, j = e % CB;
    a[i][j] = (i < kb && j <= i) ? A[(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  int bad = 0;
  const int row = t >> 2, part = t & 3;  // four threads per row of the trailing triangle
  for (int k = 0; k < CB; ++k) {
    double d = a[k][k];
    if (!(d > 0.0)) {
      if (k < kb && !bad) bad = k + 1;
      d = 1e-300;
    }
    if (t == 0) piv[k] = d;
    if (row > k) {
      const double l = a[row][k] / d;
#pragma unroll 4
      for (int j = k + 1 + part; j <= row; j += 4) a[row][j] -= l * a[j][k];
    }
    __syncthreads();
  }
  if (t == 0 && bad && *info == 0) *info = (int)(k0 + bad);
  // L[i][j] = a[i][j] / sqrt(piv[j]) below the diagonal, sqrt(piv[i]) on it
  for (int e = t; e < CB * CB; e += 256) {
    const int i = e / CB, j = e % CB;
    if (j < i) a[i][j] *= 1.0 / sqrt(piv[j]);
    else if (j == i) a[i][i] = sqrt(piv[i]);
  }
  __syncthreads();
  for (int e = t; e < kb * kb; e += 256) {
    const int i = e / kb, j = e % kb;
    A[(k0 + i) * n + k0 + j] = (j <= i) ? a[i][j] : 0.0;
  }
  __syncthreads();
  // in-place inverse, columns from the last; rows i > j: 4 lanes per row split the dot product
  for (int j = CB - 1; j >= 0; --j) {
    const double ljj = a[j][j];
    double acc = 0.0;
    if (row > j) {
      double acc2 = 0.0;
      int k = j + 1 + part;
      for (; k + 4 <= row; k += 8) {  // two independent chains, loads issued ahead
        acc += a[row][k] * a[k][j];
        acc2 += a[row][k + 4] * a[k + 4][j];
      }
      if (k <= row) acc += a[row][k] * a[k][j];
      acc += acc2;
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (part == 0 && row > j) col[row] = -acc / ljj;
    __syncthreads();
    if (part == 0 && row > j) a[row][j] = col[row];
    if (t == 0) a[j][j] = 1.0 / ljj;
    __syncthreads();
  }
  for (int e = t; e < kb * kb; e += 256) {
    const int i = e / kb, j = e % kb;
    dinv[i * kb + j] = (j <= i) ? a[i][j] : 0.0;
  }
}

// Inverse of every diagonal block of L into X (block g on workgroup g): in-place TRTI2 order in LDS as in
// chol_diag_kernel (columns from the last, four threads per row per dot product).
__global__ void __launch_bounds__(256) diag_block_inverse_kernel(int64_t n, const double* __restrict__ L,
                                                                 double* __restrict__ X) {
  __shared__ double a[CB][CB + 1];
  __shared__ double col[CB];
  const int64_t k0 = (int64_t)blockIdx.x * CB;
  const int kb = (int)((n - k0) < CB ? (n - k0) : CB);
  const int t = threadIdx.x;
  for (int e = t; e < CB * CB; e += 256) {
    const int i = e / CB, j = e % CB;
    a[i][j] = (i < kb && j <= i) ? L[(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
  }
  __syncthreads();
  const int row = t >> 2, part = t & 3;
  for (int j = CB - 1; j >= 0; --j) {
    const double ljj = a[j][j];
    double acc = 0.0;
    if (row > j) {
      double acc2 = 0.0;
      int k = j + 1 + part;
      for (; k + 4 <= row; k += 8) {  // two independent chains, loads issued ahead
        acc += a[row][k] * a[k][j];
        acc2 += a[row][k + 4] * a[k + 4][j];
      }
      if (k <= row) acc += a[row][k] * a[k][j];
      acc += acc2;
    }
    acc += __shfl_xor(acc, 1, 64);
    acc += __shfl_xor(acc, 2, 64);
    if (part == 0 && row > j) col[row] = -acc / ljj;
    __syncthreads();
    if (part == 0 && row > j) a[row][j] = col[row];
    if (t == 0) a[j][j] = 1.0 / ljj;
    __syncthreads();
  }
  for (int e = t; e < kb * kb; e += 256) {
    const int i = e / kb, j = e % kb;
    X[(k0 + i) * n + k0 + j] = (j <= i) ? a[i][j] : 0.0;
  }
}

// dst[r][c] = src[r][c] for an rows x cols block with leading dimensions lds / ldd.
__global__ void copy2d_kernel(int64_t rows, int64_t cols, const double* __restrict__ src, int64_t lds,
                              double* __restrict__ dst, int64_t ldd) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < rows * cols; e += (int64_t)gridDim.x * blockDim.x)
    dst[(e / cols) * ldd + e % cols] = src[(e / cols) * lds + e % cols];
}

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// Zero the strict upper triangle.
__global__ void zero_upper_kernel(int64_t n, double* __restrict__ A) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n; e += (int64_t)gridDim.x * blockDim.x)
    if (e % n > e / n) A[e] = 0.0;
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
  int64_t k = kb + lane;
  for (; k + 192 < ke; k += 256) {  // four independent row loads in flight per lane
    const double m0 = mr[k], m1 = mr[k + 64], m2 = mr[k + 128], m3 = mr[k + 192];
#pragma unroll
    for (int s = 0; s < NR; ++s)
      acc[s] += m0 * X[k * NR + s] + m1 * X[(k + 64) * NR + s] + m2 * X[(k + 128) * NR + s] + m3 * X[(k + 192) * NR + s];
  }
  for (; k < ke; k += 64) {
    const double m = mr[k];
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

extern "C" int cc_linalg_work_size(int64_t n, int64_t* out) {
  *out = (int64_t)CB * CB + (int64_t)CB * n;
  return CC_OK;
}

// Right-looking blocked Cholesky: diagonal block in LDS (with its inverse), the panel as a GEMM
// against that inverse, the trailing update as a GEMM.
extern "C" int cc_cholesky_f64(void* stream, int64_t n, double* A, int* info_dev, double* work) {
  if (n <= 0) { cc_set_last_error("cholesky: n must be positive"); return CC_ERR_ARG; }
  hipStream_t st = cc_stream(stream);
  double* dinv = work;
  double* tmp = work + CB * CB;
  set_int_kernel<<<1, 1, 0, st>>>(info_dev, 0);
  CC_LAUNCH_CHECK();
  for (int64_t k0 = 0; k0 < n; k0 += CB) {
    int kb = (int)((n - k0) < CB ? (n - k0) : CB);
    chol_diag_kernel<<<1, 256, 0, st>>>(n, k0, kb, A, dinv, info_dev);
    CC_LAUNCH_CHECK();
    int64_t rest = n - k0 - kb;
    if (rest > 0) {
      double* Pn = A + (k0 + kb) * n + k0;
      // tmp = A_panel * Linv_kk^T ; A_panel = tmp
      int rc = cc_gemm_f64(stream, 1, rest, kb, kb, 1.0, Pn, 0, n, 1, dinv, 0, 1, kb, 0.0, tmp, 0, kb, 1);
      if (rc) return rc;
      copy2d_kernel<<<cc_blocks(rest * kb, 256), 256, 0, st>>>(rest, kb, tmp, kb, Pn, n);
      CC_LAUNCH_CHECK();
      double* T = A + (k0 + kb) * n + (k0 + kb);
      rc = cc_gemm_f64(stream, 1, rest, rest, kb, -1.0, Pn, 0, n, 1, Pn, 0, 1, n, 1.0, T, 0, n, 1);
      if (rc) return rc;
    }
  }
  zero_upper_kernel<<<cc_blocks(n * n, 256), 256, 0, st>>>(n, A);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// X = L^-1, blocked: diagonal-block inverses, then block row i: X[i, :i] = -X_ii (L[i, :i] X[:i, :i]).
extern "C" int cc_trtri_lower_f64(void* stream, int64_t n, const double* L, double* X, double* work) {
  if (n <= 0) { cc_set_last_error("trtri: n must be positive"); return CC_ERR_ARG; }
  hipStream_t st = cc_stream(stream);
  int rc = cc_fill_f64(stream, n * n, 0.0, X);
  if (rc) return rc;
  const int64_t nblk = (n + CB - 1) / CB;
  diag_block_inverse_kernel<<<(unsigned)nblk, 256, 0, st>>>(n, L, X);
  CC_LAUNCH_CHECK();
  double* S = work;  // CB x n
  for (int64_t i = 1; i < nblk; ++i) {
    int64_t r0 = i * CB, rb = (n - r0) < CB ? (n - r0) : CB, c = r0;
    rc = cc_gemm_f64(stream, 1, rb, c, c, 1.0, L + r0 * n, 0, n, 1, X, 0, n, 1, 0.0, S, 0, c, 1);
    if (rc) return rc;
    rc = cc_gemm_f64(stream, 1, rb, c, rb, -1.0, X + r0 * n + r0, 0, n, 1, S, 0, c, 1, 0.0, X + r0 * n, 0, n, 1);
    if (rc) return rc;
  }
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
