// Dense FP64 linear algebra for the solve: blocked Cholesky, lower-triangular inverse,
// skinny products with the whitening factor, Lanczos reorthogonalisation, and the
// lowest eigenpair of a symmetric tridiagonal matrix.
#include "cc_common.h"

namespace {
constexpr int CB = 64;  // Cholesky block size

__device__ inline double bcast(double v, int lane) {
  // value of `v` in lane `lane` (a compile-time constant at every call site) through two readlanes
  int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Cholesky of the kb x kb diagonal block at (k0, k0) on one wave: lane i holds row i in registers,
// columns are broadcast with readlane.  The factor is staged in LDS and inverted (lane c solves
// column c of L x = e_c with L rows read as LDS broadcasts).  Writes L back to A and L^-1 to dinv.
__global__ void __launch_bounds__(64) chol_diag_kernel(int64_t n, int64_t k0, int kb, double* __restrict__ A,
                                                       double* __restrict__ dinv, int* __restrict__ info) {
  const int i = threadIdx.x;
  double r[CB];
#pragma unroll
  for (int c = 0; c < CB; ++c) r[c] = (i < kb && c <= i) ? A[(k0 + i) * n + k0 + c] : (i == c ? 1.0 : 0.0);
  int bad = 0;
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    double djj = bcast(r[j], j);
    if (j < kb && !(djj > 0.0)) {
      if (!bad) bad = j + 1;   // report the first non-positive pivot
      djj = 1e-300;
    }
    const double sj = sqrt(djj);
    r[j] = (i == j) ? sj : (i > j ? r[j] / sj : r[j]);
#pragma unroll
    for (int c = j + 1; c < CB; ++c) {
      double acj = bcast(r[j], c);
      if (c <= i) r[c] -= r[j] * acj;
    }
  }
  if (i == 0 && bad && *info == 0) *info = (int)(k0 + bad);
  // inverse: column `i` of X = L^-1 (row `row` of L is lane `row`'s registers)
  double x[CB];
#pragma unroll
  for (int row = 0; row < CB; ++row) {
    double acc = (row == i) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < row; ++k) acc -= bcast(r[k], row) * x[k];
    x[row] = (row < i) ? 0.0 : acc / bcast(r[row], row);
  }
  if (i < kb) {
#pragma unroll
    for (int c = 0; c < CB; ++c)
      if (c < kb) A[(k0 + i) * n + k0 + c] = (c <= i) ? r[c] : 0.0;
  }
  // dinv[row][i] = x[row]
#pragma unroll
  for (int row = 0; row < CB; ++row)
    if (row < kb && i < kb) dinv[row * kb + i] = x[row];
}

// Inverse of every diagonal block of L into X (block g on workgroup g).
__global__ void __launch_bounds__(256) diag_block_inverse_kernel(int64_t n, const double* __restrict__ L,
                                                                 double* __restrict__ X) {
  __shared__ double a[CB][CB + 1];
  __shared__ double x[CB][CB + 1];
  const int64_t k0 = (int64_t)blockIdx.x * CB;
  const int kb = (int)((n - k0) < CB ? (n - k0) : CB);
  for (int e = threadIdx.x; e < kb * kb; e += 256) {
    int i = e / kb, j = e % kb;
    a[i][j] = L[(k0 + i) * n + k0 + j];
  }
  __syncthreads();
  for (int c = threadIdx.x; c < kb; c += 256) {
    for (int i = 0; i < kb; ++i) {
      double v = (i == c) ? 1.0 : 0.0;
      for (int k = c; k < i; ++k) v -= a[i][k] * x[k][c];
      x[i][c] = (i < c) ? 0.0 : v / a[i][i];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kb * kb; e += 256) {
    int i = e / kb, j = e % kb;
    X[(k0 + i) * n + k0 + j] = x[i][j];
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
    chol_diag_kernel<<<1, 64, 0, st>>>(n, k0, kb, A, dinv, info_dev);
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
