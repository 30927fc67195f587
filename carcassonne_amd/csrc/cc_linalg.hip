// Dense FP64 linear algebra for the solve: blocked Cholesky, lower-triangular inverse,
// skinny products with the whitening factor, Lanczos reorthogonalisation, and the
// lowest eigenpair of a symmetric tridiagonal matrix.
#include "cc_common.h"

namespace {
constexpr int CB = 64;  // Cholesky block size

// Lane l's value of v (the two 32-bit halves through readlane; l uniform).
__device__ inline double readlane_f64(double v, int l) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u & 0xffffffffull), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

constexpr int CH = 16, NCH = CB / CH;  // broadcast operands are read from LDS in chunks of CH

// r[j] -= s * src(j) for every j > K.  src is read in chunks of CH, one chunk ahead of its FMAs; the
// address of each chunk's loads depends (through an empty asm that takes the FMA results) on the chunk
// two back, so at most two chunks of loaded values are live whatever order the scheduler picks (left to
// itself it issues all 63 loads of a step first and spills).  K is a template parameter so that every
// index into r is a constant and r stays in registers.
template <int K, typename F>
__device__ __forceinline__ void tri_sweep(double (&r)[CB], double s, F src) {
  constexpr int c0 = (K + 1) / CH;
  if constexpr (c0 < NCH) {
    int tok = 0;
    double cur[CH], nxt[CH];
#pragma unroll
    for (int u = 0; u < CH; ++u) cur[u] = src(c0 * CH + u);
#pragma unroll
    for (int c = c0; c < NCH; ++c) {
#pragma unroll
      for (int u = 0; u < CH; ++u) nxt[u] = (c + 1 < NCH) ? src((c + 1) * CH + u + tok) : 0.0;
#pragma unroll
      for (int u = 0; u < CH; ++u)
        if (c * CH + u > K) r[c * CH + u] -= s * cur[u];
#pragma unroll
      for (int h = 0; h < CH; h += 8)
        __asm__ volatile("" : "+v"(tok) : "v"(r[c * CH + h + 0]), "v"(r[c * CH + h + 1]), "v"(r[c * CH + h + 2]),
                         "v"(r[c * CH + h + 3]), "v"(r[c * CH + h + 4]), "v"(r[c * CH + h + 5]), "v"(r[c * CH + h + 6]),
                         "v"(r[c * CH + h + 7]));
#pragma unroll
      for (int u = 0; u < CH; ++u) cur[u] = nxt[u];
    }
  }
}

// Cholesky and inverse of a 64 x 64 block in one right-looking sweep, row i of the block held by lane i
// as c[j].  Before step i, c[j] (j > i) holds the mirrored Schur complement S[j][i]; at step i that is
// sqrt(d_i) L[j][i], i.e. -d_i X[j][i] for X = L^-1 (d_i the pivot), and afterwards the forward
// substitution for column i of X applied to y = -d_i x has the same form as the Cholesky update.  So
// every row, whether it still factors (i > K) or already inverts (i < K), does
//   s_i = c[K] / sqrt(d_K),  c[K] = s_i,  c[j] -= s_i s_j  (j > K),
// and row K keeps c[K] = sqrt(d_K) with s_K = 0.  On exit c[j] = L[i][j] (j < i), c[i] = L[i][i] and
// c[j] = -d_i X[j][i] (j > i).  The symmetric updates keep c_i[j] == c_j[i] bit for bit (the same two
// factors meet in each FMA).  A non-positive pivot of a real row records its 1-based index in bad and is
// replaced by 1e-300.
// Blocks of 64R x 64R on 4R waves: wave w holds columns 16w .. 16w + 15, lane i rows i + 64r (r < R).
// The sweep goes in panels of PB = 4 columns.  The wave owning a panel factors it alone: per step the
// pivot, the multipliers (published in LDS) and the step's update of the panel's later columns, with the
// row multipliers it needs taken by readlane.  After one barrier every wave with columns past the panel
// applies the panel's four updates, in step order, so each element sees the same FMAs in the same order
// as in a column-by-column sweep.  The owner of the next panel first brings that panel's columns up to
// date and factors it into the other LDS buffer (look-ahead), so the pivot chain overlaps the other
// columns' updates.  Panels at or past the block's kb real rows only meet the identity padding and change
// nothing, so the sweep ends there (steps past kb inside the last panel leave the real rows unchanged).
constexpr int CW = 16, PB = 4;
// Steps K .. KE - 1 of one panel in its owner wave (wave K / CW); multipliers of step K into sp[K % PB].
template <int K, int KE, int R>
__device__ __forceinline__ void chol_panel_steps(double (&c)[R][CW], double (*sp)[64 * R], int i, int kb, int& bad) {
  if constexpr (K < KE) {
    constexpr int k = K % CW;
    double d = readlane_f64(c[K / 64][k], K % 64);
    if (!(d > 0.0)) {
      if (K < kb && bad == 0) bad = K + 1;
      d = 1e-300;
    }
    // 1/sqrt(d): the hardware estimate and two Newton steps y += (y/2)(1 - d y^2) (each squares the
    // relative error; within an ulp or so of the rounded quotient), sqrt(d) = d / sqrt(d); far shorter
    // than a correctly rounded square root followed by a division on this step's critical path
    double inv = __builtin_amdgcn_rsq(d);
#pragma unroll
    for (int it = 0; it < 2; ++it) inv = fma(0.5 * inv, fma(-d * inv, inv, 1.0), inv);
    const double sd = d * inv;
    double sv[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int row = i + 64 * r;
      sv[r] = (row == K) ? 0.0 : c[r][k] * inv;
      c[r][k] = (row == K) ? sd : sv[r];
      sp[K % PB][row] = sv[r];
    }
    // this step on the panel's later columns j (row j's multiplier from its lane)
#pragma unroll
    for (int j = K + 1; j < KE; ++j) {
      const double sj = readlane_f64(sv[j / 64], j % 64);
#pragma unroll
      for (int r = 0; r < R; ++r) c[r][j % CW] -= sv[r] * sj;
    }
    chol_panel_steps<K + 1, KE, R>(c, sp, i, kb, bad);
  }
}

// Panel K0 .. K0 + PB - 1 (its multipliers already in sbuf[(K0 / PB) & 1]) applied past the panel.
template <int K0, int R>
__device__ __forceinline__ void chol_inv_panels(double (&c)[R][CW], double (*sbuf)[PB][64 * R], int w, int i, int kb, int& bad) {
  if constexpr (K0 < 64 * R) {
    if (K0 >= kb) return;
    constexpr int o = K0 / CW, k0 = K0 % CW;
    constexpr bool ahead = K0 + PB < 64 * R;
    constexpr int K1 = ahead ? K0 + PB : 0, o1 = K1 / CW, k1 = K1 % CW;
    __syncthreads();
    double (*sp)[64 * R] = sbuf[(K0 / PB) & 1];
    if (w >= o) {
      double si[PB][R];
#pragma unroll
      for (int kk = 0; kk < PB; ++kk)
#pragma unroll
        for (int r = 0; r < R; ++r) si[kk][r] = sp[kk][i + 64 * r];
      if (ahead && w == o1) {
#pragma unroll
        for (int jj = k1; jj < k1 + PB; ++jj)
#pragma unroll
          for (int kk = 0; kk < PB; ++kk) {
            const double sj = sp[kk][CW * o1 + jj];
#pragma unroll
            for (int r = 0; r < R; ++r) c[r][jj] -= si[kk][r] * sj;
          }
        if constexpr (ahead) chol_panel_steps<K1, K1 + PB, R>(c, sbuf[(K1 / PB) & 1], i, kb, bad);
      }
      // this wave's columns past the panel, less the next panel's when the look-ahead took them
      const double* sw = &sp[0][0] + CW * w;
      const int j0 = (w == o) ? k0 + PB : 0;
#pragma unroll
      for (int jj = 0; jj < CW; ++jj)
        if (jj >= j0 && !(ahead && w == o1 && jj >= k1 && jj < k1 + PB)) {
#pragma unroll
          for (int kk = 0; kk < PB; ++kk) {
            const double sj = sw[kk * 64 * R + jj];
#pragma unroll
            for (int r = 0; r < R; ++r) c[r][jj] -= si[kk][r] * sj;
          }
        }
    }
    chol_inv_panels<K0 + PB, R>(c, sbuf, w, i, kb, bad);
  }
}

// Cholesky and inverse of one diagonal block of at most 64R x 64R (see chol_inv_panels): A's block at
// (k0, k0) (lower triangle read; rows >= kb padded with the identity) is replaced by L with its upper
// triangle zeroed, X's block at (x0, x0) receives L^-1, info the first non-positive pivot (1-based,
// global index).
template <int R>
__global__ void __launch_bounds__(256 * R) chol_leaf_kernel(int64_t lda, int64_t k0, int kb, double* __restrict__ A,
                                                         double* __restrict__ X, int64_t ldx, int64_t x0,
                                                         int* __restrict__ info) {
  constexpr int BN = 64 * R, NW = BN / CW;
  __shared__ double Ls[BN][BN + 1];
  __shared__ double sbuf[2][PB][BN];
  __shared__ int badw[NW];
  const int i = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  double c[R][CW];
  // row i + 64r of the lower triangle, then its column mirrored into j > row (identity past kb); every
  // load index clamped into the block
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = i + 64 * r, rc = row < kb ? row : 0;
#pragma unroll
    for (int jj = 0; jj < CW; ++jj) {
      const int j = CW * w + jj, jc = j < kb ? j : 0;
      const double v = A[(k0 + (j <= row ? rc : jc)) * lda + k0 + (j <= row ? jc : rc)];
      c[r][jj] = (row < kb && j < kb) ? v : (row == j ? 1.0 : 0.0);
    }
  }
  int bad = 0;  // first failing pivot among this wave's columns
  if (w == 0) chol_panel_steps<0, PB, R>(c, sbuf[0], i, kb, bad);
  chol_inv_panels<0, R>(c, sbuf, w, i, kb, bad);
  if (i == 0) badw[w] = bad;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int jj = 0; jj < CW; ++jj) Ls[i + 64 * r][CW * w + jj] = c[r][jj];
  __syncthreads();
  if (threadIdx.x == 0 && *info == 0)
    for (int v = 0; v < NW; ++v)
      if (badw[v]) {
        *info = (int)(k0 + badw[v]);
        break;
      }
  for (int rr = w; rr < kb; rr += NW)
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int col = i + 64 * r;
      if (col < kb) A[(k0 + rr) * lda + k0 + col] = (col <= rr) ? Ls[rr][col] : 0.0;
    }
  // X[m][row] = -c_row[m] / d_row below the diagonal, 1 / L[row][row] on it
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int row = i + 64 * r;
    if (row < kb) {
      const double xd = 1.0 / Ls[row][row], rd = -xd * xd;
      for (int m = w; m < kb; m += NW) X[(x0 + m) * ldx + x0 + row] = (m > row) ? Ls[row][m] * rd : (m == row ? xd : 0.0);
    }
  }
}

// Column sweep of L x = e_j, steps Q..CB-1: x[Q] *= 1/L[Q][Q], then x[m] -= L[m][Q] x[Q] for m > Q.
template <int Q>
__device__ __forceinline__ void inv_steps(double (&x)[CB], const double (*Ls)[CB + 1], const double* rdiag) {
  if constexpr (Q < CB) {
    const double xq = x[Q] * rdiag[Q];
    x[Q] = xq;
    tri_sweep<Q>(x, xq, [&](int m) { return Ls[m][Q]; });
    inv_steps<Q + 1>(x, Ls, rdiag);
  }
}

// Inverse of one lower-triangular diagonal block of at most 64 x 64 on one wave: the block of A at (k0, k0)
// holds L (rows >= kb padded with the identity); lane j solves L x = e_j with L broadcast from LDS and
// writes column j of L^-1 into X at (x0, x0).
__global__ void __launch_bounds__(64) tri_leaf_kernel(int64_t lda, int64_t k0, int kb, const double* __restrict__ A,
                                                      double* __restrict__ X, int64_t ldx, int64_t x0) {
  __shared__ double Ls[CB][CB + 1];
  __shared__ double rdiag[CB];
  const int i = threadIdx.x;
  const int ic = i < kb ? i : 0;  // clamped: every load below is in bounds
  double r[CB];
  const double* arow = A + (k0 + ic) * lda + k0;
#pragma unroll
  for (int j = 0; j < CB; ++j) {
    const double v = arow[j < kb ? j : 0];
    r[j] = (i < kb && j <= i) ? v : (i == j ? 1.0 : 0.0);
  }
#pragma unroll
  for (int j = 0; j < CB; ++j) Ls[i][j] = (j <= i) ? r[j] : 0.0;
  __syncthreads();
  rdiag[i] = 1.0 / Ls[i][i];
  __syncthreads();
  double x[CB];
#pragma unroll
  for (int m = 0; m < CB; ++m) x[m] = (m == i) ? 1.0 : 0.0;
  inv_steps<0>(x, Ls, rdiag);
#pragma unroll
  for (int m = 0; m < CB; ++m)  // unrolled: x stays in registers
    if (i < kb && m < kb) X[(x0 + m) * ldx + x0 + i] = x[m];
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
// tri: 0 full, 1 lower (k <= i), 2 upper (k >= i).  Two waves per output row, each summing one half of
// the row's range (half the serial load rounds of one wave per row); the halves are added in order.
template <int NR>
__global__ void __launch_bounds__(256) skinny_kernel(int64_t n, int tri, const double* __restrict__ M,
                                                     const double* __restrict__ X, double* __restrict__ Y) {
  __shared__ double part[4][NR];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, half = wave & 1;
  const int64_t row = (int64_t)blockIdx.x * 2 + (wave >> 1);
  double acc[NR];
#pragma unroll
  for (int s = 0; s < NR; ++s) acc[s] = 0.0;
  if (row < n) {
    const int64_t rb = (tri == 2) ? row : 0, re = (tri == 1) ? row + 1 : n, mid = rb + (re - rb) / 2;
    const int64_t kb = half ? mid : rb, ke = half ? re : mid;
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
    for (int s = 0; s < NR; ++s) part[wave][s] = acc[s];
  }
  __syncthreads();
  if (lane == 0 && half == 0 && row < n) {
#pragma unroll
    for (int s = 0; s < NR; ++s) Y[row * NR + s] = part[wave][s] + part[wave + 1][s];
  }
}

// B = A^T for n x n row-major matrices: one 32 x 32 tile per workgroup (32 x 8 threads), read and written
// along rows through an LDS tile padded to 33 columns.
__global__ void __launch_bounds__(256) transpose_kernel(int64_t n, const double* __restrict__ A, double* __restrict__ B) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const int64_t r0 = (int64_t)blockIdx.y * 32, c0 = (int64_t)blockIdx.x * 32;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int64_t r = r0 + ty + k, c = c0 + tx;
    if (r < n && c < n) tile[ty + k][tx] = A[r * n + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const int64_t r = c0 + ty + k, c = r0 + tx;   // B[r][c] = A[c][r]
    if (r < n && c < n) B[r * n + c] = tile[tx][ty + k];
  }
}

// Split of a diagonal range of sz > CB rows at a block boundary near its middle.
inline int64_t split_rows(int64_t sz) { return ((sz + CB - 1) / CB / 2) * CB; }

// Recursive Cholesky with inverse of the diagonal range [off, off + sz) of A (leading dimension n):
//   (L11, X11) = chol_inv(A11); L21 = A21 X11^T; A22 -= L21 L21^T; (L22, X22) = chol_inv(A22);
//   X21 = -X22 (L21 X11).
// The GEMMs have the full K of each level (not the block width), so they fill the chip; leaves are
// chol_leaf_kernel (blocks of up to 128 rows, split at multiples of 128).  tmp holds n1 x n2 <= sz^2 / 4
// doubles.
int chol_inv_rec(void* stream, int64_t n, int64_t off, int64_t sz, double* A, double* X, double* tmp, int* info) {
  hipStream_t st = cc_stream(stream);
  if (sz <= 2 * CB) {
    if (sz <= CB) chol_leaf_kernel<1><<<1, 256, 0, st>>>(n, off, (int)sz, A, X, n, off, info);
    else chol_leaf_kernel<2><<<1, 512, 0, st>>>(n, off, (int)sz, A, X, n, off, info);
    CC_LAUNCH_CHECK();
    return CC_OK;
  }
  const int64_t n1 = ((sz + 2 * CB - 1) / (2 * CB) / 2) * (2 * CB), n2 = sz - n1;
  int rc = chol_inv_rec(stream, n, off, n1, A, X, tmp, info);
  if (rc) return rc;
  double* A21 = A + (off + n1) * n + off;
  double* A22 = A21 + n1;
  const double* X11 = X + off * n + off;
  double* X21 = X + (off + n1) * n + off;
  const double* X22 = X21 + n1;
  rc = cc_gemm_f64(stream, 1, n2, n1, n1, 1.0, A21, 0, n, 1, X11, 0, 1, n, 0.0, tmp, 0, n1, 1);
  if (rc) return rc;
  copy2d_kernel<<<cc_blocks(n2 * n1, 256), 256, 0, st>>>(n2, n1, tmp, n1, A21, n);
  CC_LAUNCH_CHECK();
  rc = cc_gemm_f64(stream, 1, n2, n2, n1, -1.0, A21, 0, n, 1, A21, 0, 1, n, 1.0, A22, 0, n, 1);
  if (rc) return rc;
  rc = chol_inv_rec(stream, n, off + n1, n2, A, X, tmp, info);
  if (rc) return rc;
  rc = cc_gemm_f64(stream, 1, n2, n1, n1, 1.0, A21, 0, n, 1, X11, 0, n, 1, 0.0, tmp, 0, n1, 1);
  if (rc) return rc;
  return cc_gemm_f64(stream, 1, n2, n1, n2, -1.0, X22, 0, n, 1, tmp, 0, n1, 1, 0.0, X21, 0, n, 1);
}

// Recursive inverse of the lower-triangular range [off, off + sz) of L: X11, X22 recursively, then
// X21 = -X22 (L21 X11); leaves are tri_leaf_kernel.
int trtri_rec(void* stream, int64_t n, int64_t off, int64_t sz, const double* L, double* X, double* tmp) {
  hipStream_t st = cc_stream(stream);
  if (sz <= CB) {
    tri_leaf_kernel<<<1, 64, 0, st>>>(n, off, (int)sz, L, X, n, off);
    CC_LAUNCH_CHECK();
    return CC_OK;
  }
  const int64_t n1 = split_rows(sz), n2 = sz - n1;
  int rc = trtri_rec(stream, n, off, n1, L, X, tmp);
  if (rc) return rc;
  rc = trtri_rec(stream, n, off + n1, n2, L, X, tmp);
  if (rc) return rc;
  const double* L21 = L + (off + n1) * n + off;
  const double* X11 = X + off * n + off;
  double* X21 = X + (off + n1) * n + off;
  rc = cc_gemm_f64(stream, 1, n2, n1, n1, 1.0, L21, 0, n, 1, X11, 0, n, 1, 0.0, tmp, 0, n1, 1);
  if (rc) return rc;
  return cc_gemm_f64(stream, 1, n2, n1, n2, -1.0, X21 + n1, 0, n, 1, tmp, 0, n1, 1, 0.0, X21, 0, n, 1);
}
}  // namespace

extern "C" int cc_linalg_work_size(int64_t n, int64_t* out) {
  const int64_t blocked = (int64_t)CB * CB + (int64_t)CB * n, recursive = n * n / 4 + 1;
  *out = blocked > recursive ? blocked : recursive;
  return CC_OK;
}

// Right-looking blocked Cholesky: diagonal block factored and inverted by chol_leaf_kernel, the panel as a GEMM
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
    chol_leaf_kernel<1><<<1, 256, 0, st>>>(n, k0, kb, A, dinv, kb, 0, info_dev);
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

// A = L L^T in place (upper zeroed) and X = L^-1 (upper zero) in one recursive pass.
extern "C" int cc_cholesky_inverse_f64(void* stream, int64_t n, double* A, double* X, int* info_dev, double* work) {
  if (n <= 0) { cc_set_last_error("cholesky_inverse: n must be positive"); return CC_ERR_ARG; }
  hipStream_t st = cc_stream(stream);
  set_int_kernel<<<1, 1, 0, st>>>(info_dev, 0);
  CC_LAUNCH_CHECK();
  int rc = cc_fill_f64(stream, n * n, 0.0, X);
  if (rc) return rc;
  rc = chol_inv_rec(stream, n, 0, n, A, X, work, info_dev);
  if (rc) return rc;
  zero_upper_kernel<<<cc_blocks(n * n, 256), 256, 0, st>>>(n, A);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// X = L^-1 for lower-triangular L, recursively (leaves on one wave each, GEMMs for the off-diagonal blocks).
extern "C" int cc_trtri_lower_f64(void* stream, int64_t n, const double* L, double* X, double* work) {
  if (n <= 0) { cc_set_last_error("trtri: n must be positive"); return CC_ERR_ARG; }
  int rc = cc_fill_f64(stream, n * n, 0.0, X);
  if (rc) return rc;
  return trtri_rec(stream, n, 0, n, L, X, work);
}

extern "C" int cc_transpose_f64(void* stream, int64_t n, const double* A, double* B) {
  if (n <= 0) return CC_OK;
  const int64_t tiles = (n + 31) / 32;
  if (tiles > 65535) { cc_set_last_error("transpose: n too large for one launch"); return CC_ERR_UNSUPPORTED; }
  transpose_kernel<<<dim3((unsigned)tiles, (unsigned)tiles), 256, 0, cc_stream(stream)>>>(n, A, B);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_skinny_f64(void* stream, int64_t n, int tri, const double* M, int nrhs, const double* X, double* Y) {
  if (n <= 0) return CC_OK;
  unsigned g = (unsigned)((n + 1) / 2);
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
