// Learned warm start: a 3-layer MLP proposing the eigensolver's start vector, with its own training.
//
//   x (K0) -> h1 = relu(W1 x + b1) (H) -> h2 = relu(W2 h1 + b2) (H) -> y = W3 h2 + b3 (NO)
//
// Parameters live in one FP32 master buffer laid out [W1 (H x K0), b1 (H), W2 (H x H), b2 (H), W3 (NO x H),
// b3 (NO)] (weights row-major out x in) with an FP16 mirror.  The forward pass multiplies FP16 activations
// by FP16 weights on matrix cores (v_mfma_f32_16x16x16f16, FP32 accumulation) and keeps each hidden layer
// in FP32 and as the FP16 copy the next layer reads.  The backward pass is FP32 (tiled FMA GEMMs over the
// same FP16 values the forward used), the loss is the sign-free misalignment 1 - (y.v)^2 / (|y|^2 |v|^2)
// with the converged state v, and Adam updates the FP32 master and refreshes the FP16 mirror.  All sums
// run in a fixed order.
#include "cc_common.h"

namespace {
typedef _Float16 h4 __attribute__((ext_vector_type(4)));

struct MlpDims {
  int64_t K0, H, NO;
  int64_t w1, b1, w2, b2, w3, b3, total;
};
MlpDims mlp_dims(int64_t K0, int64_t H, int64_t NO) {
  MlpDims d;
  d.K0 = K0;
  d.H = H;
  d.NO = NO;
  d.w1 = 0;
  d.b1 = d.w1 + H * K0;
  d.w2 = d.b1 + H;
  d.b2 = d.w2 + H * H;
  d.w3 = d.b2 + H;
  d.b3 = d.w3 + NO * H;
  d.total = d.b3 + NO;
  return d;
}

// Y[i][j] = act(sum_k X[i][k] W[j][k] + b[j]) for a 64 x 64 output tile per 256-thread workgroup; each wave
// owns a 32 x 32 quarter as 2 x 2 MFMA tiles.  X (M x K) and W (N x K) are FP16 row-major, K % 16 == 0,
// N % 64 == 0; rows past M are read as zero and not written.  Writes Y (FP32) and, if Yh, its FP16 copy.
// Operand fragments: A[i][k] in lane i + 16(k/4), element k%4; B[k][j] in lane j + 16(k/4), element k%4;
// D[i][j] in lane j + 16(i/4), register i%4.
__global__ void __launch_bounds__(256) mlp_linear_kernel(int64_t M, int64_t N, int64_t K, const _Float16* __restrict__ X,
                                                         const _Float16* __restrict__ W, const float* __restrict__ b,
                                                         int relu, float* __restrict__ Y, _Float16* __restrict__ Yh) {
  __shared__ h4 As[64][4 + 1];
  __shared__ h4 Bs[64][4 + 1];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, wy = wave >> 1, wx = wave & 1;
  const int64_t i0 = (int64_t)blockIdx.y * 64, j0 = (int64_t)blockIdx.x * 64;
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c) acc[a][c] = (f32x4){0.f, 0.f, 0.f, 0.f};
  const int lr = t >> 2, lq = t & 3;  // loader: row lr of the tile, 4-element quad lq of the 16-wide K step
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    const int64_t xi = i0 + lr;
    h4 xa = {};
    if (xi < M) xa = *reinterpret_cast<const h4*>(X + xi * K + k0 + 4 * lq);
    As[lr][lq] = xa;
    Bs[lr][lq] = *reinterpret_cast<const h4*>(W + (j0 + lr) * K + k0 + 4 * lq);
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const h4 af = As[32 * wy + 16 * a + (lane & 15)][lane >> 4];
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const h4 bf = Bs[32 * wx + 16 * c + (lane & 15)][lane >> 4];
        acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x16f16(af, bf, acc[a][c], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t i = i0 + 32 * wy + 16 * a + 4 * (lane >> 4) + r, j = j0 + 32 * wx + 16 * c + (lane & 15);
        if (i < M) {
          float v = acc[a][c][r] + b[j];
          if (relu) v = v > 0.f ? v : 0.f;
          Y[i * N + j] = v;
          if (Yh) Yh[i * N + j] = (_Float16)v;
        }
      }
}

// FP32 GEMM C[i][j] = sum_k A(i, k) B(k, j) (+ C[i][j] when accumulate), A(i, k) = A[i*sai + k*sak] and
// B(k, j) = B[k*sbk + j*sbj] in element strides, either operand FP32 or FP16 (converted on load).  64 x 64
// tile per 256-thread workgroup, 4 x 4 outputs per thread, K in steps of 16 through LDS.
template <typename TA, typename TB>
__global__ void __launch_bounds__(256) mlp_sgemm_kernel(int64_t M, int64_t N, int64_t K, const TA* __restrict__ A,
                                                        int64_t sai, int64_t sak, const TB* __restrict__ B, int64_t sbk,
                                                        int64_t sbj, float* __restrict__ C, int accumulate) {
  __shared__ float As[16][64 + 1];
  __shared__ float Bs[16][64 + 1];
  const int t = threadIdx.x, ty = t >> 4, tx = t & 15;
  const int64_t i0 = (int64_t)blockIdx.y * 64, j0 = (int64_t)blockIdx.x * 64;
  float acc[4][4] = {};
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    for (int e = t; e < 16 * 64; e += 256) {
      const int kk = e >> 6, ii = e & 63;
      const int64_t i = i0 + ii, j = j0 + ii, k = k0 + kk;
      As[kk][ii] = (i < M && k < K) ? (float)A[i * sai + k * sak] : 0.f;
      Bs[kk][ii] = (j < N && k < K) ? (float)B[k * sbk + j * sbj] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float av[4], bv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        av[r] = As[kk][ty + 16 * r];
        bv[r] = Bs[kk][tx + 16 * r];
      }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[r][c] += av[r] * bv[c];
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int64_t i = i0 + ty + 16 * r, j = j0 + tx + 16 * c;
      if (i < M && j < N) C[i * N + j] = acc[r][c] + (accumulate ? C[i * N + j] : 0.f);
    }
}

// g[j] = sum_i D[i][j] (rows in order), one thread per column
__global__ void mlp_colsum_kernel(int64_t M, int64_t N, const float* __restrict__ D, float* __restrict__ g) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int64_t i = 0; i < M; ++i) s += D[i * N + j];
  g[j] = s;
}

// D *= (Hact > 0) elementwise (ReLU backward through the stored activation)
__global__ void mlp_relu_bwd_kernel(int64_t n, const float* __restrict__ Hact, float* __restrict__ D) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
    if (!(Hact[e] > 0.f)) D[e] = 0.f;
}

// Per sample b (one 256-thread workgroup): loss[b] = 1 - (y.v)^2 / (|y|^2 |v|^2) and
// dY[b] = scale * d loss / d y = scale * (2 (y.v)^2 / (|y|^4 |v|^2) y - 2 (y.v) / (|y|^2 |v|^2) v).
__global__ void __launch_bounds__(256) mlp_loss_kernel(int64_t NO, const float* __restrict__ Y, const float* __restrict__ V,
                                                       float scale, float* __restrict__ loss, float* __restrict__ dY) {
  __shared__ float red[3][4];
  const int64_t b = blockIdx.x;
  const float* y = Y + b * NO;
  const float* v = V + b * NO;
  float yy = 0.f, vv = 0.f, yv = 0.f;
  for (int64_t i = threadIdx.x; i < NO; i += 256) {
    yy += y[i] * y[i];
    vv += v[i] * v[i];
    yv += y[i] * v[i];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    yy += __shfl_xor(yy, o, 64);
    vv += __shfl_xor(vv, o, 64);
    yv += __shfl_xor(yv, o, 64);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][wave] = yy;
    red[1][wave] = vv;
    red[2][wave] = yv;
  }
  __syncthreads();
  yy = ((red[0][0] + red[0][1]) + red[0][2]) + red[0][3];
  vv = ((red[1][0] + red[1][1]) + red[1][2]) + red[1][3];
  yv = ((red[2][0] + red[2][1]) + red[2][2]) + red[2][3];
  const float yyc = fmaxf(yy, 1e-30f), vvc = fmaxf(vv, 1e-30f);
  const float cy = scale * 2.f * yv * yv / (yyc * yyc * vvc), cv = -scale * 2.f * yv / (yyc * vvc);
  for (int64_t i = threadIdx.x; i < NO; i += 256) dY[b * NO + i] = cy * y[i] + cv * v[i];
  if (threadIdx.x == 0) loss[b] = 1.f - yv * yv / (yyc * vvc);
}

// Adam on the FP32 master parameters with bias correction; refreshes the FP16 mirror.
__global__ void mlp_adam_kernel(int64_t n, float* __restrict__ P, const float* __restrict__ G, float* __restrict__ Mv,
                                float* __restrict__ Vv, _Float16* __restrict__ P16, float lr, float b1, float b2, float eps,
                                float c1, float c2) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
    const float g = G[e];
    const float m = b1 * Mv[e] + (1.f - b1) * g, v = b2 * Vv[e] + (1.f - b2) * g * g;
    Mv[e] = m;
    Vv[e] = v;
    const float p = P[e] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[e] = p;
    P16[e] = (_Float16)p;
  }
}

// One feature row: each of the n_spec spectra (k singular values, descending) divided by its largest value,
// then h, zero padding to K0; FP16.
__global__ void mlp_features_kernel(int n_spec, int k, const double* __restrict__ S, double h, int64_t K0,
                                    _Float16* __restrict__ out) {
  for (int64_t e = threadIdx.x; e < K0; e += blockDim.x) {
    float v = 0.f;
    if (e < (int64_t)n_spec * k) {
      const int sp = (int)(e / k);
      const double top = S[(int64_t)sp * k];
      v = top > 0.0 ? (float)(S[e] / top) : 0.f;
    } else if (e == (int64_t)n_spec * k) {
      v = (float)h;
    }
    out[e] = (_Float16)v;
  }
}

template <typename TA, typename TB>
int sgemm(hipStream_t st, int64_t M, int64_t N, int64_t K, const TA* A, int64_t sai, int64_t sak, const TB* B, int64_t sbk,
          int64_t sbj, float* C, int accumulate) {
  const dim3 grid((unsigned)((N + 63) / 64), (unsigned)((M + 63) / 64));
  mlp_sgemm_kernel<TA, TB><<<grid, 256, 0, st>>>(M, N, K, A, sai, sak, B, sbk, sbj, C, accumulate);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

int check_dims(int64_t B, int64_t K0, int64_t H, int64_t NO) {
  if (B < 1 || K0 < 16 || K0 % 16 || H < 64 || H % 64 || NO < 64 || NO % 64) {
    cc_set_last_error("mlp: need batch >= 1, K0 % 16 == 0, H and NO multiples of 64");
    return CC_ERR_DIM;
  }
  return CC_OK;
}
}  // namespace

extern "C" int cc_mlp_param_count(int64_t K0, int64_t H, int64_t NO, int64_t* out) {
  *out = mlp_dims(K0, H, NO).total;
  return CC_OK;
}

extern "C" int cc_mlp_features(void* stream, int n_spec, int k, const double* S, double h, int64_t K0, void* out16) {
  if (n_spec < 1 || k < 1 || (int64_t)n_spec * k + 1 > K0) {
    cc_set_last_error("mlp features: K0 must hold n_spec * k + 1 values");
    return CC_ERR_DIM;
  }
  mlp_features_kernel<<<1, 256, 0, cc_stream(stream)>>>(n_spec, k, S, h, K0, static_cast<_Float16*>(out16));
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_mlp_forward(void* stream, int64_t B, int64_t K0, int64_t H, int64_t NO, const void* P16, const float* P32,
                              const void* X16, void* A16, float* A32, float* Y) {
  int rc = check_dims(B, K0, H, NO);
  if (rc) return rc;
  hipStream_t st = cc_stream(stream);
  const MlpDims d = mlp_dims(K0, H, NO);
  const _Float16* p = static_cast<const _Float16*>(P16);
  _Float16* h1h = static_cast<_Float16*>(A16);
  _Float16* h2h = h1h + B * H;
  float* h1 = A32;
  float* h2 = A32 + B * H;
  const unsigned gy = (unsigned)((B + 63) / 64);
  mlp_linear_kernel<<<dim3((unsigned)(H / 64), gy), 256, 0, st>>>(B, H, K0, static_cast<const _Float16*>(X16), p + d.w1,
                                                                  P32 + d.b1, 1, h1, h1h);
  CC_LAUNCH_CHECK();
  mlp_linear_kernel<<<dim3((unsigned)(H / 64), gy), 256, 0, st>>>(B, H, H, h1h, p + d.w2, P32 + d.b2, 1, h2, h2h);
  CC_LAUNCH_CHECK();
  mlp_linear_kernel<<<dim3((unsigned)(NO / 64), gy), 256, 0, st>>>(B, NO, H, h2h, p + d.w3, P32 + d.b3, 0, Y, nullptr);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_mlp_work_size(int64_t B, int64_t H, int64_t NO, int64_t* out) {
  *out = B * NO + 2 * B * H;  // floats: dY, dZ2, dZ1
  return CC_OK;
}

extern "C" int cc_mlp_backward(void* stream, int64_t B, int64_t K0, int64_t H, int64_t NO, const void* P16,
                               const void* X16, const void* A16, const float* A32, const float* Y, const float* target,
                               float* loss, float* G, float* work) {
  int rc = check_dims(B, K0, H, NO);
  if (rc) return rc;
  hipStream_t st = cc_stream(stream);
  const MlpDims d = mlp_dims(K0, H, NO);
  const _Float16* p = static_cast<const _Float16*>(P16);
  const _Float16* x = static_cast<const _Float16*>(X16);
  const _Float16* h1h = static_cast<const _Float16*>(A16);
  const _Float16* h2h = h1h + B * H;
  const float* h1 = A32;
  const float* h2 = A32 + B * H;
  float* dY = work;
  float* dZ2 = dY + B * NO;
  float* dZ1 = dZ2 + B * H;
  mlp_loss_kernel<<<(unsigned)B, 256, 0, st>>>(NO, Y, target, 1.f / (float)B, loss, dY);
  CC_LAUNCH_CHECK();
  // layer 3: dW3 = dY^T h2h, db3 = colsum dY, dZ2 = (dY W3) * [h2 > 0]
  if ((rc = sgemm<float, _Float16>(st, NO, H, B, dY, 1, NO, h2h, H, 1, G + d.w3, 0))) return rc;
  mlp_colsum_kernel<<<(unsigned)((NO + 255) / 256), 256, 0, st>>>(B, NO, dY, G + d.b3);
  CC_LAUNCH_CHECK();
  if ((rc = sgemm<float, _Float16>(st, B, H, NO, dY, NO, 1, p + d.w3, H, 1, dZ2, 0))) return rc;
  mlp_relu_bwd_kernel<<<cc_blocks(B * H, 256), 256, 0, st>>>(B * H, h2, dZ2);
  CC_LAUNCH_CHECK();
  // layer 2
  if ((rc = sgemm<float, _Float16>(st, H, H, B, dZ2, 1, H, h1h, H, 1, G + d.w2, 0))) return rc;
  mlp_colsum_kernel<<<(unsigned)((H + 255) / 256), 256, 0, st>>>(B, H, dZ2, G + d.b2);
  CC_LAUNCH_CHECK();
  if ((rc = sgemm<float, _Float16>(st, B, H, H, dZ2, H, 1, p + d.w2, H, 1, dZ1, 0))) return rc;
  mlp_relu_bwd_kernel<<<cc_blocks(B * H, 256), 256, 0, st>>>(B * H, h1, dZ1);
  CC_LAUNCH_CHECK();
  // layer 1
  if ((rc = sgemm<float, _Float16>(st, H, K0, B, dZ1, 1, H, x, K0, 1, G + d.w1, 0))) return rc;
  mlp_colsum_kernel<<<(unsigned)((H + 255) / 256), 256, 0, st>>>(B, H, dZ1, G + d.b1);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_mlp_adam(void* stream, int64_t n, float* P32, const float* G, float* M, float* V, void* P16, double lr,
                           double beta1, double beta2, double eps, int64_t step) {
  if (step < 1) { cc_set_last_error("mlp adam: step counts from 1"); return CC_ERR_ARG; }
  const float c1 = (float)(1.0 - pow(beta1, (double)step)), c2 = (float)(1.0 - pow(beta2, (double)step));
  mlp_adam_kernel<<<cc_blocks(n, 256), 256, 0, cc_stream(stream)>>>(n, P32, G, M, V, static_cast<_Float16*>(P16),
                                                                    (float)lr, (float)beta1, (float)beta2, (float)eps, c1, c2);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
