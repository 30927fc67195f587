// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64) with arbitrary element
// strides, batched over grid.z.  Tile 64x64x32 per 256-thread workgroup; each wave
// owns a 32x32 sub-tile (2x2 MFMA tiles), 32 MFMAs per K step, which covers the
// global-load latency of the next step's operands (prefetched into registers).
// Operand fragments (probed on gfx950):
//   A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i%4), reg i/4.
#include "cc_common.h"

namespace {
constexpr int BM = 64, BN = 64, BK = 32, PAD = 2;

struct GemmArgs {
  int64_t M, N, K;
  double alpha, beta;
  const double* A; int64_t sAb, sAm, sAk;
  const double* B; int64_t sBb, sBk, sBn;
  double* C; int64_t sCb, sCm, sCn;
  int64_t batch0;
};

// Loader layouts: which index runs across consecutive threads.
template <bool A_K_FAST, bool B_N_FAST>
__global__ void __launch_bounds__(256) gemm_f64_kernel(GemmArgs g) {
  __shared__ double As[BK][BM + PAD];
  __shared__ double Bs[BK][BN + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t bz = g.batch0 + blockIdx.z;
  const double* A = g.A + bz * g.sAb;
  const double* B = g.B + bz * g.sBb;
  double* C = g.C + bz * g.sCb;
  const int64_t m0 = (int64_t)blockIdx.y * BM, n0 = (int64_t)blockIdx.x * BN;

  f64x4 acc[2][2];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) acc[a][b] = (f64x4){0.0, 0.0, 0.0, 0.0};

  constexpr int NL = BM * BK / 256;  // operand elements per thread and matrix per K step
  double ra[NL], rb[NL];
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int i, k;
      if (A_K_FAST) { k = tid & (BK - 1); i = tid / BK + (256 / BK) * r; }
      else { i = tid & 63; k = (tid >> 6) + 4 * r; }
      int64_t gi = m0 + i, gk = k0 + k;
      ra[r] = (gi < g.M && gk < g.K) ? A[gi * g.sAm + gk * g.sAk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int j, k;
      if (B_N_FAST) { j = tid & 63; k = (tid >> 6) + 4 * r; }
      else { k = tid & (BK - 1); j = tid / BK + (256 / BK) * r; }
      int64_t gj = n0 + j, gk = k0 + k;
      rb[r] = (gj < g.N && gk < g.K) ? B[gk * g.sBk + gj * g.sBn] : 0.0;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int i, k;
      if (A_K_FAST) { k = tid & (BK - 1); i = tid / BK + (256 / BK) * r; }
      else { i = tid & 63; k = (tid >> 6) + 4 * r; }
      As[k][i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int j, k;
      if (B_N_FAST) { j = tid & 63; k = (tid >> 6) + 4 * r; }
      else { k = tid & (BK - 1); j = tid / BK + (256 / BK) * r; }
      Bs[k][j] = rb[r];
    }
  };

  if (g.K > 0) load(0);
  for (int64_t k0 = 0; k0 < g.K; k0 += BK) {
    store();
    __syncthreads();
    if (k0 + BK < g.K) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double af[2], bf[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        af[t] = As[kk + (lane >> 4)][wr * 32 + t * 16 + (lane & 15)];
        bf[t] = Bs[kk + (lane >> 4)][wc * 32 + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }

#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t gi = m0 + wr * 32 + a * 16 + 4 * r + (lane >> 4);
        int64_t gj = n0 + wc * 32 + b * 16 + (lane & 15);
        if (gi < g.M && gj < g.N) {
          double* p = C + gi * g.sCm + gj * g.sCn;
          double v = g.alpha * acc[a][b][r];
          *p = (g.beta == 0.0) ? v : v + g.beta * (*p);
        }
      }
}
}  // namespace

extern "C" int cc_gemm_f64(void* stream, int64_t batch, int64_t M, int64_t N, int64_t K, double alpha,
                           const double* A, int64_t sAb, int64_t sAm, int64_t sAk,
                           const double* B, int64_t sBb, int64_t sBk, int64_t sBn, double beta,
                           double* C, int64_t sCb, int64_t sCm, int64_t sCn) {
  if (batch < 0 || M < 0 || N < 0 || K < 0) { cc_set_last_error("cc_gemm_f64: negative size"); return CC_ERR_ARG; }
  if (batch == 0 || M == 0 || N == 0) return CC_OK;
  GemmArgs g{M, N, K, alpha, beta, A, sAb, sAm, sAk, B, sBb, sBk, sBn, C, sCb, sCm, sCn, 0};
  int64_t gx = (N + BN - 1) / BN, gy = (M + BM - 1) / BM;
  if (gx > 2147483647LL || gy > 65535) { cc_set_last_error("cc_gemm_f64: matrix too large for one launch"); return CC_ERR_UNSUPPORTED; }
  bool a_k_fast = (sAk == 1) || (sAm != 1 && (sAk < sAm));
  bool b_n_fast = (sBn == 1) || (sBk != 1 && (sBn < sBk));
  hipStream_t st = cc_stream(stream);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    g.batch0 = b0;
    dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)nb);
    if (a_k_fast && b_n_fast) gemm_f64_kernel<true, true><<<grid, 256, 0, st>>>(g);
    else if (a_k_fast) gemm_f64_kernel<true, false><<<grid, 256, 0, st>>>(g);
    else if (b_n_fast) gemm_f64_kernel<false, true><<<grid, 256, 0, st>>>(g);
    else gemm_f64_kernel<false, false><<<grid, 256, 0, st>>>(g);
    CC_LAUNCH_CHECK();
  }
  return CC_OK;
}
