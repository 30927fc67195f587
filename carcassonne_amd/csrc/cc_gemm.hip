// FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4f64) with arbitrary element
// strides, batched over grid.z.  A 256-thread workgroup computes a T x T tile (T = 64 or 32)
// in K steps of 32; the four waves own (T/2) x (T/2) sub-tiles, i.e. 2x2 MFMA tiles (T = 64,
// 32 MFMAs per wave and K step, which covers the global-load latency of the next step's
// operands, prefetched into registers) or one MFMA tile (T = 32, for grids that would leave
// most CUs idle with 64-wide tiles).
// Operand fragments (probed on gfx950):
//   A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16(i%4), reg i/4.
#include "cc_common.h"

namespace {
constexpr int BK = 32, PAD = 2;

struct GemmArgs {
  int64_t M, N, K;
  double alpha, beta;
  const double* A; int64_t sAb, sAm, sAk;
  const double* B; int64_t sBb, sBk, sBn;
  double* C; int64_t sCb, sCm, sCn;
  int64_t batch0;
};

// Loader layouts: which index runs across consecutive threads.
template <int T, bool A_K_FAST, bool B_N_FAST>
__global__ void __launch_bounds__(256) gemm_f64_kernel(GemmArgs g) {
  constexpr int TM = T / 32;           // 16-wide MFMA tiles per wave and dimension
  constexpr int NL = T * BK / 256;     // operand elements per thread and matrix per K step
  __shared__ double As[BK][T + PAD];
  __shared__ double Bs[BK][T + PAD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  const int64_t bz = g.batch0 + blockIdx.z;
  const double* A = g.A + bz * g.sAb;
  const double* B = g.B + bz * g.sBb;
  double* C = g.C + bz * g.sCb;
  const int64_t m0 = (int64_t)blockIdx.y * T, n0 = (int64_t)blockIdx.x * T;

  f64x4 acc[TM][TM];
  for (int a = 0; a < TM; ++a)
    for (int b = 0; b < TM; ++b) acc[a][b] = (f64x4){0.0, 0.0, 0.0, 0.0};

  double ra[NL], rb[NL];
  // element r of this thread: (row-or-column index, k) of the T x BK operand tile
  auto a_idx = [&](int r, int& i, int& k) {
    if (A_K_FAST) { k = tid & (BK - 1); i = tid / BK + (256 / BK) * r; }
    else { i = tid % T; k = tid / T + (256 / T) * r; }
  };
  auto b_idx = [&](int r, int& j, int& k) {
    if (B_N_FAST) { j = tid % T; k = tid / T + (256 / T) * r; }
    else { k = tid & (BK - 1); j = tid / BK + (256 / BK) * r; }
  };
  auto load = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int i, k;
      a_idx(r, i, k);
      int64_t gi = m0 + i, gk = k0 + k;
      ra[r] = (gi < g.M && gk < g.K) ? A[gi * g.sAm + gk * g.sAk] : 0.0;
    }
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int j, k;
      b_idx(r, j, k);
      int64_t gj = n0 + j, gk = k0 + k;
      rb[r] = (gj < g.N && gk < g.K) ? B[gk * g.sBk + gj * g.sBn] : 0.0;
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int i, k;
      a_idx(r, i, k);
      As[k][i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < NL; ++r) {
      int j, k;
      b_idx(r, j, k);
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
      double af[TM], bf[TM];
#pragma unroll
      for (int t = 0; t < TM; ++t) {
        af[t] = As[kk + (lane >> 4)][wr * (T / 2) + t * 16 + (lane & 15)];
        bf[t] = Bs[kk + (lane >> 4)][wc * (T / 2) + t * 16 + (lane & 15)];
      }
#pragma unroll
      for (int a = 0; a < TM; ++a)
#pragma unroll
        for (int b = 0; b < TM; ++b) acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[a], bf[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }

#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TM; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int64_t gi = m0 + wr * (T / 2) + a * 16 + 4 * r + (lane >> 4);
        int64_t gj = n0 + wc * (T / 2) + b * 16 + (lane & 15);
        if (gi < g.M && gj < g.N) {
          double* p = C + gi * g.sCm + gj * g.sCn;
          double v = g.alpha * acc[a][b][r];
          *p = (g.beta == 0.0) ? v : v + g.beta * (*p);
        }
      }
}

template <int T>
void launch_gemm(const GemmArgs& g, dim3 grid, bool a_k_fast, bool b_n_fast, hipStream_t st) {
  if (a_k_fast && b_n_fast) gemm_f64_kernel<T, true, true><<<grid, 256, 0, st>>>(g);
  else if (a_k_fast) gemm_f64_kernel<T, true, false><<<grid, 256, 0, st>>>(g);
  else if (b_n_fast) gemm_f64_kernel<T, false, true><<<grid, 256, 0, st>>>(g);
  else gemm_f64_kernel<T, false, false><<<grid, 256, 0, st>>>(g);
}
}  // namespace

extern "C" int cc_gemm_f64(void* stream, int64_t batch, int64_t M, int64_t N, int64_t K, double alpha,
                           const double* A, int64_t sAb, int64_t sAm, int64_t sAk,
                           const double* B, int64_t sBb, int64_t sBk, int64_t sBn, double beta,
                           double* C, int64_t sCb, int64_t sCm, int64_t sCn) {
  if (batch < 0 || M < 0 || N < 0 || K < 0) { cc_set_last_error("cc_gemm_f64: negative size"); return CC_ERR_ARG; }
  if (batch == 0 || M == 0 || N == 0) return CC_OK;
  GemmArgs g{M, N, K, alpha, beta, A, sAb, sAm, sAk, B, sBb, sBk, sBn, C, sCb, sCm, sCn, 0};
  // 64-wide tiles unless that grid would leave most of the 256 CUs idle
  const int64_t tiles64 = ((N + 63) / 64) * ((M + 63) / 64) * batch;
  const int T = tiles64 < 160 ? 32 : 64;
  int64_t gx = (N + T - 1) / T, gy = (M + T - 1) / T;
  if (gx > 2147483647LL || gy > 65535) { cc_set_last_error("cc_gemm_f64: matrix too large for one launch"); return CC_ERR_UNSUPPORTED; }
  bool a_k_fast = (sAk == 1) || (sAm != 1 && (sAk < sAm));
  bool b_n_fast = (sBn == 1) || (sBk != 1 && (sBn < sBk));
  hipStream_t st = cc_stream(stream);
  for (int64_t b0 = 0; b0 < batch; b0 += 65535) {
    int64_t nb = batch - b0 < 65535 ? batch - b0 : 65535;
    g.batch0 = b0;
    dim3 grid((unsigned)gx, (unsigned)gy, (unsigned)nb);
    if (T == 32) launch_gemm<32>(g, grid, a_k_fast, b_n_fast, st);
    else launch_gemm<64>(g, grid, a_k_fast, b_n_fast, st);
    CC_LAUNCH_CHECK();
  }
  return CC_OK;
}
