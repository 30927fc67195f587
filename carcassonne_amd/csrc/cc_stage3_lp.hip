// Low-precision fused stage-3 matvec (FP16 or BF16 operands, FP32 accumulation on matrix cores).
//
// Same arithmetic as the FP64 fused kernel.  Slabs are stored unpadded (alpha, P, P), each channel
// scaled by 1/max|slab| before rounding; the scales fold into the per-term site-operator
// coefficients.  One workgroup = NT waves (NT = ceil(P/16)); wave n owns the 16-row p' tile n:
//   stage A  W[idx = s*Q16 + q][p' in tile n] = sum_p V^T[idx][p] A_a[p'][p]   (v_mfma_f32_16x16x16)
//   stage B  Y[p' in tile n][q'][s'] += sum_q X[p'][q] B_b[q'][q],  X = sum_s O[s'][s] W[s*Q16 + q]
// The FP32 result of stage A (D[i][j] in lane j + 16(i/4), register i%4) is exactly the stage-B
// A operand (A[i][k] in lane i + 16(k/4), element k%4) once converted, so W never leaves registers.
// Per alpha the workgroup stages that alpha's slabs in LDS with coalesced 16-byte loads; MFMA
// operands are 8-byte LDS reads with zero fill outside the slab.  Partials are FP32 per workgroup
// and are summed in FP64 in a fixed order.
#include "cc_common.h"

#include <type_traits>

namespace {
constexpr int LP_MAX_CH = 16, LP_MAX_TERMS = 48, LP_MAX_D = 4;
typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

struct LPPlan {
  const void* A[LP_MAX_CH];
  const void* B[LP_MAX_CH];
  int n_a, n_b, n_terms;
  int ta[LP_MAX_TERMS], tb[LP_MAX_TERMS];
  float coef[LP_MAX_TERMS][LP_MAX_D * LP_MAX_D];  // scale_a * scale_b * O[s'][s]
  int64_t n_alpha;
};

template <typename T> struct LPType;
template <> struct LPType<_Float16> {
  typedef f16x4 vec;
  static __device__ inline float round(float x) { return (float)(_Float16)x; }
  static __device__ inline vec pack(float a, float b, float c, float d) {
    return (vec){(_Float16)a, (_Float16)b, (_Float16)c, (_Float16)d};
  }
  static __device__ inline f32x4 mfma(vec a, vec b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0); }
};
template <> struct LPType<__bf16> {
  typedef s16x4 vec;
  static __device__ inline float round(float x) { return (float)(__bf16)x; }
  static __device__ inline short bits(float x) {
    __bf16 h = (__bf16)x;
    return *reinterpret_cast<short*>(&h);
  }
  static __device__ inline vec pack(float a, float b, float c, float d) { return (vec){bits(a), bits(b), bits(c), bits(d)}; }
  static __device__ inline f32x4 mfma(vec a, vec b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }
};

// 4 consecutive elements of row `row`, columns c0..c0+3 of a P x P slab (zero outside the slab)
template <typename T, int P>
__device__ inline typename LPType<T>::vec slab4(const T* __restrict__ s, int row, int c0) {
  typedef typename LPType<T>::vec V;
  if ((P % 4) == 0) {  // c0 is a multiple of 4: the 4-vector is either fully inside or fully outside
    const bool ok = row < P && c0 < P;
    V v = *reinterpret_cast<const V*>(s + (ok ? row * P + c0 : 0));
    const V z = {0, 0, 0, 0};
    return ok ? v : z;
  }
  V v;
  for (int k = 0; k < 4; ++k) reinterpret_cast<T*>(&v)[k] = (row < P && c0 + k < P) ? s[row * P + c0 + k] : (T)0.0f;
  return v;
}

// SPLIT: the vector and the stage-B intermediate are carried as hi + lo pairs of the low-precision
// type (two MFMAs each), so only the slabs are rounded once; without it every operand is rounded.
// One wave owns a contiguous range of alpha and every p' tile; operands come straight from the
// unpadded slabs in global memory (8-byte fragments, zero outside).  The 4 waves of a workgroup
// reduce their Y through LDS in a fixed order.
template <int P, int DP, typename T, bool SPLIT>
__global__ void __launch_bounds__(256) lp_kernel(const LPPlan plan, const double* __restrict__ v,
                                                 float* __restrict__ partial) {
  typedef LPType<T> LT;
  typedef typename LT::vec V;
  constexpr int NT = (P + 15) / 16, KU = NT, Q16 = 16 * KU, MT = DP * KU;
  constexpr int SLAB = P * P, NY = P * P * DP;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ float red[4 * NY > 256 ? 4 * NY : 256];

  // max|v| over the whole vector (fixed order) -> the vector is used as v / max|v|
  {
    float mx = 0.f;
    for (int i = threadIdx.x; i < NY; i += 256) mx = fmaxf(mx, (float)fabs(v[i]));
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + st]);
      __syncthreads();
    }
  }
  const double vinv = 1.0 / (double)(red[0] > 0.f ? red[0] : 1.f);
  __syncthreads();

  V vt[MT][KU], vl[MT][KU];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int u = 0; u < KU; ++u) {
      const int idx = 16 * m + (lane & 15), s = idx / Q16, q = idx % Q16;
      float e[4], r[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int p = 16 * u + 4 * (lane >> 4) + k;
        e[k] = (s < DP && q < P && p < P) ? (float)(v[(p * P + q) * DP + s] * vinv) : 0.0f;
        r[k] = e[k] - LT::round(e[k]);
      }
      vt[m][u] = LT::pack(e[0], e[1], e[2], e[3]);
      vl[m][u] = LT::pack(r[0], r[1], r[2], r[3]);
    }

  f32x4 y[NT][NT][DP];
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int s = 0; s < DP; ++s) y[i][j][s] = (f32x4){0.f, 0.f, 0.f, 0.f};

  const int64_t gw = (int64_t)blockIdx.x * 4 + wave, nw = (int64_t)gridDim.x * 4;
  const int64_t a0 = plan.n_alpha * gw / nw, a1 = plan.n_alpha * (gw + 1) / nw;
  for (int64_t al = a0; al < a1; ++al) {
    int t = 0;
    while (t < plan.n_terms) {
      const int a = plan.ta[t];
      const T* As = reinterpret_cast<const T*>(plan.A[a]) + al * SLAB;
      int t_end = t;
      while (t_end < plan.n_terms && plan.ta[t_end] == a) ++t_end;
#pragma unroll
      for (int np = 0; np < NT; ++np) {
        f32x4 w[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) w[m] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < KU; ++u) {
          const V bop = slab4<T, P>(As, 16 * np + (lane & 15), 16 * u + 4 * (lane >> 4));
#pragma unroll
          for (int m = 0; m < MT; ++m) {
            w[m] = LT::mfma(vt[m][u], bop, w[m]);
            if (SPLIT) w[m] = LT::mfma(vl[m][u], bop, w[m]);
          }
        }
        for (int tt = t; tt < t_end; ++tt) {
          const T* Bs = reinterpret_cast<const T*>(plan.B[plan.tb[tt]]) + al * SLAB;
          const float* cf = plan.coef[tt];
#pragma unroll
          for (int u = 0; u < KU; ++u) {
            V xo[DP], xl[DP];
#pragma unroll
            for (int sp = 0; sp < DP; ++sp) {
              f32x4 acc = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
              for (int s = 0; s < DP; ++s) acc += cf[sp * DP + s] * w[s * KU + u];
              xo[sp] = LT::pack(acc[0], acc[1], acc[2], acc[3]);
              xl[sp] = LT::pack(acc[0] - LT::round(acc[0]), acc[1] - LT::round(acc[1]), acc[2] - LT::round(acc[2]),
                                acc[3] - LT::round(acc[3]));
            }
#pragma unroll
            for (int nq = 0; nq < NT; ++nq) {
              const V bop = slab4<T, P>(Bs, 16 * nq + (lane & 15), 16 * u + 4 * (lane >> 4));
#pragma unroll
              for (int sp = 0; sp < DP; ++sp) {
                y[np][nq][sp] = LT::mfma(xo[sp], bop, y[np][nq][sp]);
                if (SPLIT) y[np][nq][sp] = LT::mfma(xl[sp], bop, y[np][nq][sp]);
              }
            }
          }
        }
      }
      t = t_end;
    }
  }
  // fixed-order workgroup reduction: wave w writes slot w, the slots are summed 0..3
  __syncthreads();
#pragma unroll
  for (int np = 0; np < NT; ++np)
#pragma unroll
    for (int nq = 0; nq < NT; ++nq)
#pragma unroll
      for (int sp = 0; sp < DP; ++sp)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int pp = 16 * np + 4 * (lane >> 4) + r, qq = 16 * nq + (lane & 15);
          if (pp < P && qq < P) red[wave * NY + (pp * P + qq) * DP + sp] = y[np][nq][sp][r];
        }
  __syncthreads();
  float* out = partial + (int64_t)blockIdx.x * NY;
  for (int i = threadIdx.x; i < NY; i += 256) out[i] = ((red[i] + red[NY + i]) + red[2 * NY + i]) + red[3 * NY + i];
}

constexpr int LP_SLICES = 16;
// pass 1: slices[j][i] = sum over groups g = j, j + LP_SLICES, ... of partial[g][i]   (FP64 in a fixed order)
__global__ void __launch_bounds__(256) lp_reduce1_kernel(int64_t groups, int64_t n, const float* __restrict__ partial,
                                                         double* __restrict__ slices) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int j = blockIdx.y;
  if (i >= n) return;
  double acc = 0.0;
  for (int64_t g = j; g < groups; g += LP_SLICES) acc += (double)partial[g * n + i];
  slices[(int64_t)j * n + i] = acc;
}
// pass 2: y[i] = max|v| * sum_j slices[j][i]  (max|v| recomputed exactly as the main kernel does)
__global__ void __launch_bounds__(256) lp_reduce2_kernel(int64_t n, const double* __restrict__ slices,
                                                         const double* __restrict__ v, double* __restrict__ y) {
  __shared__ float red[256];
  float mx = 0.f;
  for (int64_t k = threadIdx.x; k < n; k += 256) mx = fmaxf(mx, (float)fabs(v[k]));
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const double scale = red[0] > 0.f ? (double)red[0] : 1.0;
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double t = 0.0;
  for (int j = 0; j < LP_SLICES; ++j) t += slices[(int64_t)j * n + i];
  y[i] = t * scale;
}

// low-precision slab = slab / max|slab|: pass 1 per-block maxima, pass 2 (one block) their maximum
__global__ void __launch_bounds__(256) lp_blockmax_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double red[256];
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) m = fmax(m, fabs(x[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void __launch_bounds__(1024) lp_absmax_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ out) {
  __shared__ double red[1024];
  double m = 0.0;
  for (int64_t i = (int64_t)threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(x[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] > 0.0 ? red[0] : 1.0;
}
template <typename T>
__global__ void lp_convert_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ scale, T* __restrict__ y) {
  const double inv = 1.0 / scale[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (T)(x[i] * inv);
}

struct LPHandle {
  LPPlan plan;
  int64_t P, d;
  int dtype, groups, split;
  float* work;  // groups * P*P*d partials + n scaled vector + 1 scale
};

template <int P, int DP, typename T>
int launch_lp(hipStream_t st, const LPHandle* h, const double* v, double* y) {
  constexpr int NT = (P + 15) / 16;
  const int64_t n = (int64_t)P * P * DP;
  float* partial = h->work;
  double* slices = reinterpret_cast<double*>(partial + (((int64_t)h->groups * n + 1) & ~1LL));
  if (h->split) lp_kernel<P, DP, T, true><<<h->groups, 256, 0, st>>>(h->plan, v, partial);
  else lp_kernel<P, DP, T, false><<<h->groups, 256, 0, st>>>(h->plan, v, partial);
  CC_LAUNCH_CHECK();
  lp_reduce1_kernel<<<dim3((unsigned)((n + 255) / 256), LP_SLICES), 256, 0, st>>>(h->groups, n, partial, slices);
  CC_LAUNCH_CHECK();
  lp_reduce2_kernel<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, slices, v, y);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

template <typename T>
int dispatch_lp(hipStream_t st, const LPHandle* h, const double* v, double* y) {
  switch (h->P) {
    case 4: return launch_lp<4, 2, T>(st, h, v, y);
    case 9: return launch_lp<9, 2, T>(st, h, v, y);
    case 16: return launch_lp<16, 2, T>(st, h, v, y);
    case 25: return launch_lp<25, 2, T>(st, h, v, y);
    case 36: return launch_lp<36, 2, T>(st, h, v, y);
  }
  cc_set_last_error("low-precision stage-3: shape not instantiated");
  return CC_ERR_UNSUPPORTED;
}
}  // namespace

int cc_internal_lp_apply(hipStream_t st, const void* plan, const double* x, double* y) {
  const LPHandle* h = static_cast<const LPHandle*>(plan);
  return h->dtype == CC_F16 ? dispatch_lp<_Float16>(st, h, x, y) : dispatch_lp<__bf16>(st, h, x, y);
}
int64_t cc_internal_lp_n(const void* plan) {
  const LPHandle* h = static_cast<const LPHandle*>(plan);
  return h->P * h->P * h->d;
}

extern "C" int cc_stage3_lp_convert(void* stream, int dtype, int64_t n, const double* slab, double* scale_out, void* out) {
  hipStream_t st = cc_stream(stream);
  if (n <= 0) return CC_OK;
  lp_blockmax_kernel<<<512, 256, 0, st>>>(n, slab, scale_out + 1);   // scale_out holds 513 doubles
  CC_LAUNCH_CHECK();
  lp_absmax_kernel<<<1, 1024, 0, st>>>(512, scale_out + 1, scale_out);
  CC_LAUNCH_CHECK();
  if (dtype == CC_F16) lp_convert_kernel<_Float16><<<cc_blocks(n, 256), 256, 0, st>>>(n, slab, scale_out, (_Float16*)out);
  else if (dtype == CC_BF16) lp_convert_kernel<__bf16><<<cc_blocks(n, 256), 256, 0, st>>>(n, slab, scale_out, (__bf16*)out);
  else { cc_set_last_error("lp convert: dtype must be f16 or bf16"); return CC_ERR_ARG; }
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_stage3_lp_work_size(int64_t n_alpha, int64_t P, int64_t d, int groups, int64_t* out) {
  *out = (int64_t)groups * P * P * d + 2 * LP_SLICES * P * P * d + 16;  // floats (the slices are doubles)
  return CC_OK;
}

extern "C" int cc_stage3_lp_plan_create(void** out, int dtype, int n_a, const void* const* A, int n_b,
                                        const void* const* B, int n_terms, const int* term_a, const int* term_b,
                                        const float* term_coef, int64_t n_alpha, int64_t P, int64_t d, int groups,
                                        int split, float* work) {
  *out = nullptr;
  if (!(dtype == CC_F16 || dtype == CC_BF16)) { cc_set_last_error("lp plan: dtype must be f16 or bf16"); return CC_ERR_ARG; }
  if (!cc_stage3_fused_supported(P, d)) { cc_set_last_error("lp plan: shape not instantiated"); return CC_ERR_UNSUPPORTED; }
  if (n_a < 1 || n_a > LP_MAX_CH || n_b < 1 || n_b > LP_MAX_CH || n_terms < 1 || n_terms > LP_MAX_TERMS || groups < 1 ||
      n_alpha < 1 || d > LP_MAX_D) {
    cc_set_last_error("lp plan: exceeds kernel limits");
    return CC_ERR_UNSUPPORTED;
  }

  LPHandle* h = new LPHandle();
  h->plan = LPPlan{};
  for (int i = 0; i < n_a; ++i) h->plan.A[i] = A[i];
  for (int i = 0; i < n_b; ++i) h->plan.B[i] = B[i];
  h->plan.n_a = n_a;
  h->plan.n_b = n_b;
  h->plan.n_terms = n_terms;
  for (int t = 0; t < n_terms; ++t) {
    if (term_a[t] < 0 || term_a[t] >= n_a || term_b[t] < 0 || term_b[t] >= n_b || (t > 0 && term_a[t] < term_a[t - 1])) {
      delete h;
      cc_set_last_error("lp plan: bad term table (indices out of range or not sorted by A channel)");
      return CC_ERR_ARG;
    }
    h->plan.ta[t] = term_a[t];
    h->plan.tb[t] = term_b[t];
    for (int e = 0; e < d * d; ++e) h->plan.coef[t][e] = term_coef[t * d * d + e];
  }
  h->plan.n_alpha = n_alpha;
  h->P = P;
  h->d = d;
  h->dtype = dtype;
  h->groups = groups;
  h->split = split ? 1 : 0;
  h->work = work;
  *out = h;
  return CC_OK;
}

extern "C" int cc_stage3_lp_plan_apply(void* stream, const void* plan, const double* x, double* y) {
  if (!plan) { cc_set_last_error("null plan"); return CC_ERR_ARG; }
  return cc_internal_lp_apply(cc_stream(stream), plan, x, y);
}

extern "C" int cc_stage3_lp_plan_destroy(void* plan) {
  delete static_cast<LPHandle*>(plan);
  return CC_OK;
}
