// Fused stage-3 matvec on CDNA4 matrix cores (FP64, v_mfma_f64_16x16x4f64).
//
//   Y[p'][q'][s'] = sum_t sum_alpha sum_{p,q,s} A_a(t)[alpha][p'][p] B_b(t)[alpha][q'][q] O_o(t)[s'][s] v[p][q][s]
//
// One wave owns a contiguous range of alpha.  Per alpha and per distinct A channel a:
//   stage A   W_a[idx = s*P4 + q][p'] = sum_p V^T[idx][p] A_a[p'][p]      (V^T operands stay in registers)
// and per term t of that channel:
//   stage B   Y[p'][q'][s'] += sum_q (sum_s O[s'][s] W_a[s*P4+q][p']) B_b[q'][q]
// The f64 MFMA result layout (D[i][j] in lane j + 16(i%4), register i/4) puts W_a[s*P4 + 4u + lane/16][p']
// in the lane/register that stage B needs as its A operand X[p' = lane%16][q = 4u + lane/16], so the
// intermediate never leaves registers.  Slabs are pre-packed in MFMA-operand order
// (packed[alpha][u][n][lane] = S[16n + lane%16][4u + lane/16], zero-padded), so each operand load is one
// coalesced 512-byte read.  Each 4-wave workgroup reduces its waves' Y through LDS in a fixed order and
// writes one partial; a second kernel sums the partials in a fixed order.
#include "cc_common.h"

#include <mutex>
#include <vector>

namespace {
// Optional timing of every fused-kernel launch (HIP events on the launch stream).

constexpr int MAX_CH = 16, MAX_TERMS = 48, MAX_OPS = 16, MAX_D = 4;

struct FusedPlan {
  const double* A[MAX_CH];
  const double* B[MAX_CH];
  int n_terms;
  int ta[MAX_TERMS], tb[MAX_TERMS], to[MAX_TERMS];  // terms sorted by channel a; to = -1 -> identity
  double ops[MAX_OPS][MAX_D * MAX_D];
  int64_t n_alpha;
};

template <int P, int DP>
struct Shape {
  static constexpr int P4 = (P + 3) / 4 * 4;      // padded p / q extent (multiple of 4)
  static constexpr int NT = (P + 15) / 16;         // 16-wide tiles over p' and q'
  static constexpr int KS = P4 / 4;                // 4-deep k steps over p (stage A) and q (stage B)
  static constexpr int IDX = DP * P4;              // stage-A rows idx = s*P4 + q
  static constexpr int MT = (IDX + 15) / 16;       // stage-A row tiles
  static constexpr int SLAB = KS * NT * 64;        // packed doubles per alpha
};

constexpr int PACK_MAX_P = 64;
// One alpha per workgroup step: the slab is gathered into LDS in source order (coalesced reads), then
// written in MFMA-operand order (packed[alpha][u][n][lane] = S[16n + lane%16][4u + lane/16], zero padded).
__global__ void __launch_bounds__(256) pack_kernel(int64_t n_alpha, int NT, int KS, CCSlabGather g,
                                                   const double* __restrict__ src, double* __restrict__ dst) {
  extern __shared__ double s[];  // P x (P + 1)
  const int P = g.P, ld = P + 1;
  const int per = KS * NT * 64;
  for (int64_t al = blockIdx.x; al < n_alpha; al += gridDim.x) {
    cc_slab_load(g, src, al, s, ld, 1.0);
    __syncthreads();
    double* out = dst + al * per;
    for (int r = threadIdx.x; r < per; r += blockDim.x) {
      const int lane = r & 63, t = r >> 6, n = t % NT, u = t / NT;
      const int row = 16 * n + (lane & 15), col = 4 * u + (lane >> 4);
      out[r] = (row < P && col < P) ? s[row * ld + col] : 0.0;
    }
    __syncthreads();
  }
}

// The same for slabs whose four legs all have extent D (constant index arithmetic).
template <int D>
__global__ void __launch_bounds__(256) pack_kernel_d(int64_t n_alpha, CCSlabGather g, const double* __restrict__ src,
                                                     double* __restrict__ dst) {
  constexpr int P = D * D, NT = (P + 15) / 16, KS = (P + 3) / 4, per = KS * NT * 64, ld = P + 1;
  __shared__ double s[P * ld];
  for (int64_t al = blockIdx.x; al < n_alpha; al += gridDim.x) {
    cc_slab_load_d<D, 256>(g, src, al, s, ld, 1.0);
    __syncthreads();
    double* out = dst + al * per;
    for (int r = threadIdx.x; r < per; r += blockDim.x) {
      const int lane = r & 63, t = r >> 6, n = t % NT, u = t / NT;
      const int row = 16 * n + (lane & 15), col = 4 * u + (lane >> 4);
      out[r] = (row < P && col < P) ? s[row * ld + col] : 0.0;
    }
    __syncthreads();
  }
}

template <int P, int DP>
__global__ void __launch_bounds__(256) fused_f64_kernel(const FusedPlan plan, const double* __restrict__ v,
                                                        double* __restrict__ partial) {
  using S = Shape<P, DP>;
  constexpr int P4 = S::P4, NT = S::NT, KS = S::KS, MT = S::MT, SLAB = S::SLAB;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t gw = (int64_t)blockIdx.x * 4 + wave, nw = (int64_t)gridDim.x * 4;
  const int64_t a0 = plan.n_alpha * gw / nw, a1 = plan.n_alpha * (gw + 1) / nw;

  // stage-A A operands V^T[idx][p] (idx = 16m + lane%16, p = 4u + lane/16), stored in LDS in lane order so
  // each read is conflict-free; the same LDS is reused for the final reduction.
  constexpr int NY = P * P * DP;
  constexpr int NVT = MT * KS * 64;
  constexpr int NLDS = (NVT > 4 * NY) ? NVT : 4 * NY;
  __shared__ double lds[NLDS];
  for (int e = threadIdx.x; e < NVT; e += 256) {
    int l = e & 63, mu = e >> 6, u = mu % KS, m = mu / KS;
    int idx = 16 * m + (l & 15), p = 4 * u + (l >> 4);
    int s = idx / P4, q = idx % P4;
    lds[e] = (s < DP && q < P && p < P) ? v[(p * P + q) * DP + s] : 0.0;
  }
  __syncthreads();

  f64x4 y[NT][NT][DP];
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int s = 0; s < DP; ++s) y[i][j][s] = (f64x4){0.0, 0.0, 0.0, 0.0};

  for (int64_t al = a0; al < a1; ++al) {
    int t = 0;
    while (t < plan.n_terms) {
      const int a = plan.ta[t];
      const double* Ap = plan.A[a] + al * SLAB;
      const int t_begin = t;
      int t_end = t;
      while (t_end < plan.n_terms && plan.ta[t_end] == a) ++t_end;
#pragma unroll
      for (int np = 0; np < NT; ++np) {
        // stage A for the p' tile np: w[m] = V^T[16m.., :] A_a^T[:, 16np..]
        f64x4 w[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) w[m] = (f64x4){0.0, 0.0, 0.0, 0.0};
        // the V^T operands are read from LDS at each use (the empty asm hides that the address is the
        // same every time): keeping all of them in registers would leave too few for operand loads
        int vo = lane;
        __asm__ volatile("" : "+v"(vo));
#pragma unroll
        for (int u = 0; u < KS; ++u) {
          double bop = Ap[(u * NT + np) * 64 + lane];
#pragma unroll
          for (int m = 0; m < MT; ++m)
            w[m] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds[(m * KS + u) * 64 + vo], bop, w[m], 0, 0, 0);
        }
        // stage B for every term of channel a, p' tile np
        for (int tt = t_begin; tt < t_end; ++tt) {
          const double* Bp = plan.B[plan.tb[tt]] + al * SLAB;
          const int o = plan.to[tt];
          double oc[DP][DP];
#pragma unroll
          for (int i = 0; i < DP; ++i)
#pragma unroll
            for (int j = 0; j < DP; ++j) oc[i][j] = (o < 0) ? (i == j ? 1.0 : 0.0) : plan.ops[o < 0 ? 0 : o][i * DP + j];
#pragma unroll
          for (int u = 0; u < KS; ++u) {
            double xs[DP];
#pragma unroll
            for (int s = 0; s < DP; ++s) {
              const int idx = s * P4 + 4 * u;  // the lane's row within the group is idx + lane/16
              xs[s] = w[idx / 16][(idx % 16) / 4];
            }
            double x[DP];
#pragma unroll
            for (int sp = 0; sp < DP; ++sp) {
              double acc = 0.0;
#pragma unroll
              for (int s = 0; s < DP; ++s) acc += oc[sp][s] * xs[s];
              x[sp] = acc;
            }
#pragma unroll
            for (int nq = 0; nq < NT; ++nq) {
              double bop = Bp[(u * NT + nq) * 64 + lane];
#pragma unroll
              for (int sp = 0; sp < DP; ++sp)
                y[np][nq][sp] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[sp], bop, y[np][nq][sp], 0, 0, 0);
            }
          }
        }
      }
      t = t_end;
    }
  }

  // workgroup reduction in a fixed order: wave w writes slot w, then every thread sums slots 0..3
  __syncthreads();
  double* red0 = lds;
#pragma unroll
  for (int np = 0; np < NT; ++np)
#pragma unroll
    for (int nq = 0; nq < NT; ++nq)
#pragma unroll
      for (int s = 0; s < DP; ++s)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int pp = 16 * np + (lane >> 4) + 4 * r, qq = 16 * nq + (lane & 15);
          if (pp < P && qq < P) red0[wave * NY + (pp * P + qq) * DP + s] = y[np][nq][s][r];
        }
  __syncthreads();
  for (int i = threadIdx.x; i < NY; i += 256)
    partial[(int64_t)blockIdx.x * NY + i] = ((red0[i] + red0[NY + i]) + red0[2 * NY + i]) + red0[3 * NY + i];
}

// y[i] = sum_g partial[g][i] in a fixed order: block = 16 consecutive outputs x 16 group slices;
// slice k sums groups k, k+16, ...; the sixteen slice sums are combined in order 0..15.
__global__ void __launch_bounds__(256) reduce_partials_kernel(int64_t groups, int64_t n, const double* __restrict__ partial,
                                                              double* __restrict__ y) {
  __shared__ double red[16][16];
  const int e = threadIdx.x & 15, slice = threadIdx.x >> 4;
  const int64_t i = (int64_t)blockIdx.x * 16 + e;
  double acc = 0.0;
  if (i < n)
    for (int64_t g = slice; g < groups; g += 16) acc += partial[g * n + i];
  red[slice][e] = acc;
  __syncthreads();
  if (slice == 0 && i < n) {
    double t = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += red[k][e];
    y[i] = t;
  }
}

// ---- the same matvec with FP32 operands on v_mfma_f32_16x16x4f32 (FP32 accumulation) ----------------------
// The f32 result layout differs from the f64 one: D[i][j] sits in lane j + 16(i/4), register i%4.  So that the
// stage-A result is again the stage-B operand without lane exchanges, the stage-A rows interleave the physical
// index, idx = 2q + s (d = 2): a lane group's four rows of a 16-row tile are q in {8m + 2g, 8m + 2g + 1} times
// s in {0, 1}.  Stage B's K step u then takes q = 8(u/2) + 2g + u%2 in lane group g (registers 2(u%2) + s of
// tile u/2), and the B slabs are packed in that K order; the A slabs keep the f64 packing (K = p = 4u + g).
template <int P>
struct Shape32 {
  static constexpr int NT = (P + 15) / 16;      // 16-wide tiles over p' and q'
  static constexpr int KSA = (P + 3) / 4;       // stage-A K steps over p
  static constexpr int Q8 = (P + 7) / 8 * 8;    // q padded to whole 16-row tiles of (q, s)
  static constexpr int MT = Q8 / 8;             // stage-A row tiles
  static constexpr int KSB = Q8 / 4;            // stage-B K steps over q
  static constexpr int SLAB_A = KSA * NT * 64;  // packed floats per alpha
  static constexpr int SLAB_B = KSB * NT * 64;
};
__host__ __device__ inline int f32_b_col(int u, int g) { return 8 * (u >> 1) + 2 * g + (u & 1); }

// role 0 (A slabs): packed[alpha][u][n][lane] = S[16n + lane%16][4u + lane/16]; role 1 (B slabs):
// packed[alpha][u][n][lane] = S[16n + lane%16][f32_b_col(u, lane/16)]; zero outside the slab
__global__ void __launch_bounds__(256) pack_f32_kernel(int64_t n_alpha, int role, int NT, int KS, CCSlabGather g,
                                                       const double* __restrict__ src, float* __restrict__ dst) {
  extern __shared__ double s[];  // P x (P + 1)
  const int P = g.P, ld = P + 1;
  const int per = KS * NT * 64;
  for (int64_t al = blockIdx.x; al < n_alpha; al += gridDim.x) {
    cc_slab_load(g, src, al, s, ld, 1.0);
    __syncthreads();
    float* out = dst + al * per;
    for (int r = threadIdx.x; r < per; r += blockDim.x) {
      const int lane = r & 63, t = r >> 6, n = t % NT, u = t / NT;
      const int row = 16 * n + (lane & 15), col = role == 0 ? 4 * u + (lane >> 4) : f32_b_col(u, lane >> 4);
      out[r] = (row < P && col < P) ? (float)s[row * ld + col] : 0.0f;
    }
    __syncthreads();
  }
}

template <int P>
__global__ void __launch_bounds__(256) fused_f32_kernel(const FusedPlan plan, const double* __restrict__ v,
                                                        double* __restrict__ partial) {
  using S = Shape32<P>;
  constexpr int NT = S::NT, KSA = S::KSA, MT = S::MT, KSB = S::KSB, SLAB_A = S::SLAB_A, SLAB_B = S::SLAB_B;
  constexpr int DP = 2, NY = P * P * DP, NVT = MT * KSA * 64;
  constexpr int NLDS = ((NVT + 1) / 2 > 4 * NY) ? (NVT + 1) / 2 : 4 * NY;
  __shared__ double lds[NLDS];
  float* ldsv = reinterpret_cast<float*>(lds);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t gw = (int64_t)blockIdx.x * 4 + wave, nw = (int64_t)gridDim.x * 4;
  const int64_t a0 = plan.n_alpha * gw / nw, a1 = plan.n_alpha * (gw + 1) / nw;
  // stage-A A operands V^T[idx][p], idx = 2q + s = 16m + lane%16, p = 4u + lane/16, in lane order
  for (int e = threadIdx.x; e < NVT; e += 256) {
    const int l = e & 63, mu = e >> 6, u = mu % KSA, m = mu / KSA;
    const int idx = 16 * m + (l & 15), q = idx >> 1, sv = idx & 1, p = 4 * u + (l >> 4);
    ldsv[e] = (q < P && p < P) ? (float)v[(p * P + q) * DP + sv] : 0.0f;
  }
  __syncthreads();

  f32x4 y[NT][NT][DP];
#pragma unroll
  for (int i = 0; i < NT; ++i)
#pragma unroll
    for (int j = 0; j < NT; ++j)
#pragma unroll
      for (int sv = 0; sv < DP; ++sv) y[i][j][sv] = (f32x4){0.f, 0.f, 0.f, 0.f};

  for (int64_t al = a0; al < a1; ++al) {
    int t = 0;
    while (t < plan.n_terms) {
      const int a = plan.ta[t];
      const float* Ap = reinterpret_cast<const float*>(plan.A[a]) + al * SLAB_A;
      const int t_begin = t;
      int t_end = t;
      while (t_end < plan.n_terms && plan.ta[t_end] == a) ++t_end;
#pragma unroll
      for (int np = 0; np < NT; ++np) {
        f32x4 w[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) w[m] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int u = 0; u < KSA; ++u) {
          const float bop = Ap[(u * NT + np) * 64 + lane];
#pragma unroll
          for (int m = 0; m < MT; ++m)
            w[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(ldsv[(m * KSA + u) * 64 + lane], bop, w[m], 0, 0, 0);
        }
        for (int tt = t_begin; tt < t_end; ++tt) {
          const float* Bp = reinterpret_cast<const float*>(plan.B[plan.tb[tt]]) + al * SLAB_B;
          const int o = plan.to[tt];
          float oc[DP][DP];
#pragma unroll
          for (int i = 0; i < DP; ++i)
#pragma unroll
            for (int j = 0; j < DP; ++j) oc[i][j] = (o < 0) ? (i == j ? 1.f : 0.f) : (float)plan.ops[o < 0 ? 0 : o][i * DP + j];
#pragma unroll
          for (int u = 0; u < KSB; ++u) {
            float xs[DP], x[DP];
#pragma unroll
            for (int sv = 0; sv < DP; ++sv) xs[sv] = w[u >> 1][2 * (u & 1) + sv];
#pragma unroll
            for (int sp = 0; sp < DP; ++sp) {
              float acc = 0.f;
#pragma unroll
              for (int sv = 0; sv < DP; ++sv) acc += oc[sp][sv] * xs[sv];
              x[sp] = acc;
            }
#pragma unroll
            for (int nq = 0; nq < NT; ++nq) {
              const float bop = Bp[(u * NT + nq) * 64 + lane];
#pragma unroll
              for (int sp = 0; sp < DP; ++sp)
                y[np][nq][sp] = __builtin_amdgcn_mfma_f32_16x16x4f32(x[sp], bop, y[np][nq][sp], 0, 0, 0);
            }
          }
        }
      }
      t = t_end;
    }
  }
  // workgroup reduction in FP64 in a fixed order (rows p' = 16 np + 4 (lane/16) + r in the f32 layout)
  __syncthreads();
#pragma unroll
  for (int np = 0; np < NT; ++np)
#pragma unroll
    for (int nq = 0; nq < NT; ++nq)
#pragma unroll
      for (int sv = 0; sv < DP; ++sv)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int pp = 16 * np + 4 * (lane >> 4) + r, qq = 16 * nq + (lane & 15);
          if (pp < P && qq < P) lds[wave * NY + (pp * P + qq) * DP + sv] = (double)y[np][nq][sv][r];
        }
  __syncthreads();
  for (int i = threadIdx.x; i < NY; i += 256)
    partial[(int64_t)blockIdx.x * NY + i] = ((lds[i] + lds[NY + i]) + lds[2 * NY + i]) + lds[3 * NY + i];
}

template <int P>
int launch_fused_f32(hipStream_t st, const FusedPlan& plan, const double* v, double* y, double* work, int groups) {
  fused_f32_kernel<P><<<groups, 256, 0, st>>>(plan, v, work);
  CC_LAUNCH_CHECK();
  const int64_t n = (int64_t)P * P * 2;
  reduce_partials_kernel<<<(unsigned)((n + 15) / 16), 256, 0, st>>>(groups, n, work, y);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

template <int P, int DP>
int launch_fused(hipStream_t st, const FusedPlan& plan, const double* v, double* y, double* work, int groups,
                 bool profile = true) {
  hipEvent_t e0 = nullptr;
  int rc = profile ? cc_prof_start(st, CC_PROF_FUSED_F64, &e0) : CC_OK;
  if (rc != CC_OK) return rc;
  fused_f64_kernel<P, DP><<<groups, 256, 0, st>>>(plan, v, work);
  CC_LAUNCH_CHECK();
  rc = profile ? cc_prof_stop(st, CC_PROF_FUSED_F64, e0) : CC_OK;
  if (rc != CC_OK) return rc;
  const int64_t n = (int64_t)P * P * DP;
  reduce_partials_kernel<<<(unsigned)((n + 15) / 16), 256, 0, st>>>(groups, n, work, y);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

int fused_groups(int64_t n_alpha) {
  // four alpha per wave at the largest shapes; never more workgroups than waves of work
  int64_t g = (n_alpha + 15) / 16;
  if (g > 1024) g = 1024;
  if (g < 1) g = 1;
  return (int)g;
}
}  // namespace

namespace {
// A persistent operator: the plan is validated and copied once, then applied many times.
struct Stage3PlanHandle {
  FusedPlan plan;
  int64_t P, d;
  int groups;
  int f32;        // 1: the slabs are FP32 packings (cc_stage3_pack_strided_f32) and the FP32 kernel runs
  double* work;
};

int apply_plan(hipStream_t st, const Stage3PlanHandle* h, const double* v, double* y) {
  if (h->f32) {
    switch (h->P) {
      case 4: return launch_fused_f32<4>(st, h->plan, v, y, h->work, h->groups);
      case 9: return launch_fused_f32<9>(st, h->plan, v, y, h->work, h->groups);
      case 16: return launch_fused_f32<16>(st, h->plan, v, y, h->work, h->groups);
      case 25: return launch_fused_f32<25>(st, h->plan, v, y, h->work, h->groups);
      case 36: return launch_fused_f32<36>(st, h->plan, v, y, h->work, h->groups);
    }
    cc_set_last_error("fused stage-3: shape not instantiated");
    return CC_ERR_UNSUPPORTED;
  }
  switch (h->P) {
    case 4: return launch_fused<4, 2>(st, h->plan, v, y, h->work, h->groups);
    case 9: return launch_fused<9, 2>(st, h->plan, v, y, h->work, h->groups);
    case 16: return launch_fused<16, 2>(st, h->plan, v, y, h->work, h->groups);
    case 25: return launch_fused<25, 2>(st, h->plan, v, y, h->work, h->groups);
    case 36: return launch_fused<36, 2>(st, h->plan, v, y, h->work, h->groups);
  }
  cc_set_last_error("fused stage-3: shape not instantiated");
  return CC_ERR_UNSUPPORTED;
}

int fill_plan(FusedPlan& plan, int n_a, const double* const* A_packed, int n_b, const double* const* B_packed,
              int n_terms, const int* term_a, const int* term_b, const int* term_o, int n_ops, const double* host_ops,
              int64_t n_alpha, int64_t d) {
  if (n_a < 1 || n_a > MAX_CH || n_b < 1 || n_b > MAX_CH || n_terms < 1 || n_terms > MAX_TERMS || n_ops < 0 ||
      n_ops > MAX_OPS || n_alpha < 1 || d < 1 || d > MAX_D) {
    cc_set_last_error("fused stage-3: plan exceeds kernel limits");
    return CC_ERR_UNSUPPORTED;
  }
  plan = FusedPlan{};
  for (int i = 0; i < n_a; ++i) plan.A[i] = A_packed[i];
  for (int i = 0; i < n_b; ++i) plan.B[i] = B_packed[i];
  plan.n_terms = n_terms;
  for (int t = 0; t < n_terms; ++t) {
    if (term_a[t] < 0 || term_a[t] >= n_a || term_b[t] < 0 || term_b[t] >= n_b || term_o[t] < -1 || term_o[t] >= n_ops) {
      cc_set_last_error("fused stage-3: term index out of range");
      return CC_ERR_ARG;
    }
    if (t > 0 && term_a[t] < term_a[t - 1]) { cc_set_last_error("fused stage-3: terms must be sorted by A channel"); return CC_ERR_ARG; }
    plan.ta[t] = term_a[t];
    plan.tb[t] = term_b[t];
    plan.to[t] = term_o[t];
  }
  for (int o = 0; o < n_ops; ++o)
    for (int e = 0; e < d * d; ++e) plan.ops[o][e] = host_ops[o * d * d + e];
  plan.n_alpha = n_alpha;
  return CC_OK;
}
}  // namespace

int cc_internal_stage3_apply(hipStream_t st, const void* plan, const double* x, double* y) {
  return apply_plan(st, static_cast<const Stage3PlanHandle*>(plan), x, y);
}
int64_t cc_internal_stage3_n(const void* plan) {
  auto h = static_cast<const Stage3PlanHandle*>(plan);
  return h->P * h->P * h->d;
}

extern "C" int cc_stage3_fused_supported(int64_t P, int64_t d) {
  return (d == 2 && (P == 4 || P == 9 || P == 16 || P == 25 || P == 36)) ? 1 : 0;
}

extern "C" int cc_stage3_packed_size(int64_t n_alpha, int64_t P, int64_t* out) {
  int64_t NT = (P + 15) / 16, KS = (P + 3) / 4;
  *out = n_alpha * KS * NT * 64;
  return CC_OK;
}

extern "C" int cc_stage3_fused_work_size(int64_t n_alpha, int64_t P, int64_t d, int64_t* out) {
  *out = (int64_t)fused_groups(n_alpha) * P * P * d;
  return CC_OK;
}

extern "C" int cc_stage3_pack_strided_f64(void* stream, int64_t na0, int64_t na1, int64_t d0, int64_t d1,
                                          const int64_t* strides, const double* src, double* packed) {
  CCSlabMap map;
  int rc = cc_slab_map(na0, na1, d0, d1, strides, &map);
  if (rc) return rc;
  const int64_t n_alpha = na0 * na1, P = d0 * d1;
  if (P > PACK_MAX_P) { cc_set_last_error("pack: slab extent above 64"); return CC_ERR_UNSUPPORTED; }
  int NT = (int)((P + 15) / 16), KS = (int)((P + 3) / 4);
  const unsigned grid = (unsigned)(n_alpha < 16384 ? n_alpha : 16384);
  const size_t lds = (size_t)P * (P + 1) * sizeof(double);
  const CCSlabGather g = cc_slab_gather(map, d0);
  hipStream_t st = cc_stream(stream);
  if (d0 == map.d1 && d0 >= 2 && d0 <= 6) {
    switch (d0) {
      case 2: pack_kernel_d<2><<<grid, 256, 0, st>>>(n_alpha, g, src, packed); break;
      case 3: pack_kernel_d<3><<<grid, 256, 0, st>>>(n_alpha, g, src, packed); break;
      case 4: pack_kernel_d<4><<<grid, 256, 0, st>>>(n_alpha, g, src, packed); break;
      case 5: pack_kernel_d<5><<<grid, 256, 0, st>>>(n_alpha, g, src, packed); break;
      default: pack_kernel_d<6><<<grid, 256, 0, st>>>(n_alpha, g, src, packed); break;
    }
  } else {
    pack_kernel<<<grid, 256, lds, st>>>(n_alpha, NT, KS, g, src, packed);
  }
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_stage3_pack_f64(void* stream, int64_t n_alpha, int64_t P, const double* slabs, double* packed) {
  if (n_alpha < 0 || P <= 0) { cc_set_last_error("pack: bad sizes"); return CC_ERR_ARG; }
  if (n_alpha == 0) return CC_OK;
  const int64_t st[6] = {P * P, 0, P, 0, 1, 0};
  return cc_stage3_pack_strided_f64(stream, n_alpha, 1, P, 1, st, slabs, packed);
}

extern "C" int cc_stage3_matvec_fused_f64(void* stream, int n_a, const double* const* A_packed, int n_b,
                                          const double* const* B_packed, int n_terms, const int* term_a,
                                          const int* term_b, const int* term_o, int n_ops, const double* host_ops,
                                          int64_t n_alpha, int64_t P, int64_t d, const double* v, double* y,
                                          double* work) {
  if (!cc_stage3_fused_supported(P, d)) { cc_set_last_error("fused stage-3: shape not instantiated"); return CC_ERR_UNSUPPORTED; }
  FusedPlan plan;
  int rc = fill_plan(plan, n_a, A_packed, n_b, B_packed, n_terms, term_a, term_b, term_o, n_ops, host_ops, n_alpha, d);
  if (rc) return rc;
  hipStream_t st = cc_stream(stream);
  int groups = fused_groups(n_alpha);
  switch (P) {
    case 4: return launch_fused<4, 2>(st, plan, v, y, work, groups);
    case 9: return launch_fused<9, 2>(st, plan, v, y, work, groups);
    case 16: return launch_fused<16, 2>(st, plan, v, y, work, groups);
    case 25: return launch_fused<25, 2>(st, plan, v, y, work, groups);
    case 36: return launch_fused<36, 2>(st, plan, v, y, work, groups);
  }
  return CC_ERR_UNSUPPORTED;
}

// ---- the allocation-free entry point -----------------------------------------------------------
// Scratch of cc_stage3_matvec_fused_f64_noalloc: the workgroup partial sums (groups x P^2 d doubles).
// Every GPU the library runs on holds its own copy until the process ends.
static __device__ __align__(256) double cc_noalloc_scratch[CC_NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kNoallocMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_scratch[kNoallocMaxDevices] = {};

// The device address of cc_noalloc_scratch on the current device, looked up once per device (the host-side
// name is only a stand-in for the symbol and is never passed to a kernel or a copy).
int noalloc_scratch(double** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) { cc_set_last_error(hipGetErrorString(e)); return CC_ERR_HIP; }
  if (dev < 0 || dev >= kNoallocMaxDevices) { cc_set_last_error("noalloc: device index above 63"); return CC_ERR_UNSUPPORTED; }
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(cc_noalloc_scratch));
    if (e != hipSuccess || !p) {
      cc_set_last_error(e != hipSuccess ? hipGetErrorString(e) : "noalloc: null scratch address");
      return CC_ERR_HIP;
    }
    g_noalloc_scratch[dev] = static_cast<double*>(p);
  }
  *out = g_noalloc_scratch[dev];
  return CC_OK;
}
}  // namespace

extern "C" int cc_stage3_matvec_fused_f64_noalloc(void* stream, int n_a, const double* const* A_packed, int n_b,
                                                  const double* const* B_packed, int n_terms, const int* term_a,
                                                  const int* term_b, const int* term_o, int n_ops,
                                                  const double* host_ops, int64_t n_alpha, int64_t P, int64_t d,
                                                  const double* v, double* y) {
  if (!cc_stage3_fused_supported(P, d)) { cc_set_last_error("fused stage-3: shape not instantiated"); return CC_ERR_UNSUPPORTED; }
  FusedPlan plan;
  int rc = fill_plan(plan, n_a, A_packed, n_b, B_packed, n_terms, term_a, term_b, term_o, n_ops, host_ops, n_alpha, d);
  if (rc) return rc;
  const int groups = fused_groups(n_alpha);
  if ((int64_t)groups * P * P * d > (int64_t)CC_NOALLOC_SCRATCH_DOUBLES) {
    cc_set_last_error("noalloc stage-3: the partial sums need more than the 1048576-double (8 MiB) scratch; "
                      "use cc_stage3_matvec_fused_f64 with a workspace");
    return CC_ERR_ARG;
  }
  double* scratch = nullptr;
  rc = noalloc_scratch(&scratch);
  if (rc) return rc;
  hipStream_t st = cc_stream(stream);
  switch (P) {
    case 4: return launch_fused<4, 2>(st, plan, v, y, scratch, groups, false);
    case 9: return launch_fused<9, 2>(st, plan, v, y, scratch, groups, false);
    case 16: return launch_fused<16, 2>(st, plan, v, y, scratch, groups, false);
    case 25: return launch_fused<25, 2>(st, plan, v, y, scratch, groups, false);
    case 36: return launch_fused<36, 2>(st, plan, v, y, scratch, groups, false);
  }
  return CC_ERR_UNSUPPORTED;
}

extern "C" int cc_stage3_plan_create(void** out, int n_a, const double* const* A_packed, int n_b,
                                     const double* const* B_packed, int n_terms, const int* term_a, const int* term_b,
                                     const int* term_o, int n_ops, const double* host_ops, int64_t n_alpha, int64_t P,
                                     int64_t d, double* work) {
  *out = nullptr;
  if (!cc_stage3_fused_supported(P, d)) { cc_set_last_error("fused stage-3: shape not instantiated"); return CC_ERR_UNSUPPORTED; }
  Stage3PlanHandle* h = new Stage3PlanHandle();
  int rc = fill_plan(h->plan, n_a, A_packed, n_b, B_packed, n_terms, term_a, term_b, term_o, n_ops, host_ops, n_alpha, d);
  if (rc) { delete h; return rc; }
  h->P = P;
  h->d = d;
  h->groups = fused_groups(n_alpha);
  h->f32 = 0;
  h->work = work;
  *out = h;
  return CC_OK;
}

extern "C" int cc_stage3_packed_size_f32(int64_t n_alpha, int64_t P, int role, int64_t* out) {
  if (n_alpha < 0 || P < 1 || (role != 0 && role != 1)) { cc_set_last_error("pack f32: bad arguments"); return CC_ERR_ARG; }
  const int64_t NT = (P + 15) / 16, KS = role == 0 ? (P + 3) / 4 : (P + 7) / 8 * 2;
  *out = n_alpha * KS * NT * 64;
  return CC_OK;
}

extern "C" int cc_stage3_pack_strided_f32(void* stream, int role, int64_t na0, int64_t na1, int64_t d0, int64_t d1,
                                          const int64_t* strides, const double* src, float* packed) {
  if (role != 0 && role != 1) { cc_set_last_error("pack f32: role must be 0 (A) or 1 (B)"); return CC_ERR_ARG; }
  CCSlabMap map;
  int rc = cc_slab_map(na0, na1, d0, d1, strides, &map);
  if (rc) return rc;
  const int64_t n_alpha = na0 * na1, P = d0 * d1;
  if (P > PACK_MAX_P) { cc_set_last_error("pack f32: slab extent above 64"); return CC_ERR_UNSUPPORTED; }
  if (n_alpha == 0) return CC_OK;
  const int NT = (int)((P + 15) / 16), KS = role == 0 ? (int)((P + 3) / 4) : (int)((P + 7) / 8 * 2);
  const unsigned grid = (unsigned)(n_alpha < 16384 ? n_alpha : 16384);
  const size_t lds = (size_t)P * (P + 1) * sizeof(double);
  pack_f32_kernel<<<grid, 256, lds, cc_stream(stream)>>>(n_alpha, role, NT, KS, cc_slab_gather(map, d0), src, packed);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_stage3_plan_create_f32(void** out, int n_a, const float* const* A_packed, int n_b,
                                         const float* const* B_packed, int n_terms, const int* term_a, const int* term_b,
                                         const int* term_o, int n_ops, const double* host_ops, int64_t n_alpha, int64_t P,
                                         int64_t d, double* work) {
  *out = nullptr;
  if (!cc_stage3_fused_supported(P, d)) { cc_set_last_error("fused stage-3: shape not instantiated"); return CC_ERR_UNSUPPORTED; }
  if (n_a < 1 || n_a > MAX_CH || n_b < 1 || n_b > MAX_CH) { cc_set_last_error("fused stage-3: plan exceeds kernel limits"); return CC_ERR_UNSUPPORTED; }
  std::vector<const double*> A(n_a), B(n_b);   // the plan stores slab addresses; the FP32 kernel reads them as float
  for (int i = 0; i < n_a; ++i) A[i] = reinterpret_cast<const double*>(A_packed[i]);
  for (int i = 0; i < n_b; ++i) B[i] = reinterpret_cast<const double*>(B_packed[i]);
  Stage3PlanHandle* h = new Stage3PlanHandle();
  int rc = fill_plan(h->plan, n_a, A.data(), n_b, B.data(), n_terms, term_a, term_b, term_o, n_ops, host_ops, n_alpha, d);
  if (rc) { delete h; return rc; }
  h->P = P;
  h->d = d;
  h->groups = fused_groups(n_alpha);
  h->f32 = 1;
  h->work = work;
  *out = h;
  return CC_OK;
}

extern "C" int cc_stage3_plan_apply(void* stream, const void* plan, const double* x, double* y) {
  if (!plan) { cc_set_last_error("null plan"); return CC_ERR_ARG; }
  return apply_plan(cc_stream(stream), static_cast<const Stage3PlanHandle*>(plan), x, y);
}

extern "C" int cc_stage3_plan_destroy(void* plan) {
  delete static_cast<Stage3PlanHandle*>(plan);
  return CC_OK;
}

