// Shared definitions for the carcassonne_amd HIP library (gfx950 / CDNA4 only).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/carcassonne_hip.h"

#define CC_HIP(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) {                            \
      cc_set_last_error(hipGetErrorString(_e));        \
      return CC_ERR_HIP;                               \
    }                                                  \
  } while (0)

#define CC_LAUNCH_CHECK()                              \
  do {                                                 \
    hipError_t _e = hipGetLastError();                 \
    if (_e != hipSuccess) {                            \
      cc_set_last_error(hipGetErrorString(_e));        \
      return CC_ERR_HIP;                               \
    }                                                  \
  } while (0)

void cc_set_last_error(const char* msg);

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

static inline hipStream_t cc_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

static inline unsigned cc_blocks(int64_t n, int threads, unsigned cap = 65535u * 8u) {
  int64_t b = (n + threads - 1) / threads;
  if (b < 1) b = 1;
  if (b > (int64_t)cap) b = cap;
  return (unsigned)b;
}

// Internal (not exported through the C ABI header): operator plans used by the solvers.
int cc_internal_stage3_apply(hipStream_t st, const void* plan, const double* x, double* y);
int64_t cc_internal_stage3_n(const void* plan);
int cc_internal_lp_apply(hipStream_t st, const void* plan, const double* x, double* y);
int64_t cc_internal_lp_n(const void* plan);

// Launch timing (between cc_profile_begin[_sampled] and cc_profile_end, single-threaded use): a kernel
// launch is bracketed by cc_prof_start / cc_prof_stop, which record HIP events on its stream when timing
// is on and the launch is one of the sampled ones of its kind (e0 stays null otherwise).
enum { CC_PROF_FUSED_F64 = 0, CC_PROF_LP = 1, CC_PROF_KINDS = 2 };
int cc_prof_start(hipStream_t st, int kind, hipEvent_t* e0);
int cc_prof_stop(hipStream_t st, int kind, hipEvent_t e0);

// A stage-3 slab channel read in place from a tensor with six index legs: slab[alpha][p'][p] with
// alpha = a0 * na1 + a1, p' = r0 * d1 + r1, p = c0 * d1 + c1 (P = d0 * d1) lives at element offset
// a0 s[0] + a1 s[1] + r0 s[2] + r1 s[3] + c0 s[4] + c1 s[5] of the source.  A contiguous
// (alpha, P, P) array is na1 = d1 = 1, s = {P*P, 0, P, 0, 1, 0}.
struct CCSlabMap {
  int64_t na1, d1;
  int64_t s[6];
};
__host__ __device__ inline int64_t cc_slab_offset(const CCSlabMap& m, int64_t al, int64_t row, int64_t col) {
  const int64_t a0 = al / m.na1, a1 = al - a0 * m.na1;
  const int64_t r0 = row / m.d1, r1 = row - r0 * m.d1, c0 = col / m.d1, c1 = col - c0 * m.d1;
  return a0 * m.s[0] + a1 * m.s[1] + r0 * m.s[2] + r1 * m.s[3] + c0 * m.s[4] + c1 * m.s[5];
}
// Validates the leg sizes (all positive, d0 * d1 = P, na0 * na1 = n_alpha) and non-negative strides.
int cc_slab_map(int64_t na0, int64_t na1, int64_t d0, int64_t d1, const int64_t* strides, CCSlabMap* out);

// Per-alpha gather of one slab into LDS in source order: the four inner legs (r0, r1, c0, c1) sorted by
// stride, largest first, so consecutive threads read consecutive source elements when some leg has
// stride 1.  Element k of the gather has digits k = ((k0 * sz1 + k1) * sz2 + k2) * sz3 + k3; its source
// offset is sum k_i st_i, its slab position row = sum k_i rw_i, col = sum k_i cw_i.
struct CCSlabGather {
  int64_t na1, sa0, sa1;  // alpha = a0 * na1 + a1 at a0 * sa0 + a1 * sa1
  int P;
  int sz[4], rw[4], cw[4];
  int64_t st[4];
};
static inline CCSlabGather cc_slab_gather(const CCSlabMap& m, int64_t d0) {
  CCSlabGather g;
  g.na1 = m.na1;
  g.sa0 = m.s[0];
  g.sa1 = m.s[1];
  g.P = (int)(d0 * m.d1);
  const int size[4] = {(int)d0, (int)m.d1, (int)d0, (int)m.d1};
  const int rw[4] = {(int)m.d1, 1, 0, 0}, cw[4] = {0, 0, (int)m.d1, 1};
  int order[4] = {0, 1, 2, 3};
  for (int i = 1; i < 4; ++i)  // stable insertion sort by stride, descending
    for (int j = i; j > 0 && m.s[2 + order[j]] > m.s[2 + order[j - 1]]; --j) {
      int t = order[j];
      order[j] = order[j - 1];
      order[j - 1] = t;
    }
  for (int i = 0; i < 4; ++i) {
    g.sz[i] = size[order[i]];
    g.st[i] = m.s[2 + order[i]];
    g.rw[i] = rw[order[i]];
    g.cw[i] = cw[order[i]];
  }
  return g;
}
// The same with every leg of extent D (P = D * D) and NTH threads: the index arithmetic divides by
// constants, and each thread issues all its loads before its LDS stores.
template <int D, int NTH>
__device__ inline void cc_slab_load_d(const CCSlabGather& g, const double* __restrict__ src, int64_t al, double* lds, int ld,
                                      double mul) {
  constexpr int n = D * D * D * D, NE = (n + NTH - 1) / NTH;
  const int64_t a0 = al / g.na1, a1 = al - a0 * g.na1;
  const double* base = src + a0 * g.sa0 + a1 * g.sa1;
  double v[NE];
  int at[NE];
#pragma unroll
  for (int u = 0; u < NE; ++u) {
    const int e = threadIdx.x + NTH * u, ec = e < n ? e : 0;  // clamped: every load in bounds
    const int k3 = ec % D, k2 = (ec / D) % D, k1 = (ec / (D * D)) % D, k0 = ec / (D * D * D);
    v[u] = base[k0 * g.st[0] + k1 * g.st[1] + k2 * g.st[2] + k3 * g.st[3]];
    at[u] = (k0 * g.rw[0] + k1 * g.rw[1] + k2 * g.rw[2] + k3 * g.rw[3]) * ld + k0 * g.cw[0] + k1 * g.cw[1] + k2 * g.cw[2] +
            k3 * g.cw[3];
  }
#pragma unroll
  for (int u = 0; u < NE; ++u)
    if (threadIdx.x + NTH * u < n) lds[at[u]] = v[u] * mul;
}
// Loads slab alpha into lds[row * ld + col] (row, col < P) with all threads of the block.
__device__ inline void cc_slab_load(const CCSlabGather& g, const double* __restrict__ src, int64_t al, double* lds, int ld,
                                    double mul) {
  const int64_t a0 = al / g.na1, a1 = al - a0 * g.na1;
  const double* base = src + a0 * g.sa0 + a1 * g.sa1;
  const int n = g.P * g.P;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    int k = e;
    const int k3 = k % g.sz[3];
    k /= g.sz[3];
    const int k2 = k % g.sz[2];
    k /= g.sz[2];
    const int k1 = k % g.sz[1];
    const int k0 = k / g.sz[1];
    const int64_t off = k0 * g.st[0] + k1 * g.st[1] + k2 * g.st[2] + k3 * g.st[3];
    const int row = k0 * g.rw[0] + k1 * g.rw[1] + k2 * g.rw[2] + k3 * g.rw[3];
    const int col = k0 * g.cw[0] + k1 * g.cw[1] + k2 * g.cw[2] + k3 * g.cw[3];
    lds[row * ld + col] = base[off] * mul;
  }
}
