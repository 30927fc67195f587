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

// Launch timing (between cc_profile_begin and cc_profile_end, single-threaded use): a kernel launch is
// bracketed by cc_prof_start / cc_prof_stop, which record HIP events on its stream when timing is on.
enum { CC_PROF_FUSED_F64 = 0, CC_PROF_LP = 1, CC_PROF_KINDS = 2 };
int cc_prof_start(hipStream_t st, hipEvent_t* e0);
int cc_prof_stop(hipStream_t st, int kind, hipEvent_t e0);
