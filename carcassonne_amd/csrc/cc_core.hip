// Data-protocol primitives: permute, elementwise, fixed-order reductions, casts.
#include "cc_common.h"

#include <vector>

#include <cstring>
#include <string>

static thread_local std::string g_last_error;
void cc_set_last_error(const char* msg) { g_last_error = msg ? msg : ""; }

extern "C" int cc_version(void) { return 1; }

int cc_slab_map(int64_t na0, int64_t na1, int64_t d0, int64_t d1, const int64_t* strides, CCSlabMap* out) {
  if (na0 < 1 || na1 < 1 || d0 < 1 || d1 < 1 || strides == nullptr) {
    cc_set_last_error("slab map: leg sizes must be positive");
    return CC_ERR_ARG;
  }
  out->na1 = na1;
  out->d1 = d1;
  for (int i = 0; i < 6; ++i) {
    if (strides[i] < 0) { cc_set_last_error("slab map: strides must be non-negative"); return CC_ERR_ARG; }
    out->s[i] = strides[i];
  }
  return CC_OK;
}
extern "C" const char* cc_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------- permute
#define CC_MAX_RANK 16
struct PermDesc {
  int rank;
  int64_t shape[CC_MAX_RANK];
  int64_t stride[CC_MAX_RANK];
};

template <typename T>
__global__ void permute_kernel(const T* __restrict__ src, T* __restrict__ dst, int64_t total, PermDesc d) {
  for (int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total;
       o += (int64_t)gridDim.x * blockDim.x) {
    int64_t rem = o, off = 0;
    for (int k = d.rank - 1; k >= 0; --k) {
      int64_t i = rem % d.shape[k];
      rem /= d.shape[k];
      off += i * d.stride[k];
    }
    dst[o] = src[off];
  }
}

// The same for fewer than 2^32 output elements, the output index split into coordinates with 32-bit
// divisions by the invariant extents done as a multiply-high and shifts (Granlund and Montgomery,
// "Division by invariant integers using multiplication", 1994: for 1 < d < 2^32, l = ceil(log2 d),
// m = floor(2^32 (2^l - d) / d) + 1, x / d = (t + ((x - t) >> 1)) >> (l - 1) with t = mulhi(m, x)).
struct PermDesc32 {
  int rank;
  uint32_t shape[CC_MAX_RANK], magic[CC_MAX_RANK], shift[CC_MAX_RANK];
  int64_t stride[CC_MAX_RANK];
};
__device__ __forceinline__ uint32_t perm_div(uint32_t x, uint32_t d, uint32_t m, uint32_t l) {
  if (d == 1) return x;
  const uint32_t t = __umulhi(m, x);
  return (t + ((x - t) >> 1)) >> (l - 1);
}
template <typename T>
__global__ void permute32_kernel(const T* __restrict__ src, T* __restrict__ dst, uint32_t total, PermDesc32 d) {
  for (uint64_t o = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; o < total; o += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t rem = (uint32_t)o;
    int64_t off = 0;
    for (int k = d.rank - 1; k >= 0; --k) {
      const uint32_t q = perm_div(rem, d.shape[k], d.magic[k], d.shift[k]);
      off += (int64_t)(rem - q * d.shape[k]) * d.stride[k];
      rem = q;
    }
    dst[o] = src[off];
  }
}

static int dtype_size(int dt) {
  switch (dt) {
    case CC_F64: return 8;
    case CC_F32: return 4;
    case CC_F16: return 2;
    case CC_BF16: return 2;
  }
  return 0;
}

extern "C" int cc_permute(void* stream, int dtype, const void* src, void* dst, int rank,
                          const int64_t* shape, const int64_t* src_strides) {
  if (rank < 0 || rank > CC_MAX_RANK) { cc_set_last_error("cc_permute: rank out of range"); return CC_ERR_ARG; }
  PermDesc d;
  d.rank = rank;
  int64_t total = 1;
  for (int k = 0; k < rank; ++k) {
    if (shape[k] < 0) { cc_set_last_error("cc_permute: negative extent"); return CC_ERR_ARG; }
    d.shape[k] = shape[k];
    d.stride[k] = src_strides[k];
    total *= shape[k];
  }
  if (total == 0) return CC_OK;
  unsigned blocks = cc_blocks(total, 256);
  if (total < (int64_t)0xFFFFFFFFLL) {
    PermDesc32 d32;
    d32.rank = rank;
    for (int k = 0; k < rank; ++k) {
      const uint64_t dk = (uint64_t)shape[k];   // >= 1: a zero extent returned above
      const uint32_t l = dk <= 1 ? 0u : (uint32_t)(64 - __builtin_clzll(dk - 1));
      d32.shape[k] = (uint32_t)dk;
      d32.shift[k] = l;
      d32.magic[k] = dk <= 1 ? 0u : (uint32_t)((((1ull << l) - dk) << 32) / dk + 1);
      d32.stride[k] = src_strides[k];
    }
    const uint32_t t32 = (uint32_t)total;
    switch (dtype_size(dtype)) {
      case 8: permute32_kernel<uint64_t><<<blocks, 256, 0, cc_stream(stream)>>>((const uint64_t*)src, (uint64_t*)dst, t32, d32); break;
      case 4: permute32_kernel<uint32_t><<<blocks, 256, 0, cc_stream(stream)>>>((const uint32_t*)src, (uint32_t*)dst, t32, d32); break;
      case 2: permute32_kernel<uint16_t><<<blocks, 256, 0, cc_stream(stream)>>>((const uint16_t*)src, (uint16_t*)dst, t32, d32); break;
      default: cc_set_last_error("cc_permute: bad dtype"); return CC_ERR_ARG;
    }
    CC_LAUNCH_CHECK();
    return CC_OK;
  }
  switch (dtype_size(dtype)) {
    case 8: permute_kernel<uint64_t><<<blocks, 256, 0, cc_stream(stream)>>>((const uint64_t*)src, (uint64_t*)dst, total, d); break;
    case 4: permute_kernel<uint32_t><<<blocks, 256, 0, cc_stream(stream)>>>((const uint32_t*)src, (uint32_t*)dst, total, d); break;
    case 2: permute_kernel<uint16_t><<<blocks, 256, 0, cc_stream(stream)>>>((const uint16_t*)src, (uint16_t*)dst, total, d); break;
    default: cc_set_last_error("cc_permute: bad dtype"); return CC_ERR_ARG;
  }
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// ---------------------------------------------------------------- elementwise
__global__ void axpby_kernel(int64_t n, double a, const double* __restrict__ x, double b, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = (b == 0.0 ? 0.0 : b * y[i]) + a * x[i];
}
__global__ void fill_kernel(int64_t n, double v, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i] = v;
}
__global__ void set_identity_kernel(int64_t n, double* __restrict__ y) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) y[i * (n + 1)] = 1.0;
}

extern "C" int cc_axpby_f64(void* stream, int64_t n, double alpha, const double* x, double beta, double* y) {
  if (n <= 0) return CC_OK;
  axpby_kernel<<<cc_blocks(n, 256), 256, 0, cc_stream(stream)>>>(n, alpha, x, beta, y);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
__global__ void axpby_dev_kernel(int64_t n, const double* __restrict__ a_dev, double a_scale, const double* __restrict__ x,
                                 double beta, double* __restrict__ y) {
  const double a = a_scale * a_dev[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = a * x[i] + beta * y[i];
}
extern "C" int cc_axpby_dev_f64(void* stream, int64_t n, const double* alpha_dev, double alpha_scale, const double* x,
                                double beta, double* y) {
  if (n <= 0) return CC_OK;
  axpby_dev_kernel<<<cc_blocks(n, 256), 256, 0, cc_stream(stream)>>>(n, alpha_dev, alpha_scale, x, beta, y);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
extern "C" int cc_fill_f64(void* stream, int64_t n, double value, double* y) {
  if (n <= 0) return CC_OK;
  fill_kernel<<<cc_blocks(n, 256), 256, 0, cc_stream(stream)>>>(n, value, y);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
extern "C" int cc_set_identity_f64(void* stream, int64_t n, double* y) {
  if (n <= 0) return CC_OK;
  set_identity_kernel<<<cc_blocks(n, 256), 256, 0, cc_stream(stream)>>>(n, y);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// ---------------------------------------------------------------- reductions
// One block of 1024 threads: each thread sums a fixed stride class, then a fixed
// tree.  The summation order depends only on n, never on scheduling.
__global__ void __launch_bounds__(1024) dot_kernel(int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                                                   double* __restrict__ out) {
  __shared__ double red[1024];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) acc += x[i] * y[i];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void __launch_bounds__(1024) count_nonfinite_kernel(int64_t n, const double* __restrict__ x, int64_t* __restrict__ out) {
  __shared__ int64_t red[1024];
  int64_t acc = 0;
  for (int64_t i = threadIdx.x; i < n; i += 1024) acc += isfinite(x[i]) ? 0 : 1;
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

extern "C" int cc_dot_f64(void* stream, int64_t n, const double* x, const double* y, double* out) {
  dot_kernel<<<1, 1024, 0, cc_stream(stream)>>>(n < 0 ? 0 : n, x, y, out);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
extern "C" int cc_count_nonfinite_f64(void* stream, int64_t n, const double* x, int64_t* out) {
  count_nonfinite_kernel<<<1, 1024, 0, cc_stream(stream)>>>(n < 0 ? 0 : n, x, out);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// ---------------------------------------------------------------- casts
template <typename S>
__device__ inline double load_as_f64(const S* p, int64_t i) { return (double)p[i]; }

template <typename S, typename D>
__global__ void cast_kernel(int64_t n, const S* __restrict__ s, D* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (D)(float)(double)s[i];
}
template <typename S>
__global__ void cast_to_f64_kernel(int64_t n, const S* __restrict__ s, double* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    d[i] = (double)(float)s[i];
}
__global__ void cast_f64_f32_kernel(int64_t n, const double* __restrict__ s, float* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = (float)s[i];
}
__global__ void cast_f32_f64_kernel(int64_t n, const float* __restrict__ s, double* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = (double)s[i];
}
// f64 -> f16/bf16 rounds once (round-to-nearest-even from the double value)
__global__ void cast_f64_f16_kernel(int64_t n, const double* __restrict__ s, _Float16* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = (_Float16)s[i];
}
__global__ void cast_f64_bf16_kernel(int64_t n, const double* __restrict__ s, __bf16* __restrict__ d) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) d[i] = (__bf16)s[i];
}

extern "C" int cc_cast(void* stream, int64_t n, int src_dtype, const void* src, int dst_dtype, void* dst) {
  if (n <= 0) return CC_OK;
  hipStream_t st = cc_stream(stream);
  unsigned b = cc_blocks(n, 256);
  if (src_dtype == CC_F64 && dst_dtype == CC_F32) cast_f64_f32_kernel<<<b, 256, 0, st>>>(n, (const double*)src, (float*)dst);
  else if (src_dtype == CC_F32 && dst_dtype == CC_F64) cast_f32_f64_kernel<<<b, 256, 0, st>>>(n, (const float*)src, (double*)dst);
  else if (src_dtype == CC_F64 && dst_dtype == CC_F16) cast_f64_f16_kernel<<<b, 256, 0, st>>>(n, (const double*)src, (_Float16*)dst);
  else if (src_dtype == CC_F64 && dst_dtype == CC_BF16) cast_f64_bf16_kernel<<<b, 256, 0, st>>>(n, (const double*)src, (__bf16*)dst);
  else if (src_dtype == CC_F16 && dst_dtype == CC_F64) cast_to_f64_kernel<_Float16><<<b, 256, 0, st>>>(n, (const _Float16*)src, (double*)dst);
  else if (src_dtype == CC_BF16 && dst_dtype == CC_F64) cast_to_f64_kernel<__bf16><<<b, 256, 0, st>>>(n, (const __bf16*)src, (double*)dst);
  else if (src_dtype == CC_F32 && dst_dtype == CC_F16) cast_kernel<float, _Float16><<<b, 256, 0, st>>>(n, (const float*)src, (_Float16*)dst);
  else if (src_dtype == CC_F16 && dst_dtype == CC_F32) cast_kernel<_Float16, float><<<b, 256, 0, st>>>(n, (const _Float16*)src, (float*)dst);
  else { cc_set_last_error("cc_cast: unsupported dtype pair"); return CC_ERR_UNSUPPORTED; }
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// ---------------------------------------------------------------- more elementwise
// binary: out = x (op) y ; unary: out = f(x; param)
__global__ void binary_kernel(int op, int64_t n, const double* __restrict__ x, const double* __restrict__ y,
                              double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double a = x[i], b = y[i];
    out[i] = (op == 0) ? a * b : (op == 1) ? a / b : (op == 2) ? a + b : a - b;
  }
}
__global__ void unary_kernel(int op, int64_t n, const double* __restrict__ x, double p, double* __restrict__ out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double a = x[i], r;
    switch (op) {
      case 0: r = sqrt(a); break;                          // sqrt
      case 1: r = fabs(a) > p ? 1.0 / a : a; break;        // reciprocal above threshold (else unchanged)
      case 2: r = fabs(a); break;                          // abs
      case 3: r = a * p; break;                            // scale
      default: r = 1.0 / a; break;                         // reciprocal
    }
    out[i] = r;
  }
}
// out[i][j] = A[i][j] * s[i] (rows) or * s[j] (cols)
__global__ void scale_rc_kernel(int cols, int64_t m, int64_t n, const double* __restrict__ A, const double* __restrict__ s,
                                double* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m * n; e += (int64_t)gridDim.x * blockDim.x)
    out[e] = A[e] * (cols ? s[e % n] : s[e / n]);
}

extern "C" int cc_binary_f64(void* stream, int op, int64_t n, const double* x, const double* y, double* out) {
  if (n <= 0) return CC_OK;
  if (op < 0 || op > 3) { cc_set_last_error("cc_binary_f64: bad op"); return CC_ERR_ARG; }
  binary_kernel<<<cc_blocks(n, 256), 256, 0, cc_stream(stream)>>>(op, n, x, y, out);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
extern "C" int cc_unary_f64(void* stream, int op, int64_t n, const double* x, double param, double* out) {
  if (n <= 0) return CC_OK;
  if (op < 0 || op > 4) { cc_set_last_error("cc_unary_f64: bad op"); return CC_ERR_ARG; }
  unary_kernel<<<cc_blocks(n, 256), 256, 0, cc_stream(stream)>>>(op, n, x, param, out);
  CC_LAUNCH_CHECK();
  return CC_OK;
}
extern "C" int cc_scale_rows_cols_f64(void* stream, int cols, int64_t m, int64_t n, const double* A, const double* s,
                                      double* out) {
  if (m <= 0 || n <= 0) return CC_OK;
  scale_rc_kernel<<<cc_blocks(m * n, 256), 256, 0, cc_stream(stream)>>>(cols, m, n, A, s, out);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

// ---------------------------------------------------------------- launch timing
// Events are recorded with a device-scope release (no system-scope cache write-back around the timed
// kernel, which would idle the GPU between the kernel and its neighbours) and come from a pool that
// lives until the process ends, so a timed launch costs two event records and no event creation.
namespace {
bool g_profile = false;
int g_profile_every = 1;
int64_t g_profile_seen[CC_PROF_KINDS];
struct ProfRecord {
  int kind;
  hipEvent_t e0, e1;
};
std::vector<ProfRecord> g_prof;
std::vector<hipEvent_t> g_event_pool;
void prof_clear() {
  for (const ProfRecord& r : g_prof) {
    g_event_pool.push_back(r.e0);
    g_event_pool.push_back(r.e1);
  }
  g_prof.clear();
}
int prof_event(hipEvent_t* e) {
  if (!g_event_pool.empty()) {
    *e = g_event_pool.back();
    g_event_pool.pop_back();
    return CC_OK;
  }
  CC_HIP(hipEventCreateWithFlags(e, hipEventReleaseToDevice));
  return CC_OK;
}
}  // namespace

int cc_prof_start(hipStream_t st, int kind, hipEvent_t* e0) {
  *e0 = nullptr;
  if (!g_profile || kind < 0 || kind >= CC_PROF_KINDS) return CC_OK;
  if (g_profile_seen[kind]++ % g_profile_every != 0) return CC_OK;
  const int rc = prof_event(e0);
  if (rc != CC_OK) return rc;
  CC_HIP(hipEventRecord(*e0, st));
  return CC_OK;
}

int cc_prof_stop(hipStream_t st, int kind, hipEvent_t e0) {
  if (!g_profile || !e0) return CC_OK;
  hipEvent_t e1 = nullptr;
  const int rc = prof_event(&e1);
  if (rc != CC_OK) return rc;
  CC_HIP(hipEventRecord(e1, st));
  g_prof.push_back(ProfRecord{kind, e0, e1});
  return CC_OK;
}

// A read of known size through direct global->LDS loads (the path the FP16 stage-3 kernel stages its slabs
// on): each workgroup takes 1024 consecutive 16-byte pieces per round, each wave four 64-lane instructions
// into its own LDS rows, waits, and adds one LDS value per thread; one value per workgroup goes to sink so
// that the loads are observable.  It calibrates the FETCH_SIZE counter for that load path.
namespace {
typedef __attribute__((address_space(3))) void cal_lds_void;
__global__ void __launch_bounds__(256) lds_read_kernel(int64_t n16, const uint4* __restrict__ src, float* __restrict__ sink) {
  __shared__ uint4 buf[1024];
  const int wave_base = threadIdx.x & ~63;
  float acc = 0.0f;
  for (int64_t base = (int64_t)blockIdx.x * 1024; base < n16; base += (int64_t)gridDim.x * 1024) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
      __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src + base + 256 * k + threadIdx.x),
                                       (cal_lds_void*)(buf + 256 * k + wave_base), 16, 0, 0);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    acc += reinterpret_cast<const float*>(buf)[threadIdx.x];
    __syncthreads();
  }
  if (threadIdx.x == 0) sink[blockIdx.x] = acc;
}
}  // namespace

extern "C" int cc_calibration_lds_read(void* stream, int64_t n_bytes, const void* src, float* sink, int blocks) {
  if (n_bytes <= 0 || n_bytes % 16384 != 0 || blocks < 1 || blocks > 65535) {
    cc_set_last_error("cc_calibration_lds_read: n_bytes must be a positive multiple of 16384 and blocks in [1, 65535]");
    return CC_ERR_ARG;
  }
  lds_read_kernel<<<blocks, 256, 0, cc_stream(stream)>>>(n_bytes / 16, static_cast<const uint4*>(src), sink);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
// Matrix-core throughput: each wave keeps eight independent accumulators and issues `iters` rounds of eight
// MFMAs (kind 0: v_mfma_f32_16x16x32_f16, 16384 FLOP each; kind 1: v_mfma_f64_16x16x4f64, 2048 FLOP each),
// then writes one value per thread to sink (blocks x 256 floats) so that the products are observable.
namespace {
// Distinct starting accumulators: with equal starts and operands the eight chains are one computation,
// which the compiler merges, and the merged loop copies accumulators around every MFMA.
__device__ inline f32x4 peak_start_f32(int j) {
  const float s0 = (float)(j + 1) * (float)threadIdx.x;
  return (f32x4){s0, s0, s0, s0};
}
__global__ void __launch_bounds__(256) mfma_peak_f16_kernel(int iters, float* __restrict__ sink) {
  f16x8 a, b;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    a[e] = (_Float16)(1e-3f * (float)((threadIdx.x + e) & 7));
    b[e] = (_Float16)(1e-3f * (float)((threadIdx.x * 3 + e) & 7));
  }
  f32x4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = peak_start_f32(j);
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, acc[j], 0, 0, 0);
  float t = 0.f;
#pragma unroll
  for (int j = 0; j < 8; ++j) t += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  sink[(int64_t)blockIdx.x * 256 + threadIdx.x] = t;
}
__global__ void __launch_bounds__(256) mfma_peak_f16k16_kernel(int iters, float* __restrict__ sink) {
  f16x4 a, b;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    a[e] = (_Float16)(1e-3f * (float)((threadIdx.x + e) & 7));
    b[e] = (_Float16)(1e-3f * (float)((threadIdx.x * 3 + e) & 7));
  }
  f32x4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = peak_start_f32(j);
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, acc[j], 0, 0, 0);
  float t = 0.f;
#pragma unroll
  for (int j = 0; j < 8; ++j) t += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  sink[(int64_t)blockIdx.x * 256 + threadIdx.x] = t;
}
__global__ void __launch_bounds__(256) mfma_peak_f64_kernel(int iters, float* __restrict__ sink) {
  const double a = 1e-3 * (double)(threadIdx.x & 7), b = 1e-3 * (double)((threadIdx.x * 3) & 7);
  f64x4 acc[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const double s0 = (double)(j + 1) * (double)threadIdx.x;
    acc[j] = (f64x4){s0, s0, s0, s0};
  }
  for (int i = 0; i < iters; ++i)
#pragma unroll
    for (int j = 0; j < 8; ++j) acc[j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[j], 0, 0, 0);
  double t = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) t += acc[j][0] + acc[j][1] + acc[j][2] + acc[j][3];
  sink[(int64_t)blockIdx.x * 256 + threadIdx.x] = (float)t;
}
}  // namespace

extern "C" int cc_calibration_mfma(void* stream, int kind, int iters, int blocks, float* sink) {
  if (kind < 0 || kind > 2 || iters < 1 || blocks < 1 || blocks > 65535) {
    cc_set_last_error("cc_calibration_mfma: kind must be 0 (f16 16x16x32), 1 (f64) or 2 (f16 16x16x16), iters >= 1, "
                      "blocks in [1, 65535]");
    return CC_ERR_ARG;
  }
  if (kind == 0) mfma_peak_f16_kernel<<<blocks, 256, 0, cc_stream(stream)>>>(iters, sink);
  else if (kind == 1) mfma_peak_f64_kernel<<<blocks, 256, 0, cc_stream(stream)>>>(iters, sink);
  else mfma_peak_f16k16_kernel<<<blocks, 256, 0, cc_stream(stream)>>>(iters, sink);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_profile_begin_sampled(int every) {
  if (every < 1) {
    cc_set_last_error("cc_profile_begin_sampled: every must be at least 1");
    return CC_ERR_ARG;
  }
  prof_clear();
  g_profile = true;
  g_profile_every = every;
  for (int k = 0; k < CC_PROF_KINDS; ++k) g_profile_seen[k] = 0;
  return CC_OK;
}

extern "C" int cc_profile_begin(void) { return cc_profile_begin_sampled(1); }

extern "C" int cc_profile_end(int n_kinds, double* total_ms, int64_t* launches) {
  g_profile = false;
  for (int k = 0; k < n_kinds; ++k) {
    total_ms[k] = 0.0;
    launches[k] = 0;
  }
  for (const ProfRecord& r : g_prof) {
    CC_HIP(hipEventSynchronize(r.e1));
    float ms = 0.0f;
    CC_HIP(hipEventElapsedTime(&ms, r.e0, r.e1));
    if (r.kind >= 0 && r.kind < n_kinds) {
      total_ms[r.kind] += ms;
      launches[r.kind] += 1;
    }
  }
  prof_clear();
  return CC_OK;
}
