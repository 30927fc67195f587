// Low-precision fused stage-3 matvec (FP16 or BF16 operands, FP32 accumulation on matrix cores).
//
//   Y[p'][q'][s'] = sum_t sum_alpha sum_{p,q,s} A_a(t)[alpha][p'][p] B_b(t)[alpha][q'][q] C_t[s'][s] v[p][q][s]
//
// Same arithmetic as the FP64 fused kernel.  Each slab channel is scaled by 1/max|slab| and rounded
// once; the scales fold into the per-term coefficient matrices C_t = scale_a * scale_b * O_t.  Slabs are
// stored (alpha, P, P) with the per-alpha stride rounded up to 8 elements, so every alpha's slab starts
// on a 16-byte boundary.
//
// Work decomposition.  The terms are cut on the host into chunks whose distinct channels fit one LDS
// buffer (a chunk never splits the terms of one A channel unless that channel alone overflows).  The grid
// is (alpha groups) x (chunks); a workgroup has NT = ceil(P/16) waves and walks its contiguous alpha
// range.  Per alpha it stages the chunk's slabs in LDS with direct global->LDS 16-byte loads, double
// buffered: the loads of alpha+1 are in flight while the waves compute alpha from the other buffer.
// Wave np owns the 16-row p' tile np:
//   stage A  W[idx = s*Q16 + q][p' in tile np] = sum_p V^T[idx][p] A_a[p'][p]     (v_mfma_f32_16x16x32)
//   stage B  Y[p' in tile np][q'][s'] += sum_q X[p'][q] B_b[q'][q],  X = sum_s C_t[s'][s] W[s*Q16 + q]
// The FP32 result of stage A (D[i][j] in lane j + 16(i/4), register i%4) is the stage-B A operand
// (A[i][k] in lane i + 16(k/8), element k%8) once converted, provided K runs in the order of lp_kidx
// (the B operands are read in the same order), so W never leaves registers.  Each workgroup builds
// the V^T fragments itself (vector staged in LDS, scaled by 1/max|v|).  Split mode BOTH carries the vector and the
// stage-B intermediate as hi + lo pairs (two MFMAs each), so only the slabs are rounded; mode VECTOR
// splits only the vector (stage B at one MFMA per product); mode NONE rounds every operand once.  Each
// workgroup writes one FP32 partial; a second kernel sums them in FP64 in a fixed order.
#include "cc_common.h"

#include <cstring>
#include <vector>

namespace {
constexpr int LP_MAX_CH = 16, LP_MAX_TERMS = 48, LP_MAX_D = 4;
constexpr int LP_MAX_CHUNKS = 16, LP_CHUNK_SLOTS = 32;
// Which operands carry a hi + lo pair: none (every operand rounded once), the vector and the stage-B
// intermediate (only the slabs rounded), or the vector only (the stage-B intermediate rounded once).
enum { LP_SPLIT_NONE = 0, LP_SPLIT_BOTH = 1, LP_SPLIT_VECTOR = 2 };
// Coefficient pattern of a term: a full C_t, a diagonal one, or (d = 2) an anti-diagonal one, which is
// diagonal once the stage-B products take the other half of W (Z'[s] = W[1 - s] B, Y[s] += C[s][1 - s] Z'[s]).
enum { LP_FULL = 0, LP_DIAG = 1, LP_ANTI = 2 };
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) void lds_void;

// Slab geometry shared by host and device: per-alpha stride (elements), and the number of slab slots
// one LDS buffer holds: 8 KB per wave of the workgroup (two buffers; 48 KB per workgroup at P = 36).
__host__ __device__ constexpr int lp_stride(int P) { return (P * P + 7) & ~7; }
__host__ __device__ constexpr int lp_slots(int P) {
  return (8192 * ((P + 15) / 16) / (2 * lp_stride(P))) < LP_CHUNK_SLOTS ? (8192 * ((P + 15) / 16) / (2 * lp_stride(P)))
                                                                       : LP_CHUNK_SLOTS;
}

// 16-byte pieces past the last slot that unconditional operand reads can touch: rows up to 16*NT - 1 and
// columns up to 32*KT - 1 of the last slab.
__host__ __device__ constexpr int lp_pad(int P) {
  return ((16 * ((P + 15) / 16) - 1) * P + 32 * ((P + 31) / 32) - lp_stride(P)) > 0
             ? (((16 * ((P + 15) / 16) - 1) * P + 32 * ((P + 31) / 32) - lp_stride(P)) * 2 + 15) / 16 + 1
             : 1;
}

struct LPChunk {
  int n_slots, n_runs;
  int run_slot[LP_MAX_TERMS];      // LDS slot of run r's A slab
  int run_t0[LP_MAX_TERMS + 1];    // run r covers chunk-ordered terms [run_t0[r], run_t0[r+1])
  int run_tm[LP_MAX_TERMS];        // ... of which [run_t0[r], run_tm[r]) are anti-diagonal (listed first)
  const void* slot_ptr[LP_CHUNK_SLOTS];
};
struct LPDevPlan {
  int64_t n_alpha;
  int n_chunks, pad;
  LPChunk chunk[LP_MAX_CHUNKS];
  int term_slot[LP_MAX_TERMS];     // LDS slot of each (chunk-ordered) term's B slab
  int term_a[LP_MAX_TERMS];        // the term's A and B channel indices (input order = chunk order)
  int term_b[LP_MAX_TERMS];
  float coef[LP_MAX_TERMS][LP_MAX_D * LP_MAX_D];       // scaled by the channel scales (read by the kernel)
  float coef_raw[LP_MAX_TERMS][LP_MAX_D * LP_MAX_D];   // as given (the source of every rescale)
};

template <typename T> struct LPType;
typedef __fp16 fp16x2 __attribute__((ext_vector_type(2)));
template <> struct LPType<_Float16> {
  typedef f16x4 vec4;
  typedef f16x8 vec8;
  static __device__ inline float round(float x) { return (float)(_Float16)x; }
  static __device__ inline _Float16 cvt(float x) { return (_Float16)x; }
  // hi + lo of two values: hi truncated to FP16 (one packed conversion), lo = the exact FP32 residual
  // truncated to FP16; |x - hi - lo| <= 2^-21 |x|
  static __device__ inline void split2(float a, float b, _Float16& ha, _Float16& hb, _Float16& la, _Float16& lb) {
    const fp16x2 h = __builtin_amdgcn_cvt_pkrtz(a, b);
    const fp16x2 l = __builtin_amdgcn_cvt_pkrtz(a - (float)h[0], b - (float)h[1]);
    ha = (_Float16)h[0];
    hb = (_Float16)h[1];
    la = (_Float16)l[0];
    lb = (_Float16)l[1];
  }
  static __device__ inline f32x4 mfma(vec8 a, vec8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
};
template <> struct LPType<__bf16> {
  typedef bf16x4 vec4;
  typedef bf16x8 vec8;
  static __device__ inline float round(float x) { return (float)(__bf16)x; }
  static __device__ inline __bf16 cvt(float x) { return (__bf16)x; }
  static __device__ inline void split2(float a, float b, __bf16& ha, __bf16& hb, __bf16& la, __bf16& lb) {
    ha = (__bf16)a;
    hb = (__bf16)b;
    la = (__bf16)(a - (float)ha);
    lb = (__bf16)(b - (float)hb);
  }
  static __device__ inline f32x4 mfma(vec8 a, vec8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
};

// K ordering of one 32-wide K tile t: lane group g = lane/16, element e (0..7) holds
//   k = 32t + 4g + e            (e < 4)
//   k = 32t + 16 + 4g + (e - 4) (e >= 4)
// so each lane reads two aligned 4-element runs of a slab row, and the stage-A result (4 consecutive rows
// per lane group of two 16-row tiles) is the stage-B A operand without any lane exchange.
__host__ __device__ constexpr int lp_kidx(int t, int g, int e) { return 32 * t + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4)); }

// 4 consecutive elements of row `row`, columns c0..c0+3 of a P x P slab in LDS (zero outside the slab)
template <typename T, int P>
__device__ inline typename LPType<T>::vec4 lds4(const T* s, int row, int c0) {
  typedef typename LPType<T>::vec4 V;
  if ((P % 4) == 0) {
    // Read unconditionally: rows past P only feed output rows/columns that are never written, and
    // columns past P meet zeros of the other operand; the buffers are padded and zero-filled so every
    // such read stays inside the workgroup's LDS and finite.
    return *reinterpret_cast<const V*>(s + row * P + c0);
  }
  V v;
  for (int k = 0; k < 4; ++k) {
    const bool ok = row < P && c0 + k < P;
    const T e = s[ok ? row * P + c0 + k : 0];
    v[k] = ok ? e : (T)0.0f;
  }
  return v;
}
// B operand of K tile t for slab row `row`: elements e hold columns lp_kidx(t, lane/16, e)
template <typename T, int P>
__device__ inline typename LPType<T>::vec8 lds8(const T* s, int row, int t, int g) {
  const typename LPType<T>::vec4 a = lds4<T, P>(s, row, 32 * t + 4 * g), b = lds4<T, P>(s, row, 32 * t + 16 + 4 * g);
  return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

// The B operand of K tile t for the slab rows of tile nq (row 16 nq + lane % 16), the same values as
// lds8(s, 16 nq + lane % 16, t, lane / 16).  For P % 4 == 0 the address is one lane-dependent offset
// from s plus a compile-time constant, which keeps every read within the few address steps over which
// the compiler's LDS lowering tags accesses with their buffer (an untagged read may be taken to alias
// the direct global->LDS loads in flight into the other buffer, and then waits for them).
template <typename T, int P>
__device__ inline typename LPType<T>::vec8 lds8b(const T* s, int nq, int t, int lane) {
  if constexpr ((P % 4) == 0) {
    typedef typename LPType<T>::vec4 V;
    const T* p = s + ((lane & 15) * P + 4 * (lane >> 4));
    const V a = *reinterpret_cast<const V*>(p + (16 * nq * P + 32 * t)),
            b = *reinterpret_cast<const V*>(p + (16 * nq * P + 32 * t + 16));
    return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
  } else {
    return lds8<T, P>(s, 16 * nq + (lane & 15), t, lane >> 4);
  }
}

// Stage A with the vector's hi and lo parts stacked along K: stacked index k < P holds hi at p = k,
// P <= k < 2P holds lo at p = k - P, and the vector fragments are zero past 2P.  That takes ceil(2P/32) K
// tiles instead of two products on each of ceil(P/32) tiles (3 instead of 4 at P = 36).  It needs
// P % 4 == 0, so that no aligned 4-element run of K crosses P or 2P.
template <int P, int SPLIT>
__host__ __device__ constexpr bool lp_stacked() { return SPLIT != LP_SPLIT_NONE && P % 4 == 0; }
template <int P, int SPLIT>
__host__ __device__ constexpr int lp_ka() { return lp_stacked<P, SPLIT>() ? (2 * P + 31) / 32 : (P + 31) / 32; }
// Slab column read for stacked index k: k mod P below 2P; past 2P (zero fragments) any finite entry, here
// k - 2P inside the row, or k itself, which stays below 32 ceil(P/32) (covered by lp_pad) whenever
// k - 2P >= P
template <int P>
__device__ inline int lp_scol(int k) { return k < P ? k : (k < 2 * P ? k - P : (k - 2 * P < P ? k - 2 * P : k)); }
template <typename T, int P>
__device__ inline typename LPType<T>::vec8 lds8s(const T* s, int row, int t, int g) {
  const typename LPType<T>::vec4 a = lds4<T, P>(s, row, lp_scol<P>(32 * t + 4 * g)),
                                 b = lds4<T, P>(s, row, lp_scol<P>(32 * t + 16 + 4 * g));
  return __builtin_shufflevector(a, b, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Direct global->LDS staging of one alpha's chunk slabs: wave w takes slots w, w + NT, ...; each
// 64-lane instruction moves 64 consecutive 16-byte pieces of one slab (slot base pointers are
// wave-uniform scalar loads, so no vector load sits between the issues).
template <int P, int NT>
__device__ __attribute__((always_inline)) inline void lp_stage(const LPChunk* __restrict__ C, int n_slots, int64_t al, uint4* buf, int wave, int lane) {
  constexpr int SPB = lp_stride(P) * 2, CPS = SPB / 16;
  for (int s = wave; s < n_slots; s += NT) {
    const char* src = reinterpret_cast<const char*>(C->slot_ptr[s]) + al * SPB;
#pragma unroll
    for (int k = 0; k < (CPS + 63) / 64; ++k) {
      const int j = 64 * k + lane;
      if (j < CPS)
        __builtin_amdgcn_global_load_lds(reinterpret_cast<const void*>(src + 16 * j), (lds_void*)(buf + s * CPS + 64 * k), 16, 0, 0);
    }
  }
}

// The chunk's term and run tables, one entry per lane (lane l: chunk-ordered term l, run l), read with
// v_readlane inside the loops so that no scalar memory load sits on the per-term path.
template <int DP>
struct LPMeta {
  int tslot;            // term l's B slot
  float coef[DP * DP];  // term l's coefficient matrix C_t[s'][s] at s' DP + s (diagonal terms too)
  float cd[DP];         // its diagonal form: C[s][s], or C[s][1 - s] for an anti-diagonal term
  int rslot, rt0, rtm;  // run l's A slot, first term (rt0 of lane n_runs: the chunk's term count) and
                        // first term past its anti-diagonal ones
};
template <int DP>
__device__ inline LPMeta<DP> lp_meta(const LPDevPlan* __restrict__ plan, const LPChunk* __restrict__ C, int lane) {
  LPMeta<DP> m;
  const int t = lane < LP_MAX_TERMS ? lane : LP_MAX_TERMS - 1;
  m.tslot = plan->term_slot[t];
#pragma unroll
  for (int k = 0; k < DP * DP; ++k) m.coef[k] = plan->coef[t][k];
  const bool anti = DP == 2 && plan->coef[t][0] == 0.f && plan->coef[t][DP * DP - 1] == 0.f;
#pragma unroll
  for (int sp = 0; sp < DP; ++sp) m.cd[sp] = plan->coef[t][sp * DP + (anti ? DP - 1 - sp : sp)];
  m.rslot = C->run_slot[t];
  m.rtm = C->run_tm[t];
  m.rt0 = C->run_t0[lane <= LP_MAX_TERMS ? lane : LP_MAX_TERMS];
  return m;
}
__device__ inline int lp_lane_int(int v, int l) { return __builtin_amdgcn_readlane(v, l); }
__device__ inline float lp_lane_float(float v, int l) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l)); }

// Y[s] += C[s] Z[s] for a term in diagonal form
template <int NT, int DP>
__device__ __forceinline__ void lp_combine_diag(f32x4 (&y)[NT][DP], const f32x4 (&z)[NT][DP], const float (&c)[DP]) {
#pragma unroll
  for (int nq = 0; nq < NT; ++nq)
#pragma unroll
    for (int s = 0; s < DP; ++s) y[nq][s] += c[s] * z[nq][s];
}

// Y[s'] += sum_s C[s'][s] Z[s] for one term's products
template <int NT, int DP>
__device__ __forceinline__ void lp_combine(f32x4 (&y)[NT][DP], const f32x4 (&z)[NT][DP], const float (&c)[DP * DP]) {
#pragma unroll
  for (int nq = 0; nq < NT; ++nq)
#pragma unroll
    for (int sp = 0; sp < DP; ++sp)
#pragma unroll
      for (int s = 0; s < DP; ++s) y[nq][sp] += c[sp * DP + s] * z[nq][s];
}

// One term of lp_alpha's diagonal-form loop: its products (with the other half of W when SWAP, an
// anti-diagonal term), the next term's operands into bop (Bn), and the pending previous term's
// diagonal-form combine, in one block so that the scheduler interleaves the combine with the MFMAs;
// the term then becomes the pending one.
template <int P, int DP, typename T, int SPLIT, bool SWAP>
__device__ __forceinline__ void lp_term_d(f32x4 (&y)[(P + 15) / 16][DP], f32x4 (&zp)[(P + 15) / 16][DP], float (&cpd)[DP],
                                          const typename LPType<T>::vec8 (&whi)[DP][(P + 31) / 32],
                                          const typename LPType<T>::vec8 (&wlo)[DP][(P + 31) / 32],
                                          typename LPType<T>::vec8 (&bop)[(P + 31) / 32][(P + 15) / 16], const T* Bn,
                                          const LPMeta<DP>& meta, int tt, int lane, int g) {
  typedef LPType<T> LT;
  constexpr int NT = (P + 15) / 16, KT = (P + 31) / 32;
  float cdn[DP];
#pragma unroll
  for (int k = 0; k < DP; ++k) cdn[k] = lp_lane_float(meta.cd[k], tt);
  f32x4 z[NT][DP];
#pragma unroll
  for (int nq = 0; nq < NT; ++nq)
#pragma unroll
    for (int s = 0; s < DP; ++s) z[nq][s] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < KT; ++t) {
#pragma unroll
    for (int nq = 0; nq < NT; ++nq)
#pragma unroll
      for (int s = 0; s < DP; ++s) z[nq][s] = LT::mfma(whi[SWAP ? DP - 1 - s : s][t], bop[t][nq], z[nq][s]);
    if (SPLIT == LP_SPLIT_BOTH) {
#pragma unroll
      for (int nq = 0; nq < NT; ++nq)
#pragma unroll
        for (int s = 0; s < DP; ++s) z[nq][s] = LT::mfma(wlo[SWAP ? DP - 1 - s : s][t], bop[t][nq], z[nq][s]);
    }
  }
#pragma unroll
  for (int t = 0; t < KT; ++t)
#pragma unroll
    for (int nq = 0; nq < NT; ++nq) bop[t][nq] = lds8b<T, P>(Bn, nq, t, lane);
  lp_combine_diag<NT, DP>(y, zp, cpd);
#pragma unroll
  for (int nq = 0; nq < NT; ++nq)
#pragma unroll
    for (int s = 0; s < DP; ++s) zp[nq][s] = z[nq][s];
#pragma unroll
  for (int k = 0; k < DP; ++k) cpd[k] = cdn[k];
}

// One alpha of one chunk for wave np (p' tile np).  Stage A: W tiles m = s*QT + mq over K tiles of p,
// converted once per A run to hi + lo stage-B operands (element e < 4 from W tile s*QT + 2t, e >= 4 from
// tile s*QT + 2t + 1, zero past QT).  Per term: Z[s] = W[s] B over K tiles of q, then
// Y[s'] += sum_s C_t[s'][s] Z[s].  The combine is software-pipelined: a term's products stay pending
// (zp, cp) and are combined during the next term's matrix products (across runs and alphas; the caller
// combines the last pending term), so the vector work overlaps the MFMAs.  The pending state starts as
// zeros with zero coefficients.
template <int P, int DP, typename T, int SPLIT, bool DIAGF>
__device__ __attribute__((always_inline)) inline void lp_alpha(const LPMeta<DP>& meta, int n_runs, const T* __restrict__ lds,
                                const typename LPType<T>::vec8 (&vt)[DP * ((P + 15) / 16)][lp_ka<P, SPLIT>()],
                                const typename LPType<T>::vec8 (&vl)[DP * ((P + 15) / 16)][(P + 31) / 32],
                                f32x4 (&y)[(P + 15) / 16][DP], f32x4 (&zp)[(P + 15) / 16][DP], float (&cp)[DP * DP],
                                float (&cpd)[DP], int np, int lane) {
  typedef LPType<T> LT;
  typedef typename LT::vec8 V;
  constexpr int NT = (P + 15) / 16, QT = NT, KT = (P + 31) / 32, MT = DP * QT, SP = lp_stride(P);
  constexpr bool STACK = lp_stacked<P, SPLIT>();
  constexpr int KA = lp_ka<P, SPLIT>();
  const int g = lane >> 4;
  // B operands of the next term, loaded during the current term's products (the chunk's terms run in
  // order across its runs; the last term reloads itself)
  const int tlast = lp_lane_int(meta.rt0, n_runs) - 1;
  V bop[KT][NT];
  {
    const T* Bs = lds + lp_lane_int(meta.tslot, lp_lane_int(meta.rt0, 0)) * SP;
#pragma unroll
    for (int t = 0; t < KT; ++t)
#pragma unroll
      for (int nq = 0; nq < NT; ++nq) bop[t][nq] = lds8b<T, P>(Bs, nq, t, lane);
  }
  // A operands of the next run, likewise loaded during the current run's terms
  V aop[KA];
  {
    const T* As = lds + lp_lane_int(meta.rslot, 0) * SP;
#pragma unroll
    for (int t = 0; t < KA; ++t)
      aop[t] = STACK ? lds8s<T, P>(As, 16 * np + (lane & 15), t, g) : lds8<T, P>(As, 16 * np + (lane & 15), t, g);
  }
  for (int r = 0; r < n_runs; ++r) {
    const int t0 = lp_lane_int(meta.rt0, r), t1 = lp_lane_int(meta.rt0, r + 1);
    V whi[DP][KT], wlo[DP][KT];
    {
      f32x4 w[MT];
#pragma unroll
      for (int m = 0; m < MT; ++m) w[m] = (f32x4){0.f, 0.f, 0.f, 0.f};
      if constexpr (STACK) {
#pragma unroll
        for (int t = 0; t < KA; ++t)
#pragma unroll
          for (int m = 0; m < MT; ++m) w[m] = LT::mfma(vt[m][t], aop[t], w[m]);
      } else {
#pragma unroll
        for (int t = 0; t < KT; ++t) {
          // hi products of all MT tiles, then the lo products: dependent MFMAs are MT issues apart
#pragma unroll
          for (int m = 0; m < MT; ++m) w[m] = LT::mfma(vt[m][t], aop[t], w[m]);
          if (SPLIT != LP_SPLIT_NONE) {
#pragma unroll
            for (int m = 0; m < MT; ++m) w[m] = LT::mfma(vl[m][t], aop[t], w[m]);
          }
        }
      }
      {
        const T* As = lds + lp_lane_int(meta.rslot, r + 1 < n_runs ? r + 1 : r) * SP;
#pragma unroll
        for (int t = 0; t < KA; ++t)
          aop[t] = STACK ? lds8s<T, P>(As, 16 * np + (lane & 15), t, g) : lds8<T, P>(As, 16 * np + (lane & 15), t, g);
      }
#pragma unroll
      for (int s = 0; s < DP; ++s)
#pragma unroll
        for (int t = 0; t < KT; ++t)
#pragma unroll
          for (int e = 0; e < 8; e += 2) {
            const int m = e < 4 ? 2 * t : 2 * t + 1;
            const float x0 = m < QT ? w[s * QT + m][e & 3] : 0.0f, x1 = m < QT ? w[s * QT + m][(e & 3) + 1] : 0.0f;
            if (SPLIT == LP_SPLIT_BOTH) {
              T h0, h1, l0, l1;
              LT::split2(x0, x1, h0, h1, l0, l1);
              whi[s][t][e] = h0;
              whi[s][t][e + 1] = h1;
              wlo[s][t][e] = l0;
              wlo[s][t][e + 1] = l1;
            } else {  // one rounding to nearest
              whi[s][t][e] = LT::cvt(x0);
              whi[s][t][e + 1] = LT::cvt(x1);
            }
          }
    }
    if constexpr (DIAGF) {
      // every term diagonal or anti-diagonal: the run's anti-diagonal terms (listed first) take the other
      // half of W, and every combine is in diagonal form
      const int tm = lp_lane_int(meta.rtm, r);
      for (int tt = t0; tt < tm; ++tt)
        lp_term_d<P, DP, T, SPLIT, true>(y, zp, cpd, whi, wlo, bop,
                                         lds + lp_lane_int(meta.tslot, tt < tlast ? tt + 1 : tlast) * SP, meta, tt, lane, g);
      for (int tt = tm; tt < t1; ++tt)
        lp_term_d<P, DP, T, SPLIT, false>(y, zp, cpd, whi, wlo, bop,
                                          lds + lp_lane_int(meta.tslot, tt < tlast ? tt + 1 : tlast) * SP, meta, tt, lane, g);
    } else {
      for (int tt = t0; tt < t1; ++tt) {
        float cn[DP * DP];
#pragma unroll
        for (int k = 0; k < DP * DP; ++k) cn[k] = lp_lane_float(meta.coef[k], tt);
        f32x4 z[NT][DP];
#pragma unroll
        for (int nq = 0; nq < NT; ++nq)
#pragma unroll
          for (int s = 0; s < DP; ++s) z[nq][s] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int t = 0; t < KT; ++t) {
          // hi products of all NT*DP accumulators, then the lo products
#pragma unroll
          for (int nq = 0; nq < NT; ++nq)
#pragma unroll
            for (int s = 0; s < DP; ++s) z[nq][s] = LT::mfma(whi[s][t], bop[t][nq], z[nq][s]);
          if (SPLIT == LP_SPLIT_BOTH) {
#pragma unroll
            for (int nq = 0; nq < NT; ++nq)
#pragma unroll
              for (int s = 0; s < DP; ++s) z[nq][s] = LT::mfma(wlo[s][t], bop[t][nq], z[nq][s]);
          }
        }
        {
          const T* Bs = lds + lp_lane_int(meta.tslot, tt < tlast ? tt + 1 : tlast) * SP;
#pragma unroll
          for (int t = 0; t < KT; ++t)
#pragma unroll
            for (int nq = 0; nq < NT; ++nq) bop[t][nq] = lds8b<T, P>(Bs, nq, t, lane);
        }
        lp_combine<NT, DP>(y, zp, cp);  // the previous term, independent of the products above
#pragma unroll
        for (int nq = 0; nq < NT; ++nq)
#pragma unroll
          for (int s = 0; s < DP; ++s) zp[nq][s] = z[nq][s];
#pragma unroll
        for (int k = 0; k < DP * DP; ++k) cp[k] = cn[k];
      }
    }
  }
}

template <int P, int DP, typename T, int SPLIT, bool DIAGF>
__global__ void __launch_bounds__(64 * ((P + 15) / 16), 2) lp_kernel(const LPDevPlan* __restrict__ plan,
                                                                     const double* __restrict__ v,
                                                                     double* __restrict__ scale,
                                                                     float* __restrict__ partial) {
  typedef LPType<T> LT;
  typedef typename LT::vec8 V;
  constexpr int NT = (P + 15) / 16, KT = (P + 31) / 32, MT = DP * NT, NY = P * P * DP;
  constexpr int CPS = lp_stride(P) * 2 / 16, BUF = lp_slots(P) * CPS + lp_pad(P);
  constexpr int NTH = 64 * NT;
  // stage-A operand fragments in LDS, vec8 [hi | lo][m][t][lane] (one stacked set when lp_stacked)
  constexpr bool STACK = lp_stacked<P, SPLIT>();
  constexpr int KA = lp_ka<P, SPLIT>();
  constexpr int FT = MT * KA * 64, NF = (SPLIT != LP_SPLIT_NONE && !STACK) ? 2 : 1;
  __shared__ uint4 buf0[BUF];
  __shared__ uint4 buf1[BUF];
  __shared__ uint4 frag[NF * FT];
  __shared__ float wmax[NT];
  const int lane = threadIdx.x & 63;
  const int np = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const LPChunk* C = &plan->chunk[blockIdx.y];
  const int n_slots = C->n_slots, n_runs = C->n_runs;
  // this workgroup's alpha range [a0, a1): 32-bit divisions whenever n_alpha * groups fits (the 64-bit
  // division is a long scalar sequence), with the same quotients
  int64_t a0, a1;
  {
    const uint64_t na = (uint64_t)plan->n_alpha;
    if (na * gridDim.x <= 0xFFFFFFFFull) {
      const uint32_t n32 = (uint32_t)na;
      a0 = (int64_t)((n32 * blockIdx.x) / gridDim.x);
      a1 = (int64_t)((n32 * (blockIdx.x + 1)) / gridDim.x);
    } else {
      a0 = (int64_t)(na * blockIdx.x / gridDim.x);
      a1 = (int64_t)(na * (blockIdx.x + 1) / gridDim.x);
    }
  }
  // The vector (all loads in flight at once; clamped index, masked value), then the first alpha's
  // slabs, so that both latencies overlap.
  // Thread t takes the DP components of the elements pq = t + NTH k (consecutive in v), so the index
  // arithmetic of an element is shared by its components.
  constexpr int NQ = (P * P + NTH - 1) / NTH;
  double xv[NQ][DP];
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int pq = threadIdx.x + NTH * k, pc = pq < P * P ? pq : P * P - 1;
#pragma unroll
    for (int sv = 0; sv < DP; ++sv) xv[k][sv] = v[pc * DP + sv];
  }
  if (a0 < a1) lp_stage<P, NT>(C, n_slots, a0, buf0, np, lane);
  // Past the staged slots the buffers hold zeros (operand reads run past the last slab), and so do the
  // fragment elements past the vector (q or p >= P).
  for (int i = n_slots * CPS + threadIdx.x; i < BUF; i += NTH) {
    buf0[i] = uint4{0u, 0u, 0u, 0u};
    buf1[i] = uint4{0u, 0u, 0u, 0u};
  }
  for (int i = threadIdx.x; i < NF * FT; i += NTH) frag[i] = uint4{0u, 0u, 0u, 0u};
  // Scale 1/max|v|: the maximum of the FP32 magnitudes, exact in any order, so every workgroup finds
  // the same scale.
  float mx = 0.f;
#pragma unroll
  for (int k = 0; k < NQ; ++k)
    if (threadIdx.x + NTH * k < P * P) {
#pragma unroll
      for (int sv = 0; sv < DP; ++sv) mx = fmaxf(mx, (float)fabs(xv[k][sv]));
    }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  if (lane == 0) wmax[np] = mx;
  __syncthreads();
  float m0 = 0.f;
#pragma unroll
  for (int w = 0; w < NT; ++w) m0 = fmaxf(m0, wmax[w]);
  if (!(m0 > 0.f)) m0 = 1.f;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) scale[0] = (double)m0;
  const double vinv = 1.0 / (double)m0;
  // Each element once per workgroup: v[(p P + q) DP + s] is element e of lane (q % 16) + 16 g of fragment
  // (m = s NT + q / 16, t = k / 32), where k % 32 = lp_kidx(0, g, e) for K index k = p (hi + lo in the
  // split modes; stacked: hi at k = p and lo at k = P + p of the one fragment set).
  T* fh = reinterpret_cast<T*>(frag);
  auto frag_index = [](int m, int kx, int q) {
    const int kk = kx & 31, g = (kk & 15) >> 2, e = (kk & 3) + ((kk >> 4) << 2);
    return (((m * KA + (kx >> 5)) * 64) + (q & 15) + 16 * g) * 8 + e;
  };
#pragma unroll
  for (int k = 0; k < NQ; ++k) {
    const int pq = threadIdx.x + NTH * k;
    if (pq < P * P) {
      const int q = pq % P, p = pq / P;
      // fragment m = sv NT + q / 16: the component sv sits NT * KA * 64 * 8 elements further on
      const int f0 = frag_index(q >> 4, p, q), f1 = STACK ? frag_index(q >> 4, P + p, q) : f0;
#pragma unroll
      for (int sv = 0; sv < DP; ++sv) {
        const int so = sv * NT * KA * 64 * 8;
        const float x = (float)(xv[k][sv] * vinv);
        fh[f0 + so] = LT::cvt(x);
        if (STACK) fh[f1 + so] = LT::cvt(x - LT::round(x));
        else if (NF == 2) fh[FT * 8 + f0 + so] = LT::cvt(x - LT::round(x));
      }
    }
  }
  __syncthreads();
  const V* fv = reinterpret_cast<const V*>(frag);
  V vt[MT][KA], vl[MT][KT];
#pragma unroll
  for (int m = 0; m < MT; ++m) {
#pragma unroll
    for (int t = 0; t < KA; ++t) vt[m][t] = fv[(m * KA + t) * 64 + lane];
#pragma unroll
    for (int t = 0; t < KT; ++t) vl[m][t] = (NF == 2) ? fv[FT + (m * KT + t) * 64 + lane] : V{};
  }
  f32x4 y[NT][DP];
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int s = 0; s < DP; ++s) y[j][s] = (f32x4){0.f, 0.f, 0.f, 0.f};

  const LPMeta<DP> meta = lp_meta<DP>(plan, C, lane);
  f32x4 zp[NT][DP];
  float cp[DP * DP], cpd[DP];
#pragma unroll
  for (int k = 0; k < DP; ++k) cpd[k] = 0.f;
#pragma unroll
  for (int j = 0; j < NT; ++j)
#pragma unroll
    for (int s = 0; s < DP; ++s) zp[j][s] = (f32x4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k = 0; k < DP * DP; ++k) cp[k] = 0.f;

  for (int64_t al = a0; al < a1; al += 2) {
    // buffer 0 holds alpha al; everyone has finished reading buffer 1 (alpha al-1)
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    if (al + 1 < a1) lp_stage<P, NT>(C, n_slots, al + 1, buf1, np, lane);
    lp_alpha<P, DP, T, SPLIT, DIAGF>(meta, n_runs, reinterpret_cast<const T*>(buf0), vt, vl, y, zp, cp, cpd, np, lane);
    if (al + 1 >= a1) break;
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    if (al + 2 < a1) lp_stage<P, NT>(C, n_slots, al + 2, buf0, np, lane);
    lp_alpha<P, DP, T, SPLIT, DIAGF>(meta, n_runs, reinterpret_cast<const T*>(buf1), vt, vl, y, zp, cp, cpd, np, lane);
  }
  if constexpr (DIAGF) lp_combine_diag<NT, DP>(y, zp, cpd);  // the last pending term
  else lp_combine<NT, DP>(y, zp, cp);
  // rows p' = 16 np + 4(lane/16) + r, columns q' = 16 nq + lane%16 of this workgroup's partial
  // (d = 2: both physical components of an element in one 8-byte store)
  float* out = partial + ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * NY;
#pragma unroll
  for (int nq = 0; nq < NT; ++nq)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int pp = 16 * np + 4 * (lane >> 4) + r, qq = 16 * nq + (lane & 15);
      if (pp < P && qq < P) {
        if constexpr (DP == 2) {
          *reinterpret_cast<float2*>(out + (pp * P + qq) * 2) = make_float2(y[nq][0][r], y[nq][1][r]);
        } else {
#pragma unroll
          for (int sp = 0; sp < DP; ++sp) out[(pp * P + qq) * DP + sp] = y[nq][sp][r];
        }
      }
    }
}

// y[i] = max|v| * sum over partial rows in row order (FP64).  A workgroup owns LP_RQ * VW columns; its
// LP_RG row groups take consecutive row ranges, each thread loading LP_RU rows of VW columns at once,
// and the group sums are added in group order.
constexpr int LP_RQ = 8, LP_RG = 32, LP_RU = 16;
template <int VW> struct LPVec;
template <> struct LPVec<4> { typedef float4 type; static __device__ float at(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; } };
template <> struct LPVec<2> { typedef float2 type; static __device__ float at(const float2& v, int k) { return k == 0 ? v.x : v.y; } };
template <> struct LPVec<1> { typedef float type; static __device__ float at(const float& v, int) { return v; } };
template <int VW>
__global__ void __launch_bounds__(LP_RQ * LP_RG) lp_reduce_kernel(int64_t rows, int64_t n, const float* __restrict__ partial,
                                                                 const double* __restrict__ scale, double* __restrict__ y) {
  typedef typename LPVec<VW>::type FV;
  __shared__ double part[LP_RG][LP_RQ * VW];
  const int q = threadIdx.x % LP_RQ, g = threadIdx.x / LP_RQ;
  const int64_t c0 = ((int64_t)blockIdx.x * LP_RQ + q) * VW;
  const int64_t rpt = (rows + LP_RG - 1) / LP_RG, r0 = g * rpt, r1 = rows < r0 + rpt ? rows : r0 + rpt;
  double acc[VW];
#pragma unroll
  for (int k = 0; k < VW; ++k) acc[k] = 0.0;
  if (c0 < n) {
    for (int64_t rb = r0; rb < r1; rb += LP_RU) {
      FV v[LP_RU];
#pragma unroll
      for (int u = 0; u < LP_RU; ++u) {  // clamped row: every load in bounds, the value masked below
        const int64_t r = rb + u < r1 ? rb + u : r0;
        v[u] = *reinterpret_cast<const FV*>(partial + r * n + c0);
      }
#pragma unroll
      for (int u = 0; u < LP_RU; ++u)
        if (rb + u < r1) {
#pragma unroll
          for (int k = 0; k < VW; ++k) acc[k] += (double)LPVec<VW>::at(v[u], k);
        }
    }
  }
#pragma unroll
  for (int k = 0; k < VW; ++k) part[g][q * VW + k] = acc[k];
  __syncthreads();
  if ((int)threadIdx.x < LP_RQ * VW) {
    const int64_t i = (int64_t)blockIdx.x * LP_RQ * VW + threadIdx.x;
    if (i < n) {
      double t = 0.0;
      for (int h = 0; h < LP_RG; ++h) t += part[h][threadIdx.x];
      y[i] = t * scale[0];
    }
  }
}

// the sum of the workgroup partials with the widest loads the row length allows (rows start 16-byte
// aligned: the partial area is)
inline void lp_reduce(hipStream_t st, int64_t rows, int64_t n, const float* partial, const double* scale, double* y) {
  if (n % 4 == 0) lp_reduce_kernel<4><<<(unsigned)((n + LP_RQ * 4 - 1) / (LP_RQ * 4)), LP_RQ * LP_RG, 0, st>>>(rows, n, partial, scale, y);
  else if (n % 2 == 0) lp_reduce_kernel<2><<<(unsigned)((n + LP_RQ * 2 - 1) / (LP_RQ * 2)), LP_RQ * LP_RG, 0, st>>>(rows, n, partial, scale, y);
  else lp_reduce_kernel<1><<<(unsigned)((n + LP_RQ - 1) / LP_RQ), LP_RQ * LP_RG, 0, st>>>(rows, n, partial, scale, y);
}

// low-precision slab = slab / max|slab|: pass 1 per-block maxima, pass 2 (one block) their maximum
// max |x| over n contiguous elements: block partials (four independent loads in flight per thread),
// then one block
__global__ void __launch_bounds__(256) lp_blockmax_kernel(int64_t n, const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double red[256];
  double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
  const int64_t stride = (int64_t)gridDim.x * 256;
  int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + 3 * stride < n; i += 4 * stride) {
    m0 = fmax(m0, fabs(x[i]));
    m1 = fmax(m1, fabs(x[i + stride]));
    m2 = fmax(m2, fabs(x[i + 2 * stride]));
    m3 = fmax(m3, fabs(x[i + 3 * stride]));
  }
  for (; i < n; i += stride) m0 = fmax(m0, fabs(x[i]));
  double m = fmax(fmax(m0, m1), fmax(m2, m3));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
// the same over a channel read through the slab map (a map that does not cover its source exactly)
__global__ void __launch_bounds__(256) lp_blockmax_map_kernel(int64_t n_alpha, int64_t P, CCSlabMap map,
                                                              const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double red[256];
  const int64_t PP = P * P, n = n_alpha * PP;
  double m = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const int64_t al = i / PP, e = i - al * PP;
    m = fmax(m, fabs(x[cc_slab_offset(map, al, e / P, e % P)]));
  }
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
constexpr int LP_CONVERT_MAX_P = 64;  // dynamic LDS of P x (P + 1) doubles
// y[alpha][i] = slab[alpha][i / P][i % P] / scale for i < P*P, zero in the stride padding; each slab is
// gathered into LDS in source order first (coalesced reads), then written row-major
template <typename T>
__global__ void __launch_bounds__(256) lp_convert_kernel(int64_t n_alpha, int64_t SP, CCSlabGather g,
                                                         const double* __restrict__ x, const double* __restrict__ scale,
                                                         T* __restrict__ y) {
  extern __shared__ double s[];  // P x (P + 1)
  const double inv = 1.0 / scale[0];
  const int P = g.P, ld = P + 1, PP = P * P;
  for (int64_t al = blockIdx.x; al < n_alpha; al += gridDim.x) {
    cc_slab_load(g, x, al, s, ld, inv);
    __syncthreads();
    T* out = y + al * SP;
    for (int i = threadIdx.x; i < SP; i += blockDim.x) out[i] = i < PP ? (T)s[(i / P) * ld + i % P] : (T)0.0f;
    __syncthreads();
  }
}

// The same for slabs whose four legs all have extent D (constant index arithmetic).
template <typename T, int D>
__global__ void __launch_bounds__(256) lp_convert_kernel_d(int64_t n_alpha, CCSlabGather g, const double* __restrict__ x,
                                                           const double* __restrict__ scale, T* __restrict__ y) {
  constexpr int P = D * D, ld = P + 1, PP = P * P, SP = lp_stride(P);
  __shared__ double s[P * ld];
  const double inv = 1.0 / scale[0];
  for (int64_t al = blockIdx.x; al < n_alpha; al += gridDim.x) {
    cc_slab_load_d<D, 256>(g, x, al, s, ld, inv);
    __syncthreads();
    T* out = y + al * SP;
    for (int i = threadIdx.x; i < SP; i += blockDim.x) out[i] = i < PP ? (T)s[(i / P) * ld + i % P] : (T)0.0f;
    __syncthreads();
  }
}
template <typename T>
void launch_lp_convert(hipStream_t st, unsigned grid, size_t lds, int64_t n_alpha, int64_t SP, int64_t d0, int64_t d1,
                       const CCSlabGather& g, const double* src, const double* scale, T* out) {
  if (d0 == d1 && d0 >= 2 && d0 <= 6) {
    switch (d0) {
      case 2: lp_convert_kernel_d<T, 2><<<grid, 256, 0, st>>>(n_alpha, g, src, scale, out); return;
      case 3: lp_convert_kernel_d<T, 3><<<grid, 256, 0, st>>>(n_alpha, g, src, scale, out); return;
      case 4: lp_convert_kernel_d<T, 4><<<grid, 256, 0, st>>>(n_alpha, g, src, scale, out); return;
      case 5: lp_convert_kernel_d<T, 5><<<grid, 256, 0, st>>>(n_alpha, g, src, scale, out); return;
      default: lp_convert_kernel_d<T, 6><<<grid, 256, 0, st>>>(n_alpha, g, src, scale, out); return;
    }
  }
  lp_convert_kernel<T><<<grid, 256, lds, st>>>(n_alpha, SP, g, src, scale, out);
}

// Workspace layout (in floats): device plan | partials [rows][n] | scale (one double, the vector's
// max|v|, written by the matvec kernel).  Pieces start on 16-byte boundaries.
struct LPLayout {
  int64_t plan, partial, scale, total;
};
inline int64_t lp_align4(int64_t x) { return (x + 3) & ~3LL; }
inline LPLayout lp_layout(int64_t P, int64_t d, int64_t rows) {
  const int64_t n = P * P * d;
  LPLayout L;
  L.plan = 0;
  L.partial = lp_align4((int64_t)(sizeof(LPDevPlan) + 3) / 4);
  L.scale = lp_align4(L.partial + rows * n);
  L.total = L.scale + 4;
  return L;
}

struct LPHandle {
  int64_t P, d;
  int dtype, groups, chunks, split;
  int diagf;           // 1 when every term's coefficient matrix is diagonal or anti-diagonal (d = 2)
  int n_terms, n_a;    // term count and A-channel count (the scale layout: A channels, then B channels)
  float* work;         // lp_layout(P, d, groups * LP_MAX_CHUNKS); the device plan sits at its start
  LPDevPlan* host;     // the host plan, kept until destroy (source of the stream-ordered upload)
  hipEvent_t uploaded; // recorded after the upload; destroy waits for it before freeing `host`
};

template <int P, int DP, typename T>
int launch_lp(hipStream_t st, const LPHandle* h, const double* v, double* y) {
  constexpr int NT = (P + 15) / 16;
  typedef typename LPType<T>::vec8 V;
  const int64_t n = (int64_t)P * P * DP;
  const LPLayout L = lp_layout(P, DP, (int64_t)h->groups * LP_MAX_CHUNKS);
  const LPDevPlan* plan = reinterpret_cast<const LPDevPlan*>(h->work + L.plan);
  float* partial = h->work + L.partial;
  double* scale = reinterpret_cast<double*>(h->work + L.scale);
  const dim3 grid((unsigned)h->groups, (unsigned)h->chunks);
  hipEvent_t e0 = nullptr;
  int rc = cc_prof_start(st, CC_PROF_LP, &e0);
  if (rc != CC_OK) return rc;
  if (h->diagf) {
    if (h->split == LP_SPLIT_BOTH) lp_kernel<P, DP, T, LP_SPLIT_BOTH, true><<<grid, 64 * NT, 0, st>>>(plan, v, scale, partial);
    else if (h->split == LP_SPLIT_VECTOR) lp_kernel<P, DP, T, LP_SPLIT_VECTOR, true><<<grid, 64 * NT, 0, st>>>(plan, v, scale, partial);
    else lp_kernel<P, DP, T, LP_SPLIT_NONE, true><<<grid, 64 * NT, 0, st>>>(plan, v, scale, partial);
  } else {
    if (h->split == LP_SPLIT_BOTH) lp_kernel<P, DP, T, LP_SPLIT_BOTH, false><<<grid, 64 * NT, 0, st>>>(plan, v, scale, partial);
    else if (h->split == LP_SPLIT_VECTOR) lp_kernel<P, DP, T, LP_SPLIT_VECTOR, false><<<grid, 64 * NT, 0, st>>>(plan, v, scale, partial);
    else lp_kernel<P, DP, T, LP_SPLIT_NONE, false><<<grid, 64 * NT, 0, st>>>(plan, v, scale, partial);
  }
  CC_LAUNCH_CHECK();
  rc = cc_prof_stop(st, CC_PROF_LP, e0);
  if (rc != CC_OK) return rc;
  lp_reduce(st, (int64_t)h->groups * h->chunks, n, partial, scale, y);
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

// Host: cut the (A-sorted) term table into chunks whose distinct slabs fit one LDS buffer.
// coef[t] = coef_raw[t] * scale[a(t)] * scale[n_a + b(t)] on the device (channel scales computed by the
// conversion); idempotent, so slabs converted again into the same buffers only need this kernel again
__global__ void lp_scale_coef_kernel(LPDevPlan* plan, int n_terms, int dd, int n_a, const double* __restrict__ scale) {
  for (int i = threadIdx.x; i < n_terms * dd; i += blockDim.x) {
    const int t = i / dd, e = i % dd;
    plan->coef[t][e] = (float)(scale[plan->term_a[t]] * scale[n_a + plan->term_b[t]] * (double)plan->coef_raw[t][e]);
  }
}

// Pattern of a d x d coefficient matrix (row-major): diagonal, anti-diagonal (d = 2, zero diagonal), or full.
int lp_term_kind(const float* c, int d) {
  bool diag = true, zero_diag = true;
  for (int e = 0; e < d * d; ++e) {
    if (c[e] != 0.0f && e / d != e % d) diag = false;
    if (c[e] != 0.0f && e / d == e % d) zero_diag = false;
  }
  return diag ? LP_DIAG : (d == 2 && zero_diag) ? LP_ANTI : LP_FULL;
}

int lp_build_plan(LPDevPlan* dp, int64_t P, int64_t d, const void* const* A, const void* const* B, int n_terms,
                  const int* ta, const int* tb, const float* coef, int64_t n_alpha) {
  const int cap = lp_slots((int)P);
  std::memset(dp, 0, sizeof(*dp));
  dp->n_alpha = n_alpha;
  int chunk = -1, out_t = 0;
  std::vector<const void*> slots;
  auto slot_of = [&](const void* ptr) {
    for (size_t i = 0; i < slots.size(); ++i)
      if (slots[i] == ptr) return (int)i;
    return -1;
  };
  auto open_chunk = [&]() -> bool {
    if (chunk + 1 >= LP_MAX_CHUNKS) return false;
    ++chunk;
    slots.clear();
    dp->chunk[chunk].n_runs = 0;
    dp->chunk[chunk].run_t0[0] = out_t;
    return true;
  };
  if (!open_chunk()) return CC_ERR_UNSUPPORTED;
  int t = 0;
  while (t < n_terms) {
    const void* a_ptr = A[ta[t]];
    // terms of this A channel
    int t_end = t;
    while (t_end < n_terms && ta[t_end] == ta[t]) ++t_end;
    // anti-diagonal terms first (each run lists them before the others; run_tm marks the split)
    std::vector<int> order;
    for (int pass = 0; pass < 2; ++pass)
      for (int k = t; k < t_end; ++k)
        if ((lp_term_kind(coef + k * d * d, (int)d) == LP_ANTI) == (pass == 0)) order.push_back(k);
    // place as many of them as fit; open a new chunk when even the first does not fit
    bool run_open = false;
    for (int k : order) {
      const void* b_ptr = B[tb[k]];
      const int need = (slot_of(a_ptr) < 0 ? 1 : 0) + (slot_of(b_ptr) < 0 ? 1 : 0);
      if ((int)slots.size() + need > cap) {
        // close the current run and chunk
        if (run_open) {
          LPChunk& c = dp->chunk[chunk];
          c.run_t0[c.n_runs + 1] = out_t;
          ++c.n_runs;
          run_open = false;
        }
        dp->chunk[chunk].n_slots = (int)slots.size();
        for (size_t i = 0; i < slots.size(); ++i) dp->chunk[chunk].slot_ptr[i] = slots[i];
        if (!open_chunk()) return CC_ERR_UNSUPPORTED;
      }
      if (!run_open) {
        if (slot_of(a_ptr) < 0) slots.push_back(a_ptr);
        LPChunk& c = dp->chunk[chunk];
        c.run_slot[c.n_runs] = slot_of(a_ptr);
        c.run_t0[c.n_runs] = out_t;
        c.run_tm[c.n_runs] = out_t;
        run_open = true;
      }
      if (slot_of(b_ptr) < 0) slots.push_back(b_ptr);
      if ((int)slots.size() > cap) return CC_ERR_UNSUPPORTED;  // cap < 2
      dp->term_slot[out_t] = slot_of(b_ptr);
      dp->term_a[out_t] = ta[k];
      dp->term_b[out_t] = tb[k];
      for (int e = 0; e < d * d; ++e) dp->coef[out_t][e] = coef[k * d * d + e];
      if (lp_term_kind(coef + k * d * d, (int)d) == LP_ANTI) dp->chunk[chunk].run_tm[dp->chunk[chunk].n_runs] = out_t + 1;
      ++out_t;
    }
    if (run_open) {
      LPChunk& c = dp->chunk[chunk];
      c.run_t0[c.n_runs + 1] = out_t;
      ++c.n_runs;
    }
    t = t_end;
  }
  dp->chunk[chunk].n_slots = (int)slots.size();
  for (size_t i = 0; i < slots.size(); ++i) dp->chunk[chunk].slot_ptr[i] = slots[i];
  dp->n_chunks = chunk + 1;
  return CC_OK;
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

extern "C" int cc_stage3_lp_slab_size(int64_t n_alpha, int64_t P, int64_t* out) {
  if (n_alpha < 0 || P < 1) { cc_set_last_error("lp slab size: bad arguments"); return CC_ERR_ARG; }
  *out = n_alpha * lp_stride((int)P);
  return CC_OK;
}

extern "C" int cc_stage3_lp_convert_strided(void* stream, int dtype, int64_t na0, int64_t na1, int64_t d0, int64_t d1,
                                            const int64_t* strides, int64_t src_elems, const double* src,
                                            double* scale_out, void* out) {
  CCSlabMap map;
  int rc = cc_slab_map(na0, na1, d0, d1, strides, &map);
  if (rc) return rc;
  if (!(dtype == CC_F16 || dtype == CC_BF16)) { cc_set_last_error("lp convert: dtype must be f16 or bf16"); return CC_ERR_ARG; }
  hipStream_t st = cc_stream(stream);
  const int64_t n_alpha = na0 * na1, P = d0 * d1, SP = lp_stride((int)P);
  if (P > LP_CONVERT_MAX_P) { cc_set_last_error("lp convert: slab extent above 64"); return CC_ERR_UNSUPPORTED; }
  // scale_out holds 513 doubles: block maxima in [1, 513), the maximum in [0]
  if (src_elems == n_alpha * P * P) lp_blockmax_kernel<<<512, 256, 0, st>>>(src_elems, src, scale_out + 1);
  else lp_blockmax_map_kernel<<<512, 256, 0, st>>>(n_alpha, P, map, src, scale_out + 1);
  CC_LAUNCH_CHECK();
  lp_absmax_kernel<<<1, 1024, 0, st>>>(512, scale_out + 1, scale_out);
  CC_LAUNCH_CHECK();
  const unsigned grid = (unsigned)(n_alpha < 16384 ? n_alpha : 16384);
  const CCSlabGather g = cc_slab_gather(map, d0);
  const size_t lds = (size_t)P * (P + 1) * sizeof(double);
  if (dtype == CC_F16) launch_lp_convert<_Float16>(st, grid, lds, n_alpha, SP, d0, d1, g, src, scale_out, (_Float16*)out);
  else launch_lp_convert<__bf16>(st, grid, lds, n_alpha, SP, d0, d1, g, src, scale_out, (__bf16*)out);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_stage3_lp_convert(void* stream, int dtype, int64_t n_alpha, int64_t P, const double* slab,
                                    double* scale_out, void* out) {
  if (n_alpha <= 0) return CC_OK;
  if (P < 1) { cc_set_last_error("lp convert: bad sizes"); return CC_ERR_ARG; }
  const int64_t st[6] = {P * P, 0, P, 0, 1, 0};
  return cc_stage3_lp_convert_strided(stream, dtype, n_alpha, 1, P, 1, st, n_alpha * P * P, slab, scale_out, out);
}

extern "C" int cc_stage3_lp_work_size(int64_t n_alpha, int64_t P, int64_t d, int groups, int64_t* out) {
  (void)n_alpha;
  *out = lp_layout(P, d, (int64_t)groups * LP_MAX_CHUNKS).total;  // floats
  return CC_OK;
}

extern "C" int cc_stage3_lp_plan_create(void** out, int dtype, int n_a, const void* const* A, int n_b,
                                        const void* const* B, int n_terms, const int* term_a, const int* term_b,
                                        const float* term_coef, int64_t n_alpha, int64_t P, int64_t d, int groups,
                                        int split, float* work, void* stream, const double* channel_scales) {
  *out = nullptr;
  if (!(dtype == CC_F16 || dtype == CC_BF16)) { cc_set_last_error("lp plan: dtype must be f16 or bf16"); return CC_ERR_ARG; }
  if (split < LP_SPLIT_NONE || split > LP_SPLIT_VECTOR) { cc_set_last_error("lp plan: split must be 0, 1 or 2"); return CC_ERR_ARG; }
  if (!cc_stage3_fused_supported(P, d)) { cc_set_last_error("lp plan: shape not instantiated"); return CC_ERR_UNSUPPORTED; }
  if (n_a < 1 || n_a > LP_MAX_CH || n_b < 1 || n_b > LP_MAX_CH || n_terms < 1 || n_terms > LP_MAX_TERMS || groups < 1 ||
      n_alpha < 1 || d > LP_MAX_D) {
    cc_set_last_error("lp plan: exceeds kernel limits");
    return CC_ERR_UNSUPPORTED;
  }
  for (int t = 0; t < n_terms; ++t)
    if (term_a[t] < 0 || term_a[t] >= n_a || term_b[t] < 0 || term_b[t] >= n_b || (t > 0 && term_a[t] < term_a[t - 1])) {
      cc_set_last_error("lp plan: bad term table (indices out of range or not sorted by A channel)");
      return CC_ERR_ARG;
    }
  LPDevPlan* dp = new LPDevPlan();
  int rc = lp_build_plan(dp, P, d, A, B, n_terms, term_a, term_b, term_coef, n_alpha);
  if (rc != CC_OK) {
    delete dp;
    cc_set_last_error("lp plan: the term table needs more chunks than the kernel supports");
    return rc;
  }
  std::memcpy(dp->coef_raw, dp->coef, sizeof(dp->coef));
  // stream-ordered upload (no host wait); the handle owns the host copy until destroy
  hipStream_t st = cc_stream(stream);
  hipError_t e = hipMemcpyAsync(work, dp, sizeof(LPDevPlan), hipMemcpyHostToDevice, st);
  if (e == hipSuccess && channel_scales != nullptr) {
    lp_scale_coef_kernel<<<1, 256, 0, st>>>(reinterpret_cast<LPDevPlan*>(work), n_terms, (int)(d * d), n_a, channel_scales);
    e = hipGetLastError();
  }
  if (e != hipSuccess) {
    hipStreamSynchronize(st);   // the copy may still read dp
    delete dp;
    cc_set_last_error(hipGetErrorString(e));
    return CC_ERR_HIP;
  }
  hipEvent_t ev = nullptr;
  e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventRecord(ev, st);
  if (e != hipSuccess) {
    hipStreamSynchronize(st);
    if (ev) hipEventDestroy(ev);
    delete dp;
    cc_set_last_error(hipGetErrorString(e));
    return CC_ERR_HIP;
  }
  LPHandle* h = new LPHandle();
  h->P = P;
  h->d = d;
  h->dtype = dtype;
  h->groups = groups;
  h->chunks = dp->n_chunks;
  h->split = split;
  h->diagf = d == 2;
  for (int t = 0; t < n_terms; ++t)
    if (lp_term_kind(dp->coef[t], (int)d) == LP_FULL) h->diagf = 0;
  h->n_terms = n_terms;
  h->n_a = n_a;
  h->work = work;
  h->host = dp;
  h->uploaded = ev;
  *out = h;
  return CC_OK;
}

extern "C" int cc_stage3_lp_plan_rescale(void* stream, void* plan, const double* channel_scales) {
  if (plan == nullptr || channel_scales == nullptr) { cc_set_last_error("lp plan rescale: null argument"); return CC_ERR_ARG; }
  const LPHandle* h = static_cast<const LPHandle*>(plan);
  const LPLayout L = lp_layout(h->P, h->d, (int64_t)h->groups * LP_MAX_CHUNKS);
  lp_scale_coef_kernel<<<1, 256, 0, cc_stream(stream)>>>(reinterpret_cast<LPDevPlan*>(h->work + L.plan), h->n_terms,
                                                         (int)(h->d * h->d), h->n_a, channel_scales);
  CC_LAUNCH_CHECK();
  return CC_OK;
}

extern "C" int cc_stage3_lp_plan_chunks(const void* plan, int* out) {
  if (!plan) { cc_set_last_error("null plan"); return CC_ERR_ARG; }
  *out = static_cast<const LPHandle*>(plan)->chunks;
  return CC_OK;
}

extern "C" int cc_stage3_lp_plan_diagonal_form(const void* plan, int* out) {
  if (!plan) { cc_set_last_error("null plan"); return CC_ERR_ARG; }
  *out = static_cast<const LPHandle*>(plan)->diagf;
  return CC_OK;
}

extern "C" int cc_stage3_lp_plan_apply(void* stream, const void* plan, const double* x, double* y) {
  if (!plan) { cc_set_last_error("null plan"); return CC_ERR_ARG; }
  return cc_internal_lp_apply(cc_stream(stream), plan, x, y);
}

extern "C" int cc_stage3_lp_plan_destroy(void* plan) {
  LPHandle* h = static_cast<LPHandle*>(plan);
  if (!h) return CC_OK;
  if (h->uploaded) {
    hipEventSynchronize(h->uploaded);   // the upload has read the host plan
    hipEventDestroy(h->uploaded);
  }
  delete h->host;
  delete h;
  return CC_OK;
}
