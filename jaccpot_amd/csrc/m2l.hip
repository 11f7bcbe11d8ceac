// M2L over the far interaction list: rotate the source multipole onto the
// pair axis, apply the z-axis M2L core, rotate the local back and sum per
// target node.
//
// Mapping: one wave per target node; lane l takes the target's far pairs
// l, l+64, ...  Every lane runs the whole per-pair cascade in registers:
// the rotation blocks are pair-independent constants folded into the
// instruction stream (tables_gen.h), so the per-pair state is just the
// coefficient vector and the cos/sin(m az), cos/sin(m ax) ladders.  The lane
// partial sums are reduced through LDS in a fixed order, so the result is
// deterministic and written once per target (no atomics).

#include <hip/hip_runtime.h>

#include <atomic>
#include <type_traits>

#include "device_math.hpp"
#include "launch.hpp"
#include "rows.hpp"
#include "xcd.hpp"

namespace jfm {
namespace {

constexpr int kWave = 64;
// coefficients per pass of the LDS reductions: a 64 x 33 buffer (8.4 KiB for
// float32) keeps LDS from limiting the waves resident per CU
constexpr int kChunk = 32;
// Coefficients per reduction pass.  The LDS buffer holds kWave rows of them,
// and a workgroup's LDS decides how many one-wave workgroups share a CU, so
// the C coefficients are split evenly into passes of at most kChunk (float32)
// or kChunk / 2 (float64): float64 order 4 takes two passes of 13 through
// 6.5 KiB instead of one pass through 17 KiB (3 -> 5 waves per SIMD), float32
// order 3 one pass of 16 through 4.5 KiB.  Rows have an odd length (row_for),
// so a lane's row starts in its own bank.
constexpr int chunk_for(int c, int bytes) {
  const int most = bytes == 4 ? kChunk : kChunk / 2;
  const int passes = (c + most - 1) / most;
  return (c + passes - 1) / passes;
}
constexpr int row_for(int c, int bytes) { return chunk_for(c, bytes) | 1; }

// ---------------------------------------------------------------------------
// Matrix-core form of the B_U applications (float32; m2l_engine "mfma").
//
// Each degree block of B_U (and B_U^T) splits into four connected components
// of at most five indices (tables_gen.h, kBuGroups).  v_mfma_f32_4x4x1_16b_f32
// computes sixteen independent 4x4 outer products per instruction: block b
// (lanes 4b..4b+3) forms D[b][i][j] += A[b][i] * B[b][j] with B[b][j] in lane
// 4b+j, and lane 4b+j receives D[b][0..3][j] in its four result registers.
// With A[b][i] = M[rows[i]][cols[k]] (the same in every block) and B[b][j] =
// the lane's own coefficient cols[k], a chain of one instruction per
// component column leaves in every lane the four output rows of its own pair:
// each lane keeps its pair's coefficient vector in registers, as in the VALU
// form, and the products run on the matrix pipe while the VALU does the Dz
// rotations and the z-axis core.
//
// The A operands are broadcast: with CBSZ = 4 the instruction takes the A
// values of block ABID (lanes 4 ABID .. 4 ABID + 3) for all sixteen blocks,
// so one VGPR carries the A operands of sixteen instructions.  Instruction q
// of matrix t reads register kBuRegs * t + q / 16 with ABID = q % 16, lane
// 4 (q % 16) + r holding row r; the twelve registers (85 instructions per
// matrix at degree 8) are loaded once per wave.
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBuRegs = (kBuInstMax + 15) / 16;
struct BuBcastTable {
  float v[2 * kBuRegs * 64];
};
constexpr BuBcastTable make_bu_bcast_table() {
  BuBcastTable tab{};
  for (int t = 0; t < 2; ++t)
    for (int g = 0; g < kBuRegs; ++g)
      for (int l = 0; l < 64; ++l) {
        const int q = 16 * g + l / 4;
        tab.v[(t * kBuRegs + g) * 64 + l] = q < kBuInstMax ? kBuMfmaA[t][q][l % 4] : 0.0f;
      }
  return tab;
}
__constant__ BuBcastTable kBuBcast = make_bu_bcast_table();

// This lane's A registers (kBuBcast rows of lane `lane`).
struct BuRegs {
  float r[2 * kBuRegs];
  __device__ __forceinline__ void load(int lane) {
#pragma unroll
    for (int q = 0; q < 2 * kBuRegs; ++q) r[q] = kBuBcast.v[q * 64 + lane];
  }
};

// Columns K.. of component GI (instruction q = g.inst0 + K).
template <bool TR, int GI, int K>
__device__ __forceinline__ void bu_mfma_cols(const float* in, f32x4& acc, const BuRegs& a) {
  constexpr BuGroup g = kBuGroups[TR ? 1 : 0][GI];
  if constexpr (K < g.ncols) {
    constexpr int q = g.inst0 + K;
    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(a.r[(TR ? kBuRegs : 0) + q / 16], in[g.cols[K]], acc, 4, q % 16, 0);
    bu_mfma_cols<TR, GI, K + 1>(in, acc, a);
  }
}

// blk <- M blk for the components GI.. of degree L (M = B_U, or B_U^T if TR);
// in holds the block's input.
template <int L, bool TR, int GI>
__device__ __forceinline__ void bu_mfma_groups(const float* in, float* blk, const BuRegs& a) {
  if constexpr (GI < kBuGroupOff[TR ? 1 : 0][L + 1]) {
    constexpr BuGroup g = kBuGroups[TR ? 1 : 0][GI];
    f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    bu_mfma_cols<TR, GI, 0>(in, acc, a);
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (g.rows[i] >= 0) blk[g.rows[i]] = acc[i];
    bu_mfma_groups<L, TR, GI + 1>(in, blk, a);
  }
}

// Degrees below kMfmaMinDegree stay on the VALU: their components are one or
// two indices, a single FMA each, cheaper than a matrix instruction.
constexpr int kMfmaMinDegree = 3;

template <int L, bool TR>
__device__ __forceinline__ void apply_bu_mf(float* blk, const BuRegs& a) {
  if constexpr (L < kMfmaMinDegree) {
    apply_bu<float, L, TR>(blk);
  } else {
    constexpr int W = 2 * L + 1;
    float in[W];
#pragma unroll
    for (int j = 0; j < W; ++j) in[j] = blk[j];
    bu_mfma_groups<L, TR, kBuGroupOff[TR ? 1 : 0][L]>(in, blk, a);
  }
}

template <int P, int L, Rot R>
__device__ __forceinline__ void rotate_block_mf(float* v, const Align<float, P>& al, const BuRegs& a) {
  float* blk = v + L * L;
  if constexpr (R == Rot::ToZMultipole) {  // B Dz(-ax) B Dz(az)
    apply_dz<float, L, false>(blk, al.ca, al.sa);
    apply_bu_mf<L, false>(blk, a);
    apply_dz<float, L, true>(blk, al.cx, al.sx);
    apply_bu_mf<L, false>(blk, a);
  } else {  // FromZLocal: Dz(-az) B^T Dz(ax) B^T
    static_assert(R == Rot::FromZLocal, "the M2L uses the multipole-to-z and local-from-z rotations");
    apply_bu_mf<L, true>(blk, a);
    apply_dz<float, L, false>(blk, al.cx, al.sx);
    apply_bu_mf<L, true>(blk, a);
    apply_dz<float, L, true>(blk, al.ca, al.sa);
  }
}

template <int P, Rot R, int L = 1>
__device__ __forceinline__ void rotate_all_mf(float* v, const Align<float, P>& al, const BuRegs& a) {
  if constexpr (L <= P) {
    rotate_block_mf<P, L, R>(v, al, a);
    rotate_all_mf<P, R, L + 1>(v, al, a);
  }
}

// One pair's M2L contribution w = D_from^L Z(r) D_to^M v at order G (v is
// overwritten); MF: B_U on the matrix cores (float32 only), else the VALU.
template <typename T, int G, bool MF>
__device__ __forceinline__ void m2l_pair(T* v, T* w, T dx, T dy, T dz, const BuRegs& bu) {
  Align<T, G> a;
  a.init(dx, dy, dz);
  if constexpr (MF) {
    static_assert(sizeof(T) == 4, "matrix-core B_U is float32 only");
    rotate_all_mf<G, Rot::ToZMultipole>(v, a, bu);
    zm2l<T, G>(v, w, a.r);
    rotate_all_mf<G, Rot::FromZLocal>(w, a, bu);
  } else {
    rotate_all<T, G, Rot::ToZMultipole>(v, a);
    zm2l<T, G>(v, w, a.r);
    rotate_all<T, G, Rot::FromZLocal>(w, a);
  }
}

// G <= P is the pair's expansion order (gear): the source multipole is read
// to degree G, translated at order G and added into degrees <= G of the
// target local.  ACC adds into the existing rows instead of overwriting them
// (used when a target's pairs are split across several gears).
template <typename T, int P, int G, bool ACC, bool MF>
__global__ __launch_bounds__(kWave) void m2l_kernel(const int32_t* __restrict__ targets,
                                                    const int32_t* __restrict__ far_off,
                                                    const int32_t* __restrict__ far_src, const T* __restrict__ mult,
                                                    const T* __restrict__ centers4, T* __restrict__ locals,
                                                    int32_t cs) {
  constexpr int C = ncoef(G);
  constexpr int CH = chunk_for(C, sizeof(T));
  __shared__ T red[kWave][row_for(C, sizeof(T))];
  const int tgt = targets[xcd_contiguous(blockIdx.x, gridDim.x)];
  const int lane = threadIdx.x;
  BuRegs bu;
  if constexpr (MF) bu.load(lane);
  const int b = far_off[tgt], e = far_off[tgt + 1];
  const V4<T> ct = load4(centers4, tgt);
  T acc[C];
#pragma unroll
  for (int j = 0; j < C; ++j) acc[j] = T(0);
  // MF: the matrix instructions need every lane of the wave, so the loop runs
  // to the wave's last round; lanes past the list repeat the target's first
  // pair and drop the result
  const int rounds = (e - b + kWave - 1) / kWave;
  for (int q = 0; q < rounds; ++q) {
    const int k = b + q * kWave + lane;
    if (!MF && k >= e) break;
    const bool live = k < e;
    const int src = far_src[live ? k : b];
    const V4<T> cs4 = load4(centers4, src);
    T v[C];
    load_row<T, C>(mult + static_cast<int64_t>(src) * cs, v);
    T w[C];
    m2l_pair<T, G, MF>(v, w, ct.x - cs4.x, ct.y - cs4.y, ct.z - cs4.z, bu);  // target - source
    if (live) {
#pragma unroll
      for (int j = 0; j < C; ++j) acc[j] += w[j];
    }
  }
  // lane partial sums reduced through LDS, CH coefficients at a time
  T* row = locals + static_cast<int64_t>(tgt) * cs;
#pragma unroll
  for (int c0 = 0; c0 < C; c0 += CH) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (c0 + j < C) red[lane][j] = acc[c0 + j];
    __syncthreads();
    if (lane < CH && c0 + lane < C) {
      T s = T(0);
      for (int l = 0; l < kWave; ++l) s += red[l][lane];
      if (ACC)
        row[c0 + lane] += s;
      else
        row[c0 + lane] = s;
    }
  }
  if (!ACC)
    for (int j = C + lane; j < cs; j += kWave) row[j] = T(0);
}

// Per-pair weights of the flat M2L (jfm_m2l_flat_weighted): pair (t <- s)
// is scaled by level_w[clamp(max(node_rung[t], node_rung[s]), 0, nlev - 1)].
template <typename T>
struct PairWeights {
  const int32_t* node_rung;
  const T* level_w;
  int32_t nlev;
};

// Flat mapping (jfm_m2l_flat): one wave per 64 consecutive pairs of a CSR
// far list (pairs grouped by target, pair_tgt[k] = target of pair k); lane l
// runs pair 64 w + l.  The lanes' contributions go to LDS and lane j sums
// coefficient j down the 64 rows, closing a sum at every change of target.
// A target whose pairs lie inside the wave is accumulated into its local
// directly; one that continues from the previous wave or into the next one
// leaves a partial sum in scratch (slot 0: the segment holding lane 0 when
// it continues an earlier wave, slot 1: a segment that starts in this wave
// and continues), and m2l_fixup_kernel adds those partials in wave order.
// Deterministic, no atomics; every lane carries a pair whatever the number
// of pairs per target.
// WT: every pair's contribution is scaled by its PairWeights weight.
template <typename T, int P, int G, bool MF, bool WT = false>
__global__ __launch_bounds__(kWave) void m2l_flat_kernel(int32_t K, const int32_t* __restrict__ pair_tgt,
                                                         const int32_t* __restrict__ pair_src,
                                                         const T* __restrict__ mult, const T* __restrict__ centers4,
                                                         T* __restrict__ locals, int32_t cs,
                                                         T* __restrict__ scratch, PairWeights<T> pw) {
  constexpr int C = ncoef(G);
  constexpr int CH = chunk_for(C, sizeof(T));
  __shared__ T red[kWave][row_for(C, sizeof(T))];
  __shared__ int32_t tg[kWave];
  const int lane = threadIdx.x;
  BuRegs bu;
  if constexpr (MF) bu.load(lane);
  const int blk = xcd_contiguous(blockIdx.x, gridDim.x);  // logical 64-pair block
  const int64_t k0 = static_cast<int64_t>(blk) * kWave;
  const int64_t k = k0 + lane;
  T w[C];
  int t = -1;
  if (MF || k < K) {
    // MF: every lane runs the cascade (the matrix instructions need the whole
    // wave); lanes past the list take pair k0 and contribute nothing
    const bool live = k < K;
    const int64_t kk = live ? k : k0;
    const int tt = pair_tgt[kk];
    const int src = pair_src[kk];
    const V4<T> ct = load4(centers4, tt);
    const V4<T> cs4 = load4(centers4, src);
    T v[C];
    load_row<T, C>(mult + static_cast<int64_t>(src) * cs, v);
    m2l_pair<T, G, MF>(v, w, ct.x - cs4.x, ct.y - cs4.y, ct.z - cs4.z, bu);  // target - source
    if constexpr (WT) {
      const int lv = min(max(max(pw.node_rung[tt], pw.node_rung[src]), 0), pw.nlev - 1);
      const T f = pw.level_w[lv];
#pragma unroll
      for (int j = 0; j < C; ++j) w[j] *= f;
    }
    if (live) {
      t = tt;
    } else {
#pragma unroll
      for (int j = 0; j < C; ++j) w[j] = T(0);
    }
  } else {
#pragma unroll
    for (int j = 0; j < C; ++j) w[j] = T(0);
  }
  // segment ends: the last lane, or a lane whose right neighbour has another target
  const int tn = __shfl_down(t, 1, kWave);
  const uint64_t ends0 = __ballot(lane == kWave - 1 || tn != t);
  tg[lane] = t;
  // neighbours of the wave's pair range (uniform)
  const int before = k0 > 0 ? pair_tgt[k0 - 1] : -1;
  const int after = k0 + kWave < K ? pair_tgt[k0 + kWave] : -1;
  T* part = scratch + static_cast<int64_t>(blk) * 2 * C;
  // the lanes' contributions go through LDS CH coefficients at a time
#pragma unroll
  for (int c0 = 0; c0 < C; c0 += CH) {
    __syncthreads();
#pragma unroll
    for (int j = 0; j < CH; ++j)
      if (c0 + j < C) red[lane][j] = w[c0 + j];
    __syncthreads();
    const int j = c0 + lane;
    uint64_t ends = ends0;
    int start = 0;
    while (ends) {  // scalar loop over the wave's segments
      const int last = __builtin_ctzll(ends);
      ends &= ends - 1;
      const int tl = __builtin_amdgcn_readfirstlane(tg[last]);  // the same for every lane
      if (tl >= 0 && lane < CH && j < C) {
        const bool cont_before = start == 0 && before == tl;
        const bool cont_after = last == kWave - 1 && after == tl;
        T s = T(0);
        for (int l = start; l <= last; ++l) s += red[l][lane];
        if (cont_before)
          part[j] = s;
        else if (cont_after)
          part[C + j] = s;
        else
          locals[static_cast<int64_t>(tl) * cs + j] += s;
      }
      start = last + 1;
    }
  }
}

// Targets whose pairs span waves w0 < w1 of the flat launch:
// local += partial(w0, slot 1) + sum_{w0 < w <= w1} partial(w, slot 0).
// Above 16 coefficients each thread takes four (target, coefficient)
// entries, a workgroup's entries kFixupThreads apart so that every load stays
// coalesced over the coefficients.  The loads a sum needs whatever the span
// (its span record, slot 1 of w0, slot 0 of w0 + 1 and the local) are issued
// for all the thread's entries before the first add: a span of two waves, the
// usual one at the high gears, costs two memory round trips instead of four.
// The low gears, whose targets span many blocks, keep one entry per thread
// and load eight blocks at a time (one memory round trip per eight blocks).
// Block partials are added in block order.
constexpr int kFixupThreads = 256;
constexpr int fixup_items(int c) { return c <= 16 ? 1 : 4; }
inline int fixup_blocks(int32_t num_span, int c) {
  const int64_t per = int64_t(kFixupThreads) * fixup_items(c);
  return static_cast<int>((static_cast<int64_t>(num_span) * c + per - 1) / per);
}

template <typename T, int G>
__global__ __launch_bounds__(kFixupThreads) void m2l_fixup_kernel(int32_t num_span,
                                                                  const int32_t* __restrict__ span_tgt,
                                                                  const int32_t* __restrict__ span_w0,
                                                                  const int32_t* __restrict__ span_w1,
                                                                  const T* __restrict__ scratch,
                                                                  T* __restrict__ locals, int32_t cs) {
  constexpr int C = ncoef(G);
  constexpr int Q = fixup_items(C);
  const int64_t total = static_cast<int64_t>(num_span) * C;
  const int64_t base = static_cast<int64_t>(blockIdx.x) * (kFixupThreads * Q) + threadIdx.x;
  bool live[Q];
  int j[Q], w0[Q], w1[Q];
  int64_t at[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int64_t g = base + static_cast<int64_t>(q) * kFixupThreads;
    live[q] = g < total;
    const int sp = live[q] ? static_cast<int>(g / C) : 0;  // dead entries read span 0 and store nothing
    j[q] = live[q] ? static_cast<int>(g - static_cast<int64_t>(sp) * C) : 0;
    at[q] = static_cast<int64_t>(span_tgt[sp]) * cs + j[q];
    w0[q] = span_w0[sp];
    w1[q] = span_w1[sp];
  }
  T s[Q], nx[Q], loc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    s[q] = scratch[static_cast<int64_t>(w0[q]) * 2 * C + C + j[q]];
    nx[q] = scratch[static_cast<int64_t>(w0[q] + 1) * 2 * C + j[q]];  // w1 > w0 for every span
    loc[q] = locals[at[q]];
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    T v = s[q] + nx[q];
    int w = w0[q] + 2;
    const int last = w1[q];
    if constexpr (C <= kWave) {
      for (; w + 7 <= last; w += 8) {
        T b[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) b[r] = scratch[static_cast<int64_t>(w + r) * 2 * C + j[q]];
#pragma unroll
        for (int r = 0; r < 8; ++r) v += b[r];
      }
    }
    for (; w <= last; ++w) v += scratch[static_cast<int64_t>(w) * 2 * C + j[q]];
    if (live[q]) locals[at[q]] = loc[q] + v;
  }
}

}  // namespace
}  // namespace jfm

using namespace jfm;

namespace {
// Engine of the float32 M2L rotations: 0 = VALU (B_U literals), 1 = matrix
// cores (v_mfma_f32_4x4x1_16b_f32 for degrees >= kMfmaMinDegree).  float64
// always runs on the VALU.
std::atomic<int32_t> g_m2l_engine{JFM_M2L_VALU};

template <typename T>
bool use_mfma() {
  return std::is_same<T, float>::value && g_m2l_engine.load(std::memory_order_relaxed) == JFM_M2L_MFMA;
}

template <typename T, int P, int G, bool ACC>
void launch_target(hipStream_t s, int32_t nt, const int32_t* targets, const int32_t* off, const int32_t* src,
                   const void* mult, const void* centers4, void* locals) {
  const int cs = coef_stride(P, sizeof(T));
  if constexpr (std::is_same<T, float>::value) {
    if (use_mfma<T>()) {
      m2l_kernel<T, P, G, ACC, true><<<nt, kWave, 0, s>>>(targets, off, src, (const T*)mult, (const T*)centers4,
                                                           (T*)locals, cs);
      return;
    }
  }
  m2l_kernel<T, P, G, ACC, false><<<nt, kWave, 0, s>>>(targets, off, src, (const T*)mult, (const T*)centers4,
                                                        (T*)locals, cs);
}
}  // namespace

extern "C" int32_t jfm_set_m2l_engine(int32_t engine) {
  if (engine != JFM_M2L_VALU && engine != JFM_M2L_MFMA) return fail(JFM_ERR_ARG, "jfm_set_m2l_engine: unknown engine");
  g_m2l_engine.store(engine, std::memory_order_relaxed);
  return JFM_OK;
}

extern "C" int32_t jfm_m2l_engine(void) { return g_m2l_engine.load(std::memory_order_relaxed); }

extern "C" int32_t jfm_m2l(void* stream, int32_t dtype, int32_t p, const int32_t* targets, int32_t num_targets,
                           const int32_t* far_off, const int32_t* far_src, const void* multipoles,
                           const void* centers4, void* locals) {
  if (num_targets < 0) return fail(JFM_ERR_ARG, "jfm_m2l: negative target count");
  if (num_targets == 0) return JFM_OK;
  if (!targets || !far_off || !far_src || !multipoles || !centers4 || !locals)
    return fail(JFM_ERR_ARG, "jfm_m2l: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    launch_target<T, P, P, false>(s, num_targets, targets, far_off, far_src, multipoles, centers4, locals);
  }));
  return check_launch("jfm_m2l");
}

namespace {
template <typename T, int P>
int32_t launch_gear(hipStream_t s, int32_t gear, bool acc, int32_t nt, const int32_t* targets, const int32_t* off,
                    const int32_t* src, const void* mult, const void* centers4, void* locals) {
#define JFM_GEAR(G)                                                                                \
  if (gear == G) {                                                                                 \
    if constexpr (G <= P) {                                                                        \
      if (acc)                                                                                     \
        launch_target<T, P, G, true>(s, nt, targets, off, src, mult, centers4, locals);            \
      else                                                                                         \
        launch_target<T, P, G, false>(s, nt, targets, off, src, mult, centers4, locals);           \
      return check_launch("jfm_m2l_gear");                                                         \
    }                                                                                              \
  }
  JFM_GEAR(1) JFM_GEAR(2) JFM_GEAR(3) JFM_GEAR(4) JFM_GEAR(5) JFM_GEAR(6) JFM_GEAR(7) JFM_GEAR(8)
#undef JFM_GEAR
  return fail(JFM_ERR_UNSUPPORTED, "jfm_m2l_gear: gear must be in [1, p]");
}
}  // namespace

extern "C" int32_t jfm_m2l_gear(void* stream, int32_t dtype, int32_t p, int32_t gear, int32_t accumulate,
                                const int32_t* targets, int32_t num_targets, const int32_t* far_off,
                                const int32_t* far_src, const void* multipoles, const void* centers4,
                                void* locals) {
  if (num_targets < 0) return fail(JFM_ERR_ARG, "jfm_m2l_gear: negative target count");
  if (gear < 1 || gear > p) return fail(JFM_ERR_ARG, "jfm_m2l_gear: gear must be in [1, p]");
  if (num_targets == 0) return JFM_OK;
  if (!targets || !far_off || !far_src || !multipoles || !centers4 || !locals)
    return fail(JFM_ERR_ARG, "jfm_m2l_gear: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    return launch_gear<T, P>(s, gear, accumulate != 0, num_targets, targets, far_off, far_src, multipoles,
                             centers4, locals);
  }));
  return JFM_OK;
}

extern "C" int64_t jfm_m2l_flat_scratch_size(int32_t gear, int32_t num_pairs) {
  if (gear < 1 || gear > 8 || num_pairs < 0) return -1;
  return ((static_cast<int64_t>(num_pairs) + 63) / 64) * 2 * ncoef(gear);
}

namespace {
template <typename T, int P, int G>
void launch_flat_pairs(hipStream_t s, int nw, int32_t K, const int32_t* pair_tgt, const int32_t* pair_src,
                       const void* mult, const void* centers4, void* locals, void* scratch) {
  const int cs = coef_stride(P, sizeof(T));
  const PairWeights<T> none{nullptr, nullptr, 0};
  if constexpr (std::is_same<T, float>::value) {
    if (use_mfma<T>()) {
      m2l_flat_kernel<T, P, G, true><<<nw, kWave, 0, s>>>(K, pair_tgt, pair_src, (const T*)mult,
                                                           (const T*)centers4, (T*)locals, cs, (T*)scratch, none);
      return;
    }
  }
  m2l_flat_kernel<T, P, G, false><<<nw, kWave, 0, s>>>(K, pair_tgt, pair_src, (const T*)mult, (const T*)centers4,
                                                        (T*)locals, cs, (T*)scratch, none);
}

template <typename T, int P>
int32_t launch_flat(hipStream_t s, int32_t gear, int32_t K, const int32_t* pair_tgt, const int32_t* pair_src,
                    const void* mult, const void* centers4, void* locals, void* scratch, int32_t num_span,
                    const int32_t* span_tgt, const int32_t* span_w0, const int32_t* span_w1) {
  const int cs = coef_stride(P, sizeof(T));
  const int nw = (K + kWave - 1) / kWave;
#define JFM_FLAT(G)                                                                                              \
  if (gear == G) {                                                                                               \
    if constexpr (G <= P) {                                                                                      \
      launch_flat_pairs<T, P, G>(s, nw, K, pair_tgt, pair_src, mult, centers4, locals, scratch);                 \
      if (num_span > 0)                                                                                          \
        m2l_fixup_kernel<T, G><<<fixup_blocks(num_span, ncoef(G)), kFixupThreads, 0, s>>>(                      \
            num_span, span_tgt, span_w0, span_w1, (const T*)scratch, (T*)locals, cs);                           \
      return check_launch("jfm_m2l_flat");                                                                       \
    }                                                                                                            \
  }
  JFM_FLAT(1) JFM_FLAT(2) JFM_FLAT(3) JFM_FLAT(4) JFM_FLAT(5) JFM_FLAT(6) JFM_FLAT(7) JFM_FLAT(8)
#undef JFM_FLAT
  return fail(JFM_ERR_UNSUPPORTED, "jfm_m2l_flat: gear must be in [1, p]");
}
}  // namespace

extern "C" int32_t jfm_m2l_flat(void* stream, int32_t dtype, int32_t p, int32_t gear, int32_t num_pairs,
                                const int32_t* pair_tgt, const int32_t* pair_src, const void* multipoles,
                                const void* centers4, void* locals, void* scratch, int32_t num_span,
                                const int32_t* span_tgt, const int32_t* span_w0, const int32_t* span_w1) {
  if (num_pairs < 0 || num_span < 0) return fail(JFM_ERR_ARG, "jfm_m2l_flat: negative count");
  if (gear < 1 || gear > p) return fail(JFM_ERR_ARG, "jfm_m2l_flat: gear must be in [1, p]");
  if (num_pairs == 0) return JFM_OK;
  if (!pair_tgt || !pair_src || !multipoles || !centers4 || !locals || !scratch)
    return fail(JFM_ERR_ARG, "jfm_m2l_flat: NULL buffer");
  if (num_span > 0 && (!span_tgt || !span_w0 || !span_w1)) return fail(JFM_ERR_ARG, "jfm_m2l_flat: NULL span list");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    return launch_flat<T, P>(s, gear, num_pairs, pair_tgt, pair_src, multipoles, centers4, locals, scratch,
                             num_span, span_tgt, span_w0, span_w1);
  }));
  return JFM_OK;
}

extern "C" int32_t jfm_m2l_flat_weighted(void* stream, int32_t dtype, int32_t p, int32_t num_pairs,
                                         const int32_t* pair_tgt, const int32_t* pair_src, const int32_t* node_rung,
                                         const void* level_weights, int32_t num_levels, const void* multipoles,
                                         const void* centers4, void* locals, void* scratch, int32_t num_span,
                                         const int32_t* span_tgt, const int32_t* span_w0, const int32_t* span_w1) {
  if (num_pairs < 0 || num_span < 0) return fail(JFM_ERR_ARG, "jfm_m2l_flat_weighted: negative count");
  if (num_levels < 1) return fail(JFM_ERR_ARG, "jfm_m2l_flat_weighted: need at least one level weight");
  if (num_pairs == 0) return JFM_OK;
  if (!pair_tgt || !pair_src || !node_rung || !level_weights || !multipoles || !centers4 || !locals || !scratch)
    return fail(JFM_ERR_ARG, "jfm_m2l_flat_weighted: NULL buffer");
  if (num_span > 0 && (!span_tgt || !span_w0 || !span_w1))
    return fail(JFM_ERR_ARG, "jfm_m2l_flat_weighted: NULL span list");
  hipStream_t s = as_stream(stream);
  const int nw = (num_pairs + kWave - 1) / kWave;
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    const int cs = coef_stride(P, sizeof(T));
    const PairWeights<T> pw{node_rung, (const T*)level_weights, num_levels};
    m2l_flat_kernel<T, P, P, false, true><<<nw, kWave, 0, s>>>(num_pairs, pair_tgt, pair_src, (const T*)multipoles,
                                                               (const T*)centers4, (T*)locals, cs, (T*)scratch, pw);
    if (num_span > 0)
      m2l_fixup_kernel<T, P><<<fixup_blocks(num_span, ncoef(P)), kFixupThreads, 0, s>>>(
          num_span, span_tgt, span_w0, span_w1, (const T*)scratch, (T*)locals, cs);
    return check_launch("jfm_m2l_flat_weighted");
  }));
  return JFM_OK;
}
