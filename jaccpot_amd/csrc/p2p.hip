// Near-field P2P: softened Newtonian sums over the self leaf and the near-list
// leaves of every target leaf, plus the fused evaluate entry point
// (L2P far field + P2P near field + combine + un-permute).
//
// Mapping: one 256-thread workgroup per target leaf.  Thread t owns target
// t % WT of the leaf (WT = leaf width rounded up to a power of two) and source
// partition t / WT.  The leaf's source slots (self leaf first, then the near
// leaves, each leaf_size slots wide) stream through LDS in tiles; inside a
// tile partition q takes slots q, q + NPART, ... so every lane of a wave reads
// one broadcast address per interaction and the tail tile stays balanced.
// The next tile's sources are fetched into registers while the current one is
// consumed, with the near-list entries loaded one tile further ahead than
// the positions they address (no load is waited for before a tile's
// interactions).
// Interactions use the hardware v_rsq_f32 (float32); per-tile partial sums are
// folded into the running total with Kahan compensation when requested; the
// partitions are reduced through LDS in fixed order (deterministic).

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "launch.hpp"
#include "rows.hpp"
#include "xcd.hpp"

namespace jfm {
namespace {

constexpr int kThreads = 256;
// LDS per CU is 160 KiB; a workgroup holding just over 1/(N+1) of it caps the
// CU at N workgroups (LDS is allocated in 512-byte granules).
constexpr int kLdsPerCu = 160 * 1024;
#ifndef JFM_P2P_BLOCKS_PER_CU
#define JFM_P2P_BLOCKS_PER_CU 5
#endif
constexpr int kP2PBlocksPerCu = JFM_P2P_BLOCKS_PER_CU;
constexpr int kP2PLdsBytes = (kLdsPerCu / (kP2PBlocksPerCu + 1) / 512 + 1) * 512;
static_assert(kP2PLdsBytes * (kP2PBlocksPerCu + 1) > kLdsPerCu && kP2PLdsBytes * kP2PBlocksPerCu <= kLdsPerCu,
              "P2P LDS reservation must cap the CU at kP2PBlocksPerCu workgroups");


template <typename T>
__device__ __forceinline__ T fast_rsqrt(T x) {
  if constexpr (sizeof(T) == 4)
    return __builtin_amdgcn_rsqf(x);
  else
    return 1.0 / sqrt(x);
}

template <typename T>
__device__ __forceinline__ void kahan_add(T& s, T& c, T v) {
  const T y = v - c;
  const T t = s + y;
  c = (t - s) - y;
  s = t;
}

// Source fetch in two stages.  A target leaf's slot sequence is its own
// leaf, then its nnear near leaves, leaf_size slots each; thread t owns slots
// first + l * STRIDE (l < LOADS).  fetch_list loads the raw near-list entries
// of those slots (indices clamped into range; an empty list reads
// `fallback`, a valid int32 location, and the value is discarded; no
// arithmetic on the loaded values, so nothing waits for them here);
// fetch_positions decodes the same slots with those entries one tile later
// and loads the positions (slots past `total` or past the last particle are
// marked invalid and become zero-mass sources).  The P2P loop keeps list
// loads one tile ahead of position loads, so neither is waited for before the
// interactions of the current tile.  For leaf_size == WT (power-of-two
// leaves) the slot decode is a shift and a mask.
// FIXED: leaf_size == WT, known at compile time (shift and mask, no branch).
template <int WT, bool FIXED>
__device__ __forceinline__ void slot_decode(int leaf_size, int slot, int& qi, int& o) {
  if constexpr (FIXED) {
    qi = static_cast<int>(static_cast<unsigned>(slot) / WT);
    o = static_cast<int>(static_cast<unsigned>(slot) % WT);
  } else {
    qi = slot / leaf_size;
    o = slot - qi * leaf_size;
  }
}

template <int WT, bool FIXED, int LOADS, int STRIDE>
__device__ __forceinline__ void fetch_list(int leaf_size, const int32_t* __restrict__ near_list, int nnear,
                                           const int32_t* __restrict__ fallback, int first, int (&v)[LOADS]) {
  const int32_t* list = nnear > 0 ? near_list : fallback;
  const int hi = max(nnear - 1, 0);
#pragma unroll
  for (int l = 0; l < LOADS; ++l) {
    int qi, o;
    slot_decode<WT, FIXED>(leaf_size, first + l * STRIDE, qi, o);
    v[l] = list[min(max(qi - 1, 0), hi)];
  }
}

template <typename T, int WT, bool FIXED, int LOADS, int STRIDE>
__device__ __forceinline__ void fetch_positions(const T* __restrict__ pos4, int n, int leaf_size, int leaf, int total,
                                                int first, const int (&v)[LOADS], V4<T> (&out)[LOADS],
                                                bool (&valid)[LOADS]) {
#pragma unroll
  for (int l = 0; l < LOADS; ++l) {
    int qi, o;
    slot_decode<WT, FIXED>(leaf_size, first + l * STRIDE, qi, o);
    const int sl = qi > 0 ? v[l] : leaf;
    const int g = first + l * STRIDE < total ? sl * leaf_size + o : n;
    valid[l] = g < n;
    out[l] = load4(pos4, min(g, n - 1));
  }
}

// Leaf-granular fetch (power-of-two leaves, leaf_size == WT): thread t takes
// the LOADS consecutive slots first .. first + LOADS - 1, which lie in one
// source leaf, so one near-list entry and one slot decode serve them all.
template <int WT>
__device__ __forceinline__ int fetch_list_leaf(const int32_t* __restrict__ near_list, int nnear,
                                               const int32_t* __restrict__ fallback, int first) {
  const int32_t* list = nnear > 0 ? near_list : fallback;
  const int qi = static_cast<int>(static_cast<unsigned>(first) / WT);
  return list[min(max(qi - 1, 0), max(nnear - 1, 0))];
}

template <typename T, int WT, int LOADS>
__device__ __forceinline__ void fetch_positions_leaf(const T* __restrict__ pos4, int n, int leaf, int total, int first,
                                                     int v, V4<T> (&out)[LOADS], bool (&valid)[LOADS]) {
  const int qi = static_cast<int>(static_cast<unsigned>(first) / WT);
  const int o = static_cast<int>(static_cast<unsigned>(first) % WT);
  const int g0 = (qi > 0 ? v : leaf) * WT + o;
#pragma unroll
  for (int l = 0; l < LOADS; ++l) {
    const int g = first + l < total ? g0 + l : n;
    valid[l] = g < n;
    out[l] = load4(pos4, min(g, n - 1));
  }
}

// float32 leaf-granular fetch through a raw buffer resource over the n rows
// of pos4 (16 bytes each).  The hardware range check returns zero for a row
// at or past n, i.e. a zero-mass source at the origin, so the missing rows of
// a ragged last leaf and slots past `total` (sent to row n) need no clamp and
// no mask, and the LOADS rows of a thread share one offset register (the row
// steps go into the instructions' immediate offsets).  Requires total and
// first to be multiples of LOADS (slot ranges are whole leaves); row offsets
// reach 16 * (n + WT) with WT <= 128 on this path, kept below 2^31.
constexpr int kBufMaxRows = (1 << 27) - 512;
static_assert((static_cast<int64_t>(kBufMaxRows) + 128) * 16 + 64 < (int64_t(1) << 31), "buffer offsets fit int32");
constexpr int kBufDword3 = 0x00020000;  // untyped 32-bit data format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t pos4_rsrc(const void* pos4, int n) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(pos4), static_cast<short>(0), n * 16, kBufDword3);
}

template <int WT, int LOADS>
__device__ __forceinline__ void fetch_positions_buf(__amdgpu_buffer_rsrc_t rsrc, int n, int leaf, int total,
                                                    int first, int v, V4<float> (&out)[LOADS]) {
  const int qi = static_cast<int>(static_cast<unsigned>(first) / WT);
  const int o = static_cast<int>(static_cast<unsigned>(first) % WT);
  const int g0 = first < total ? (qi > 0 ? v : leaf) * WT + o : n;
#pragma unroll
  for (int l = 0; l < LOADS; ++l)
    out[l] = __builtin_bit_cast(V4<float>, __builtin_amdgcn_raw_buffer_load_b128(rsrc, g0 * 16 + 16 * l, 0, 0));
}

// Work items split the source slots of leaves with long near lists: item k
// covers slots [s0, s1) of leaf `leaf`; part < 0 -> the item is the leaf's
// only one and writes the output rows, else it writes the unscaled partial
// sums of its targets to partials[(part * leaf_size + i) * 4 + c] and
// p2p_reduce_kernel adds a leaf's partials in item order.  leaf == NULL: one
// item per leaf covering all of its slots.
template <typename T>
struct P2PItems {
  const int32_t* leaf;
  const int32_t* s0;
  const int32_t* s1;
  const int32_t* part;
  T* partials;
  int32_t num_part;  // rows of partials (items with part >= num_part write nothing)
};

// Workgroup size of the float32 leaf-granular path for leaves of 8: 128
// threads (16 source partitions), so each thread runs 64 interactions per
// 1024-source tile and the per-item set-up and partition reduction are shared
// by half as many threads as with 256.
#ifndef JFM_P2P_NARROW_THREADS
#define JFM_P2P_NARROW_THREADS 128
#endif
constexpr int kNarrowThreads = JFM_P2P_NARROW_THREADS;
// Narrow workgroups per CU (LDS reservation, as for the 256-thread form): 8 x
// 2 waves, so their registers leave room on every SIMD for a tree-pass wave.
#ifndef JFM_P2P_NARROW_BLOCKS_PER_CU
#define JFM_P2P_NARROW_BLOCKS_PER_CU 8
#endif

// SAFE: guard r^2 == 0 (unsoftened self term / coincident points).
// THREADS: workgroup size (kThreads, or kNarrowThreads for narrow leaves).
template <typename T, int WT, bool POT, bool COMP, bool SAFE, bool FIXED, int THREADS>
__global__ __launch_bounds__(THREADS) void p2p_kernel(const T* __restrict__ pos4, int32_t n, int32_t leaf_size,
                                                      const int32_t* __restrict__ near_off,
                                                      const int32_t* __restrict__ near_src, T eps2, T G,
                                                      const int32_t* __restrict__ out_index, T* __restrict__ acc,
                                                      T* __restrict__ pot, int32_t accumulate, P2PItems<T> items) {
  constexpr int NPART = THREADS / WT;
  constexpr int TILE = sizeof(T) == 4 ? 1024 : 512;
  constexpr int LOADS = TILE / THREADS;
  static_assert(2 * THREADS <= TILE, "reduction buffer must fit in the tile");
  // REMAP: leaf-granular fetch; slot s of a tile is stored at
  // (s % LOADS) * ROW + s / LOADS, so the fetching thread writes its LOADS
  // slots to LOADS conflict-free rows and partition p still reads one address
  // per interaction, (p % LOADS) * ROW + p / LOADS + k * NPART / LOADS.  Rows
  // are padded so that the eight partitions a wave reads fall in distinct
  // 4-bank groups: with LOADS = 4 they sit in rows p % 4 at columns p / 4
  // (ROW = THREADS + 2), with LOADS = 8 in rows p at column 0 (ROW = THREADS + 1).
  constexpr bool REMAP = FIXED && (NPART % LOADS == 0);
  constexpr bool BUF = REMAP && sizeof(T) == 4;  // float32: buffer-resource fetch
  constexpr int ROW = REMAP ? THREADS + (LOADS == 8 ? 1 : 2) : THREADS;
  // The tile array is padded so that at most kBlocksPerCu workgroups share a
  // CU: the rest of its registers and LDS stay free for the tree-pass kernels
  // that run beside P2P in an overlapped step (at full occupancy their small
  // dependent launches wait for P2P workgroups to end).
  constexpr int kBlocksPerCu = THREADS == kThreads ? kP2PBlocksPerCu : JFM_P2P_NARROW_BLOCKS_PER_CU;
  constexpr int kPadBytes = (kLdsPerCu / (kBlocksPerCu + 1) / 512 + 1) * 512;
  constexpr int kTileElems = LOADS * ROW;
  constexpr int kPadElems = static_cast<int>(kPadBytes / sizeof(V4<T>));
  __shared__ V4<T> tile[kTileElems > kPadElems ? kTileElems : kPadElems];

  // XCD-aware order: workgroups are dealt round-robin to the XCDs, so hand
  // each XCD a contiguous run of leaves (neighbouring leaves share most of
  // their source leaves, which then stay in that XCD's L2).  Bijective:
  // XCD x < r owns q+1 leaves, the others q, where nblk = 8q + r.
  const int item = xcd_contiguous(blockIdx.x, gridDim.x);
  const int leaf = items.leaf ? items.leaf[item] : item;
  const int i = threadIdx.x % WT;
  const int part = threadIdx.x / WT;
  const int first = leaf * leaf_size;
  const int cnt = min(leaf_size, n - first);
  const V4<T> xi = load4(pos4, first + (i < cnt ? i : 0));
  const int nb = near_off[leaf];
  const int nsl = 1 + (near_off[leaf + 1] - nb);
  // slots [start, total) of the leaf's sequence (self leaf, then near leaves)
  const int start = items.leaf ? items.s0[item] : 0;
  const int total = items.leaf ? items.s1[item] : nsl * leaf_size;
  const int part_out = items.leaf ? items.part[item] : -1;

  // The next tile's sources are fetched into registers while the current
  // tile is being consumed, hiding the global-load latency behind the
  // interaction loop.
  V4<T> next[LOADS];
  bool ok[LOADS];
  int lst[LOADS];  // near-list entries of the tile after `next`
  const int tid = threadIdx.x;
  const __amdgpu_buffer_rsrc_t rsrc = pos4_rsrc(pos4, BUF ? n : 0);  // used by the float32 fetch only
  if constexpr (REMAP) {
    lst[0] = fetch_list_leaf<WT>(near_src + nb, nsl - 1, near_off, start + LOADS * tid);
    if constexpr (BUF)
      fetch_positions_buf<WT, LOADS>(rsrc, n, leaf, total, start + LOADS * tid, lst[0],
                                     reinterpret_cast<V4<float>(&)[LOADS]>(next));
    else
      fetch_positions_leaf<T, WT, LOADS>(pos4, n, leaf, total, start + LOADS * tid, lst[0], next, ok);
    lst[0] = fetch_list_leaf<WT>(near_src + nb, nsl - 1, near_off, start + TILE + LOADS * tid);
  } else {
    fetch_list<WT, FIXED, LOADS, THREADS>(leaf_size, near_src + nb, nsl - 1, near_off, start + tid, lst);
    fetch_positions<T, WT, FIXED, LOADS, THREADS>(pos4, n, leaf_size, leaf, total, start + tid, lst, next, ok);
    fetch_list<WT, FIXED, LOADS, THREADS>(leaf_size, near_src + nb, nsl - 1, near_off, start + TILE + tid, lst);
  }
  const int rbase = REMAP ? (part % LOADS) * ROW + part / LOADS : part;
  constexpr int RSTEP = REMAP ? NPART / LOADS : NPART;

  T sx = 0, sy = 0, sz = 0, sp = 0;
  T kx = 0, ky = 0, kz = 0, kp = 0;
  for (int base = start; base < total; base += TILE) {
    __syncthreads();
#pragma unroll
    for (int l = 0; l < LOADS; ++l) {
      V4<T> q = next[l];
      if (!BUF && !ok[l]) q.w = T(0);
      tile[l * ROW + threadIdx.x] = q;
    }
    __syncthreads();
    // unconditional (past the end they yield invalid slots): a branch here
    // would make the compiler wait for the loads at its join
    if constexpr (BUF) {
      fetch_positions_buf<WT, LOADS>(rsrc, n, leaf, total, base + TILE + LOADS * tid, lst[0],
                                     reinterpret_cast<V4<float>(&)[LOADS]>(next));
      lst[0] = fetch_list_leaf<WT>(near_src + nb, nsl - 1, near_off, base + 2 * TILE + LOADS * tid);
    } else if constexpr (REMAP) {
      fetch_positions_leaf<T, WT, LOADS>(pos4, n, leaf, total, base + TILE + LOADS * tid, lst[0], next, ok);
      lst[0] = fetch_list_leaf<WT>(near_src + nb, nsl - 1, near_off, base + 2 * TILE + LOADS * tid);
    } else {
      fetch_positions<T, WT, FIXED, LOADS, THREADS>(pos4, n, leaf_size, leaf, total, base + TILE + tid, lst, next,
                                                    ok);
      fetch_list<WT, FIXED, LOADS, THREADS>(leaf_size, near_src + nb, nsl - 1, near_off, base + 2 * TILE + tid,
                                            lst);
    }
    // Slots past `total` hold zero-mass sources, so every partition runs the
    // same (wave-uniform, scalar-controlled) trip count.
    const int trips = (min(TILE, total - base) + NPART - 1) / NPART;
    T tx = 0, ty = 0, tz = 0, tp = 0;
#pragma unroll 8
    for (int k = 0; k < trips; ++k) {
      const V4<T> s = tile[rbase + k * RSTEP];
      const T dx = s.x - xi.x;
      const T dy = s.y - xi.y;
      const T dz = s.z - xi.z;
      const T r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
      T rinv = fast_rsqrt(r2);
      if (SAFE) rinv = r2 > T(0) ? rinv : T(0);
      const T mr = s.w * rinv;
      const T mr3 = mr * (rinv * rinv);
      tx = fma(mr3, dx, tx);
      ty = fma(mr3, dy, ty);
      tz = fma(mr3, dz, tz);
      if (POT) tp += mr;
    }
    if (COMP) {
      kahan_add(sx, kx, tx);
      kahan_add(sy, ky, ty);
      kahan_add(sz, kz, tz);
      if (POT) kahan_add(sp, kp, tp);
    } else {
      sx += tx;
      sy += ty;
      sz += tz;
      if (POT) sp += tp;
    }
  }
  // partitions reduced pairwise in LDS (fixed tree order, every thread takes part)
  __syncthreads();
  tile[threadIdx.x] = V4<T>{sx - kx, sy - ky, sz - kz, sp - kp};
  __syncthreads();
#pragma unroll
  for (int w = NPART / 2; w >= 1; w >>= 1) {
    if (part < w) {
      const V4<T> a = tile[threadIdx.x];
      const V4<T> b = tile[threadIdx.x + w * WT];
      tile[threadIdx.x] = V4<T>{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w};
    }
    __syncthreads();
  }
  if (part == 0 && i < cnt) {
    const V4<T> r = tile[i];
    T ax = r.x, ay = r.y, az = r.z, ap = r.w;
    if (part_out >= 0) {
      if (part_out >= items.num_part) return;
      T* dst = items.partials + (static_cast<int64_t>(part_out) * leaf_size + i) * 4;
      dst[0] = ax;
      dst[1] = ay;
      dst[2] = az;
      dst[3] = ap;
      return;
    }
    if (POT && eps2 > T(0)) ap -= xi.w * fast_rsqrt(eps2);  // remove the softened self term
    const int64_t row = out_index ? out_index[first + i] : first + i;
    if (accumulate) {
      acc[3 * row + 0] += G * ax;
      acc[3 * row + 1] += G * ay;
      acc[3 * row + 2] += G * az;
      if (POT) pot[row] += -G * ap;
    } else {
      acc[3 * row + 0] = G * ax;
      acc[3 * row + 1] = G * ay;
      acc[3 * row + 2] = G * az;
      if (POT) pot[row] = -G * ap;
    }
  }
}

// Leaves split over several items: one workgroup per split leaf.  Target i
// of the leaf has nsub = 256 / leaf_size lanes; lane j adds the partials of
// items p0 + j, p0 + j + nsub, ... (compensated), and lane 0 then adds the
// nsub lane sums in lane order: a fixed summation order for every run.  The
// result gets what the single-item path applies (self potential, G, output row).
template <typename T, bool POT>
__global__ __launch_bounds__(kThreads) void p2p_reduce_kernel(const T* __restrict__ pos4, int32_t n, int32_t leaf_size,
                                                              const int32_t* __restrict__ split_leaf,
                                                              const int32_t* __restrict__ p0,
                                                              const int32_t* __restrict__ p1,
                                                              const T* __restrict__ partials, T eps2, T G,
                                                              const int32_t* __restrict__ out_index,
                                                              T* __restrict__ acc, T* __restrict__ pot,
                                                              int32_t accumulate) {
  __shared__ T red[4][kThreads];
  const int k = blockIdx.x;
  const int nsub = kThreads / leaf_size;
  const int i = threadIdx.x % leaf_size, j = threadIdx.x / leaf_size;
  T s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
  if (j < nsub) {
    const int a = p0[k], b = p1[k];
    for (int p = a + j; p < b; p += nsub) {
      const T* src = partials + (static_cast<int64_t>(p) * leaf_size + i) * 4;
      for (int q = 0; q < 4; ++q) kahan_add(s[q], c[q], src[q]);
    }
  }
  for (int q = 0; q < 4; ++q) red[q][threadIdx.x] = s[q] - c[q];
  __syncthreads();
  const int first = split_leaf[k] * leaf_size;
  if (j != 0 || first + i >= n) return;
  T t[4] = {0, 0, 0, 0}, e[4] = {0, 0, 0, 0};
  for (int jj = 0; jj < nsub; ++jj)
    for (int q = 0; q < 4; ++q) kahan_add(t[q], e[q], red[q][jj * leaf_size + i]);
  for (int q = 0; q < 4; ++q) t[q] -= e[q];
  T ap = t[3];
  if (POT && eps2 > T(0)) ap -= pos4[4 * static_cast<int64_t>(first + i) + 3] * fast_rsqrt(eps2);
  const int64_t row = out_index ? out_index[first + i] : first + i;
  if (accumulate) {
    acc[3 * row + 0] += G * t[0];
    acc[3 * row + 1] += G * t[1];
    acc[3 * row + 2] += G * t[2];
    if (POT) pot[row] += -G * ap;
  } else {
    acc[3 * row + 0] = G * t[0];
    acc[3 * row + 1] = G * t[1];
    acc[3 * row + 2] = G * t[2];
    if (POT) pot[row] = -G * ap;
  }
}

template <typename T, int WT>
int32_t launch_p2p_wt(hipStream_t s, const T* pos4, int32_t n, int32_t leaf_size, int32_t num_blocks,
                      const int32_t* near_off, const int32_t* near_src, T eps2, T G, const int32_t* out_index,
                      T* acc, T* pot, bool comp, bool accumulate, P2PItems<T> items) {
  const bool want_pot = pot != nullptr;
  const bool safe = !(eps2 > T(0));
  // the common cases (softened, compensated, leaf_size == WT) get the compile-time
  // slot decode and leaf-granular fetch; every other combination the general one
  // float32 leaves of 8 on the fast path run in narrow workgroups
  constexpr int kFastThreads = (sizeof(T) == 4 && WT == 8) ? kNarrowThreads : kThreads;
#define JFM_P2P_LAUNCH(POT, COMP, SAFE)                                                                      \
  if (COMP && !SAFE && leaf_size == WT && (sizeof(T) == 8 || n <= kBufMaxRows))                              \
    p2p_kernel<T, WT, POT, COMP, SAFE, true, kFastThreads><<<num_blocks, kFastThreads, 0, s>>>(              \
        pos4, n, leaf_size, near_off, near_src, eps2, G, out_index, acc, pot, accumulate ? 1 : 0, items);    \
  else                                                                                                       \
    p2p_kernel<T, WT, POT, COMP, SAFE, false, kThreads><<<num_blocks, kThreads, 0, s>>>(                      \
        pos4, n, leaf_size, near_off, near_src, eps2, G, out_index, acc, pot, accumulate ? 1 : 0, items)
  if (want_pot) {
    if (comp) {
      if (safe) JFM_P2P_LAUNCH(true, true, true); else JFM_P2P_LAUNCH(true, true, false);
    } else {
      if (safe) JFM_P2P_LAUNCH(true, false, true); else JFM_P2P_LAUNCH(true, false, false);
    }
  } else {
    if (comp) {
      if (safe) JFM_P2P_LAUNCH(false, true, true); else JFM_P2P_LAUNCH(false, true, false);
    } else {
      if (safe) JFM_P2P_LAUNCH(false, false, true); else JFM_P2P_LAUNCH(false, false, false);
    }
  }
#undef JFM_P2P_LAUNCH
  return check_launch("jfm_p2p");
}

template <typename T>
int32_t launch_p2p(hipStream_t s, const T* pos4, int32_t n, int32_t leaf_size, int32_t num_leaves,
                   const int32_t* near_off, const int32_t* near_src, T eps2, T G, const int32_t* out_index, T* acc,
                   T* pot, bool comp, bool accumulate, P2PItems<T> items = P2PItems<T>{}) {
  if (leaf_size <= 8)
    return launch_p2p_wt<T, 8>(s, pos4, n, leaf_size, num_leaves, near_off, near_src, eps2, G, out_index, acc,
                               pot, comp, accumulate, items);
  if (leaf_size <= 16)
    return launch_p2p_wt<T, 16>(s, pos4, n, leaf_size, num_leaves, near_off, near_src, eps2, G, out_index, acc,
                                pot, comp, accumulate, items);
  if (leaf_size <= 32)
    return launch_p2p_wt<T, 32>(s, pos4, n, leaf_size, num_leaves, near_off, near_src, eps2, G, out_index, acc,
                                pot, comp, accumulate, items);
  if (leaf_size <= 64)
    return launch_p2p_wt<T, 64>(s, pos4, n, leaf_size, num_leaves, near_off, near_src, eps2, G, out_index, acc,
                                pot, comp, accumulate, items);
  if (leaf_size <= 128)
    return launch_p2p_wt<T, 128>(s, pos4, n, leaf_size, num_leaves, near_off, near_src, eps2, G, out_index, acc,
                                 pot, comp, accumulate, items);
  return launch_p2p_wt<T, 256>(s, pos4, n, leaf_size, num_leaves, near_off, near_src, eps2, G, out_index, acc,
                               pot, comp, accumulate, items);
}

}  // namespace
}  // namespace jfm

using namespace jfm;

extern "C" {

int32_t jfm_l2p(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t leaf_size,
                int32_t num_internal, const void* locals, const void* centers4, double G, const int32_t* out_index,
                void* acc, void* pot, int32_t accumulate);

int32_t jfm_p2p(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves, int32_t leaf_size,
                const int32_t* near_off, const int32_t* near_src, double eps2, double G, const int32_t* out_index,
                void* acc, void* pot, int32_t flags, int32_t accumulate) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_leaves != (n + leaf_size - 1) / leaf_size)
    return fail(JFM_ERR_ARG, "jfm_p2p: inconsistent sizes (leaf_size must be 1..256)");
  if (!pos4 || !near_off || !near_src || !acc) return fail(JFM_ERR_ARG, "jfm_p2p: NULL buffer");
  if (!(eps2 >= 0.0)) return fail(JFM_ERR_ARG, "jfm_p2p: eps2 must be >= 0");
  hipStream_t s = as_stream(stream);
  const bool comp = (flags & JFM_EVAL_COMPENSATED) != 0;
  void* potp = (flags & JFM_EVAL_POTENTIAL) ? pot : nullptr;
  if ((flags & JFM_EVAL_POTENTIAL) && !pot) return fail(JFM_ERR_ARG, "jfm_p2p: potential requested, pot is NULL");
  if (dtype == JFM_F32)
    return launch_p2p<float>(s, (const float*)pos4, n, leaf_size, num_leaves, near_off, near_src,
                             static_cast<float>(eps2), static_cast<float>(G), out_index, (float*)acc, (float*)potp,
                             comp, accumulate != 0);
  if (dtype == JFM_F64)
    return launch_p2p<double>(s, (const double*)pos4, n, leaf_size, num_leaves, near_off, near_src, eps2, G,
                              out_index, (double*)acc, (double*)potp, comp, accumulate != 0);
  return fail(JFM_ERR_ARG, "jfm_p2p: bad dtype");
}

int32_t jfm_p2p_items(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves, int32_t leaf_size,
                      int32_t num_items, const int32_t* item_leaf, const int32_t* item_s0, const int32_t* item_s1,
                      const int32_t* item_part, int32_t num_part, void* partials, const int32_t* near_off,
                      const int32_t* near_src, double eps2, double G, const int32_t* out_index, void* acc, void* pot,
                      int32_t flags, int32_t accumulate) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_leaves != (n + leaf_size - 1) / leaf_size || num_items < 1 ||
      num_part < 0)
    return fail(JFM_ERR_ARG, "jfm_p2p_items: inconsistent sizes");
  if (!pos4 || !near_off || !near_src || !acc || !item_leaf || !item_s0 || !item_s1 || !item_part ||
      (num_part > 0 && !partials))
    return fail(JFM_ERR_ARG, "jfm_p2p_items: NULL buffer");
  if (!(eps2 >= 0.0)) return fail(JFM_ERR_ARG, "jfm_p2p_items: eps2 must be >= 0");
  if ((flags & JFM_EVAL_POTENTIAL) && !pot) return fail(JFM_ERR_ARG, "jfm_p2p_items: pot is NULL");
  hipStream_t s = as_stream(stream);
  const bool comp = (flags & JFM_EVAL_COMPENSATED) != 0;
  void* potp = (flags & JFM_EVAL_POTENTIAL) ? pot : nullptr;
  if (dtype == JFM_F32)
    return launch_p2p<float>(s, (const float*)pos4, n, leaf_size, num_items, near_off, near_src,
                             static_cast<float>(eps2), static_cast<float>(G), out_index, (float*)acc, (float*)potp,
                             comp, accumulate != 0,
                             P2PItems<float>{item_leaf, item_s0, item_s1, item_part, (float*)partials, num_part});
  if (dtype == JFM_F64)
    return launch_p2p<double>(s, (const double*)pos4, n, leaf_size, num_items, near_off, near_src, eps2, G,
                              out_index, (double*)acc, (double*)potp, comp, accumulate != 0,
                              P2PItems<double>{item_leaf, item_s0, item_s1, item_part, (double*)partials, num_part});
  return fail(JFM_ERR_ARG, "jfm_p2p_items: bad dtype");
}

int32_t jfm_p2p_reduce(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t leaf_size, int32_t num_split,
                       const int32_t* split_leaf, const int32_t* split_p0, const int32_t* split_p1,
                       const void* partials, double eps2, double G, const int32_t* out_index, void* acc, void* pot,
                       int32_t flags, int32_t accumulate) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_split < 0) return fail(JFM_ERR_ARG, "jfm_p2p_reduce: sizes");
  if (num_split == 0) return JFM_OK;
  if (!pos4 || !split_leaf || !split_p0 || !split_p1 || !partials || !acc)
    return fail(JFM_ERR_ARG, "jfm_p2p_reduce: NULL buffer");
  const bool want_pot = (flags & JFM_EVAL_POTENTIAL) != 0;
  if (want_pot && !pot) return fail(JFM_ERR_ARG, "jfm_p2p_reduce: pot is NULL");
  hipStream_t s = as_stream(stream);
#define JFM_P2P_REDUCE(T, POT)                                                                                  \
  p2p_reduce_kernel<T, POT><<<num_split, kThreads, 0, s>>>((const T*)pos4, n, leaf_size, split_leaf, split_p0,   \
                                                           split_p1, (const T*)partials, static_cast<T>(eps2), \
                                                           static_cast<T>(G), out_index, (T*)acc, (T*)pot,     \
                                                           accumulate)
  if (dtype == JFM_F32) {
    if (want_pot) JFM_P2P_REDUCE(float, true); else JFM_P2P_REDUCE(float, false);
  } else if (dtype == JFM_F64) {
    if (want_pot) JFM_P2P_REDUCE(double, true); else JFM_P2P_REDUCE(double, false);
  } else {
    return fail(JFM_ERR_ARG, "jfm_p2p_reduce: bad dtype");
  }
#undef JFM_P2P_REDUCE
  return check_launch("jfm_p2p_reduce");
}

int32_t jfm_evaluate(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t num_leaves,
                     int32_t leaf_size, int32_t num_internal, const int32_t* near_off, const int32_t* near_src,
                     const void* locals, const void* centers4, double eps2, double G, const int32_t* out_index,
                     void* acc, void* pot, int32_t flags) {
  const bool near = (flags & JFM_EVAL_NEAR) != 0;
  const bool far = (flags & JFM_EVAL_FAR) != 0;
  if (!near && !far) return fail(JFM_ERR_ARG, "jfm_evaluate: neither JFM_EVAL_NEAR nor JFM_EVAL_FAR set");
  if (num_internal != num_leaves - 1) return fail(JFM_ERR_ARG, "jfm_evaluate: num_internal must be L-1");
  if ((flags & JFM_EVAL_POTENTIAL) && !pot) return fail(JFM_ERR_ARG, "jfm_evaluate: pot is NULL");
  void* potp = (flags & JFM_EVAL_POTENTIAL) ? pot : nullptr;
  if (far) {
    int32_t rc = jfm_l2p(stream, dtype, p, pos4, n, leaf_size, num_internal, locals, centers4, G, out_index, acc,
                         potp, 0);
    if (rc != JFM_OK) return rc;
  }
  if (near)
    return jfm_p2p(stream, dtype, pos4, n, num_leaves, leaf_size, near_off, near_src, eps2, G, out_index, acc, potp,
                   flags, far ? 1 : 0);
  return JFM_OK;
}

}  // extern "C"
