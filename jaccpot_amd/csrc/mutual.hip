// Momentum-conserving (mutual) lane: rung-weighted forces on the fixed topology.
//
// Counterpart of the reference's mutual FMM (jaccpot/mutual/force.py:985-1068,
// nearfield.py, farfield.py).  The topology here is already symmetric: the
// dual-tree walk visits every unordered node pair once with a symmetric MAC
// and emits both directions (topology.cpp), so
//
//  * near field: each target sums over its self leaf and near leaves with a
//    pair force built so that f_ij == -f_ji to the last bit
//      c = m_i * m_j (commutative), c *= w[max(rung_i, rung_j)],
//      f = c * inv_r^3, F_i += f * (x_j - x_i), and F_i *= G once per target,
//    where r^2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps^2))) is the same
//    for d and -d; the pair is evaluated from both ends, with bit-negated
//    results, instead of being scattered (no atomics, deterministic);
//  * far field: the M2L of pair (t <- s) is scaled by
//    w[max(rung_t, rung_s)] with the cell rung the maximum over the cell's
//    particles (jfm_m2l_flat_weighted); the order-p L2P force on a cell is
//    the gradient of the pair energy truncated at n + k <= p - 1, which is
//    symmetric in the two cells, so F_A + F_B cancels algebraically;
//  * forces are accumulated and divided by the mass once, at the end.
//
// Weights (level_weights[k], k = 0..num_levels-1) select or combine block-step
// interaction levels; rungs outside [0, num_levels) are clamped, as the
// reference clips them (force.py:950-955).

#include <hip/hip_runtime.h>

#include "launch.hpp"
#include "rows.hpp"
#include "xcd.hpp"

namespace jfm {
namespace {

constexpr int kMutualThreads = 256;
constexpr int kMaxLevels = 64;

__global__ __launch_bounds__(256) void leaf_rungs_kernel(int32_t n, int32_t num_leaves, int32_t leaf_size,
                                                         int32_t num_internal, const int32_t* __restrict__ perm,
                                                         const int32_t* __restrict__ rung,
                                                         int32_t* __restrict__ rung_sorted,
                                                         int32_t* __restrict__ node_rung) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= num_leaves) return;
  int mx = -1;
  const int a = l * leaf_size;
  const int b = min(a + leaf_size, n);
  for (int j = a; j < b; ++j) {
    const int r = rung[perm[j]];
    rung_sorted[j] = r;
    mx = max(mx, r);
  }
  node_rung[num_internal + l] = mx;
}

__global__ __launch_bounds__(256) void node_rungs_kernel(const int32_t* __restrict__ nodes, int32_t count,
                                                         const int32_t* __restrict__ left,
                                                         const int32_t* __restrict__ right,
                                                         int32_t* __restrict__ node_rung) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const int u = nodes[k];
  node_rung[u] = max(node_rung[left[u]], node_rung[right[u]]);
}

// r2^(-3/2); with SAFE, 0 for r2 <= 0 (no softening: the self pair and a
// zero-mass slot at a target's own position).  float32: v_rsq_f32 (about
// 1 ulp) cubed.  float64: the v_rsq_f64 estimate y refined in the cube
// itself, r2^(-3/2) = y^3 (1 - e)^(-3/2) with e = 1 - r2 y^2, expanded to
// y^3 (1 + 3e/2 + 15e^2/8) (the dropped terms are (35/16) e^3 relative).  A
// deterministic function of r2 alone, so the two ends of a pair get the
// same value.
// k holds 15/8 in a vector register (see vgpr_const): an fma with two
// constants that are not inline operands would need a copy per call.
template <typename T, bool SAFE>
__device__ __forceinline__ T inv_cube(T r2, T k) {
  T z;
  if constexpr (sizeof(T) == 4) {
    const T y = __builtin_amdgcn_rsqf(r2);
    z = y * y * y;
  } else {
    const T y = __builtin_amdgcn_rsq(r2);
    const T y2 = y * y;
    const T e = fma(-r2, y2, T(1));
    const T y3 = y2 * y;
    z = fma(y3 * e, fma(e, k, T(1.5)), y3);
  }
  if constexpr (SAFE) z = r2 > T(0) ? z : T(0);
  return z;
}

// A loop-invariant constant kept in a vector register: the empty asm
// statement hides the value from the compiler, which then neither folds it
// into an instruction nor rematerialises it inside a loop.
template <typename T>
__device__ __forceinline__ T vgpr_const(T v) {
  asm volatile("" : "+v"(v));
  return v;
}

template <typename T>
__device__ __forceinline__ void kahan_add(T& sum, T& comp, T x) {
  const T y = x - comp;
  const T t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}

// Work items over the source slots of the target leaves (whole source
// leaves, a bounded number of slots each, so the few leaves with very long
// near lists are spread over the GPU instead of running alone at the end of
// the launch): item k covers slots [s0[k], s1[k]) of leaf leaf[k]'s sequence
// (self leaf, then its near leaves, leaf_size slots each).  part[k] < 0: the
// item is its leaf's only one and writes the force rows; otherwise it writes
// its targets' partial sums to partials[(part[k] * leaf_size + slot) * 4 + c]
// (c < 3; num_part rows of room) and mutual_reduce_kernel adds a split leaf's
// partials in item order.  leaf == nullptr: one item per leaf over all of its
// slots.
template <typename T>
struct MutualItems {
  const int32_t* leaf;
  const int32_t* s0;
  const int32_t* s1;
  const int32_t* part;
  T* partials;
  int32_t num_part;
};

// One workgroup per work item.  Thread t holds target slot t mod WT
// (WT = leaf width rounded up to a power of two) and the source partition
// t / WT; the item's sources stream through a kTile-entry LDS tile, partition
// q taking the tile entries q, q + NP, q + 2 NP, ... (NP = THREADS / WT
// partitions; the partitions of a wave read neighbouring entries, so the LDS
// reads hit distinct banks); the tile's memory then holds the partition sums.
//
// Fetch pipeline (kLoads slots per thread and tile): the near-list entries of
// tile k + 2 and the positions (and rungs) of tile k + 1 are loaded while
// tile k is in use.  Every load is unconditional, its index clamped into
// range, and no loaded value is tested before the next tile boundary: a list
// entry is decoded one tile after it is loaded, and a slot past the item's
// end or past the last particle is flagged from values already in registers,
// loads particle 0 and gets mass zero when the tile is written.
// So the only wait for memory is at the tile store, one tile after the loads
// were issued, and every partition runs the same wave-uniform trip count.
// The self pair has d = 0 and a zero-mass source adds exactly zero (r^2 > 0
// with softening; without it SAFE maps r^2 == 0 to zero), so neither needs a
// branch.  Per-tile partials are folded with Kahan
// compensation and the partitions are added in fixed order.  Sums are scaled
// by G once per target (partials unscaled).  Rungs are clamped into
// [0, nlev) as they are loaded, so a pair's level index is one max.
// SAFE: guard r^2 == 0 (no softening).  LEAF > 0: leaf_size == LEAF, a power
// of two known at compile time (the partition stride then goes into the LDS
// reads' immediate offsets); LEAF == 0: any leaf size, decoded at run time.
template <typename T, bool WEIGHTED, bool SAFE, int LEAF, int THREADS>
__global__ __launch_bounds__(THREADS) void mutual_p2p_kernel(
    const T* __restrict__ pos4, int32_t n, int32_t leaf_size, int32_t wt_log2, int32_t leaf_log2,
    const int32_t* __restrict__ near_off, const int32_t* __restrict__ near_src,
    const int32_t* __restrict__ rung_sorted, const T* __restrict__ level_w, int32_t nlev, T eps2, T G,
    const int32_t* __restrict__ out_index, T* __restrict__ force, MutualItems<T> items) {
  constexpr int kLoads = sizeof(T) == 4 ? 4 : 2;  // 16 KiB of positions per tile either way
  constexpr int kTile = THREADS * kLoads;
  static_assert(kTile * 4 >= THREADS * 3, "the tile must hold the reduction");
  __shared__ V4<T> sp[kTile];
  __shared__ int32_t sr[WEIGHTED ? kTile : 1];
  __shared__ T lw[kMaxLevels];
  T(*red)[3] = reinterpret_cast<T(*)[3]>(sp);  // the tile, reused for the partition sums
  const int item = xcd_contiguous(blockIdx.x, gridDim.x);  // neighbouring leaves share an XCD's L2
  const int leaf = items.leaf ? items.leaf[item] : item;
  const int t = threadIdx.x;
  static_assert(LEAF == 0 || (LEAF & (LEAF - 1)) == 0, "compile-time leaves are powers of two");
  if constexpr (LEAF > 0) {
    wt_log2 = __builtin_ctz(LEAF);
    leaf_log2 = wt_log2;
    leaf_size = LEAF;
  }
  const int wt = 1 << wt_log2;
  const int slot = t & (wt - 1);
  const int part = t >> wt_log2;
  const int np = THREADS >> wt_log2;  // source partitions
  const int i = leaf * leaf_size + slot;
  const bool live = slot < leaf_size && i < n;
  V4<T> xi{T(0), T(0), T(0), T(0)};
  int ri = 0;
  if (live) {
    xi = load4(pos4, i);
    if constexpr (WEIGHTED) ri = min(max(rung_sorted[i], 0), nlev - 1);
  }
  if constexpr (WEIGHTED) {
    if (t < nlev) lw[t] = level_w[t];
  }
  const int b = near_off[leaf];
  const int nnear = near_off[leaf + 1] - b;
  const int start = items.leaf ? items.s0[item] : 0;
  const int total = items.leaf ? items.s1[item] : (nnear + 1) * leaf_size;
  const int part_out = items.leaf ? items.part[item] : -1;
  const int32_t* list = nnear > 0 ? near_src + b : near_off;  // an empty list reads a valid address
  const int hi = max(nnear - 1, 0);
  auto seq_leaf = [&](int s) { return leaf_log2 >= 0 ? s >> leaf_log2 : s / leaf_size; };
  int lst[kLoads];  // raw near-list entries, decoded one tile later
  V4<T> nx[kLoads];
  bool ok[kLoads];
  int nr[kLoads];
  auto list_entries = [&](int base) {
#pragma unroll
    for (int q = 0; q < kLoads; ++q) lst[q] = list[min(max(seq_leaf(base + q * THREADS + t) - 1, 0), hi)];
  };
  auto positions = [&](int base) {
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      const int s = base + q * THREADS + t;
      const int ql = seq_leaf(s);
      const int g = (ql == 0 ? leaf : lst[q]) * leaf_size + (s - ql * leaf_size);
      ok[q] = s < total && g < n;
      const int gi = ok[q] ? g : 0;
      nx[q] = load4(pos4, gi);
      if constexpr (WEIGHTED) nr[q] = rung_sorted[gi];
    }
  };
  list_entries(start);
  positions(start);
  list_entries(start + kTile);
  const T k158 = vgpr_const(T(1.875));
  T fx = T(0), fy = T(0), fz = T(0), cx = T(0), cy = T(0), cz = T(0);
  for (int base = start; base < total; base += kTile) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      V4<T> v = nx[q];
      if (!ok[q]) v.w = T(0);
      sp[q * THREADS + t] = v;
      if constexpr (WEIGHTED) sr[q * THREADS + t] = min(max(nr[q], 0), nlev - 1);
    }
    __syncthreads();
    // unconditional (past the end they give flagged slots): a branch here
    // would make the compiler wait for the loads at its join
    positions(base + kTile);
    list_entries(base + 2 * kTile);
    const int trips = (min(kTile, total - base) + np - 1) / np;
    T px = T(0), py = T(0), pz = T(0);
#pragma unroll 4
    for (int k = 0; k < trips; ++k) {
      const int e = part + k * np;
      const V4<T> xj = sp[e];
      const T dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
      const T r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
      T c = xi.w * xj.w;
      if constexpr (WEIGHTED) c = c * lw[max(ri, sr[e])];
      const T f = c * inv_cube<T, SAFE>(r2, k158);
      px = fma(f, dx, px);
      py = fma(f, dy, py);
      pz = fma(f, dz, pz);
    }
    kahan_add(fx, cx, px);
    kahan_add(fy, cy, py);
    kahan_add(fz, cz, pz);
  }
  __syncthreads();  // every partition is done with the tile
  red[t][0] = fx;
  red[t][1] = fy;
  red[t][2] = fz;
  __syncthreads();
  if (live && part == 0) {
    T sx = T(0), sy = T(0), sz = T(0);
    for (int q = 0; q < np; ++q) {
      sx += red[q * wt + slot][0];
      sy += red[q * wt + slot][1];
      sz += red[q * wt + slot][2];
    }
    if (part_out >= 0) {
      if (part_out < items.num_part) {
        T* dst = items.partials + (static_cast<int64_t>(part_out) * leaf_size + slot) * 4;
        dst[0] = sx;
        dst[1] = sy;
        dst[2] = sz;
        dst[3] = T(0);
      }
      return;
    }
    const int64_t r = out_index ? out_index[i] : i;
    force[3 * r + 0] = G * sx;
    force[3 * r + 1] = G * sy;
    force[3 * r + 2] = G * sz;
  }
}

// Split leaves: one thread per (split leaf, target slot) adds the leaf's item
// partials p0[k] .. p1[k]-1 in item order (Kahan) and writes G times the sum.
template <typename T>
__global__ __launch_bounds__(256) void mutual_reduce_kernel(int32_t n, int32_t leaf_size, int32_t num_split,
                                                            const int32_t* __restrict__ split_leaf,
                                                            const int32_t* __restrict__ split_p0,
                                                            const int32_t* __restrict__ split_p1,
                                                            const T* __restrict__ partials, T G,
                                                            const int32_t* __restrict__ out_index, T* __restrict__ force) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (g >= static_cast<int64_t>(num_split) * leaf_size) return;
  const int k = static_cast<int>(g / leaf_size);
  const int slot = static_cast<int>(g - static_cast<int64_t>(k) * leaf_size);
  const int i = split_leaf[k] * leaf_size + slot;
  if (i >= n) return;
  T sum[3] = {T(0), T(0), T(0)}, comp[3] = {T(0), T(0), T(0)};
  for (int p = split_p0[k]; p < split_p1[k]; ++p) {
    const T* src = partials + (static_cast<int64_t>(p) * leaf_size + slot) * 4;
#pragma unroll
    for (int c = 0; c < 3; ++c) kahan_add(sum[c], comp[c], src[c]);
  }
  const int64_t r = out_index ? out_index[i] : i;
#pragma unroll
  for (int c = 0; c < 3; ++c) force[3 * r + c] = G * sum[c];
}

template <typename T>
__global__ __launch_bounds__(256) void mutual_combine_kernel(int32_t n, const T* __restrict__ masses,
                                                             const T* __restrict__ far_acc,
                                                             const T* __restrict__ near_force, T* __restrict__ acc,
                                                             T* __restrict__ far_force) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const T m = masses[r];
  for (int k = 0; k < 3; ++k) {
    const T ff = m * far_acc[3 * r + k];
    const T f = near_force[3 * r + k] + ff;
    if (far_force) far_force[3 * r + k] = ff;
    acc[3 * r + k] = m > T(0) ? f / m : T(0);
  }
}

int nblocks(int64_t n, int t) { return static_cast<int>((n + t - 1) / t); }

template <typename T>
struct MutualArgs {
  const T* pos4;
  int32_t n, leaf_size, wt_log2, leaf_log2;
  const int32_t* near_off;
  const int32_t* near_src;
  const int32_t* rung_sorted;
  const T* level_w;
  int32_t nlev;
  T eps2, G;
  const int32_t* out_index;
  T* force;
};

template <typename T, bool W, bool SAFE, int LEAF, int TH>
void launch_mutual_leaf(const MutualArgs<T>& a, const MutualItems<T>& items, int num_blocks, hipStream_t s) {
  mutual_p2p_kernel<T, W, SAFE, LEAF, TH><<<num_blocks, TH, 0, s>>>(a.pos4, a.n, a.leaf_size, a.wt_log2, a.leaf_log2,
                                                        a.near_off, a.near_src, a.rung_sorted, a.level_w, a.nlev,
                                                        a.eps2, a.G, a.out_index, a.force, items);
}

template <typename T, bool W, bool SAFE, int TH>
void launch_mutual(const MutualArgs<T>& a, const MutualItems<T>& items, int num_blocks, hipStream_t s) {
  switch (a.leaf_size) {
    case 8:
      return launch_mutual_leaf<T, W, SAFE, 8, TH>(a, items, num_blocks, s);
    case 16:
      return launch_mutual_leaf<T, W, SAFE, 16, TH>(a, items, num_blocks, s);
    case 32:
      return launch_mutual_leaf<T, W, SAFE, 32, TH>(a, items, num_blocks, s);
    case 64:
      return launch_mutual_leaf<T, W, SAFE, 64, TH>(a, items, num_blocks, s);
    default:
      return launch_mutual_leaf<T, W, SAFE, 0, TH>(a, items, num_blocks, s);
  }
}

// Shared by the per-leaf and the work-item entry points: argument checks,
// slot-decode constants, the near-field launch.
int32_t mutual_near(const char* name, hipStream_t s, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves,
                    int32_t leaf_size, int32_t num_blocks, const int32_t* item_leaf, const int32_t* item_s0,
                    const int32_t* item_s1, const int32_t* item_part, int32_t num_part, void* partials,
                    const int32_t* near_off, const int32_t* near_src, const int32_t* rung_sorted,
                    const void* level_weights, int32_t num_levels, double eps2, double G, const int32_t* out_index,
                    void* force) {
  if (n < 1 || leaf_size < 1 || leaf_size > kMutualThreads || num_leaves != (n + leaf_size - 1) / leaf_size)
    return fail(JFM_ERR_ARG, std::string(name) + ": inconsistent sizes (leaf_size must be 1..256)");
  if (!pos4 || !near_off || !near_src || !force) return fail(JFM_ERR_ARG, std::string(name) + ": NULL buffer");
  const bool weighted = level_weights != nullptr;
  if (weighted && (!rung_sorted || num_levels < 1 || num_levels > kMaxLevels))
    return fail(JFM_ERR_ARG, std::string(name) + ": weights need rungs and 1..64 levels");
  int wt_log2 = 0;
  while ((1 << wt_log2) < leaf_size) ++wt_log2;
  const int leaf_log2 = (1 << wt_log2) == leaf_size ? wt_log2 : -1;
  JFM_FOR_DTYPE(dtype, {
    const T* lw = weighted ? (const T*)level_weights : nullptr;
    const int32_t* rs = weighted ? rung_sorted : nullptr;
    const int nl = weighted ? num_levels : 1;
    const MutualArgs<T> a{(const T*)pos4, n, leaf_size, wt_log2, leaf_log2, near_off, near_src, rs, lw, nl,
                          static_cast<T>(eps2), static_cast<T>(G), out_index, (T*)force};
    const MutualItems<T> items{item_leaf, item_s0, item_s1, item_part, (T*)partials, num_part};
    const bool safe = !(a.eps2 > T(0));
    if (weighted && safe)
      launch_mutual<T, true, true, kMutualThreads>(a, items, num_blocks, s);
    else if (weighted)
      launch_mutual<T, true, false, kMutualThreads>(a, items, num_blocks, s);
    else if (safe)
      launch_mutual<T, false, true, kMutualThreads>(a, items, num_blocks, s);
    else
      launch_mutual<T, false, false, kMutualThreads>(a, items, num_blocks, s);
  });
  return check_launch(name);
}

}  // namespace
}  // namespace jfm

using namespace jfm;

extern "C" {

int32_t jfm_mutual_rungs(void* stream, int32_t n, int32_t num_leaves, int32_t leaf_size, const int32_t* perm,
                         const int32_t* rung, int32_t* rung_sorted, int32_t* node_rung) {
  if (n < 1 || leaf_size < 1 || num_leaves != (n + leaf_size - 1) / leaf_size)
    return fail(JFM_ERR_ARG, "jfm_mutual_rungs: inconsistent sizes");
  if (!perm || !rung || !rung_sorted || !node_rung) return fail(JFM_ERR_ARG, "jfm_mutual_rungs: NULL buffer");
  leaf_rungs_kernel<<<nblocks(num_leaves, 256), 256, 0, as_stream(stream)>>>(n, num_leaves, leaf_size, num_leaves - 1,
                                                                            perm, rung, rung_sorted, node_rung);
  return check_launch("jfm_mutual_rungs");
}

int32_t jfm_mutual_rungs_level(void* stream, const int32_t* nodes, int32_t count, const int32_t* left,
                               const int32_t* right, int32_t* node_rung) {
  if (count < 0) return fail(JFM_ERR_ARG, "jfm_mutual_rungs_level: negative count");
  if (count == 0) return JFM_OK;
  if (!nodes || !left || !right || !node_rung) return fail(JFM_ERR_ARG, "jfm_mutual_rungs_level: NULL buffer");
  node_rungs_kernel<<<nblocks(count, 256), 256, 0, as_stream(stream)>>>(nodes, count, left, right, node_rung);
  return check_launch("jfm_mutual_rungs_level");
}

int32_t jfm_mutual_p2p(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves,
                       int32_t leaf_size, const int32_t* near_off, const int32_t* near_src, const int32_t* rung_sorted,
                       const void* level_weights, int32_t num_levels, double eps2, double G, const int32_t* out_index,
                       void* force) {
  return mutual_near("jfm_mutual_p2p", as_stream(stream), dtype, pos4, n, num_leaves, leaf_size, num_leaves, nullptr,
                     nullptr, nullptr, nullptr, 0, nullptr, near_off, near_src, rung_sorted, level_weights,
                     num_levels, eps2, G, out_index, force);
}

int32_t jfm_mutual_p2p_items(void* stream, int32_t dtype, const void* pos4, int32_t n, int32_t num_leaves,
                             int32_t leaf_size, int32_t num_items, const int32_t* item_leaf, const int32_t* item_s0,
                             const int32_t* item_s1, const int32_t* item_part, int32_t num_part, void* partials,
                             int32_t num_split, const int32_t* split_leaf, const int32_t* split_p0,
                             const int32_t* split_p1, const int32_t* near_off, const int32_t* near_src,
                             const int32_t* rung_sorted, const void* level_weights, int32_t num_levels, double eps2,
                             double G, const int32_t* out_index, void* force) {
  if (num_items < 1 || num_part < 0 || num_split < 0)
    return fail(JFM_ERR_ARG, "jfm_mutual_p2p_items: need num_items >= 1, num_part >= 0, num_split >= 0");
  if (!item_leaf || !item_s0 || !item_s1 || !item_part || (num_part > 0 && !partials) ||
      (num_split > 0 && (!split_leaf || !split_p0 || !split_p1)))
    return fail(JFM_ERR_ARG, "jfm_mutual_p2p_items: NULL item buffer");
  hipStream_t s = as_stream(stream);
  const int32_t rc = mutual_near("jfm_mutual_p2p_items", s, dtype, pos4, n, num_leaves, leaf_size, num_items,
                                 item_leaf, item_s0, item_s1, item_part, num_part, partials, near_off, near_src,
                                 rung_sorted, level_weights, num_levels, eps2, G, out_index, force);
  if (rc != JFM_OK || num_split == 0) return rc;
  const int64_t rows = static_cast<int64_t>(num_split) * leaf_size;
  JFM_FOR_DTYPE(dtype, {
    mutual_reduce_kernel<T><<<nblocks(rows, 256), 256, 0, s>>>(n, leaf_size, num_split, split_leaf, split_p0,
                                                               split_p1, (const T*)partials, static_cast<T>(G),
                                                               out_index, (T*)force);
  });
  return check_launch("jfm_mutual_p2p_items (reduce)");
}

int32_t jfm_mutual_combine(void* stream, int32_t dtype, int32_t n, const void* masses, const void* far_acc,
                           const void* near_force, void* acc, void* far_force) {
  if (n < 0) return fail(JFM_ERR_ARG, "jfm_mutual_combine: negative count");
  if (n == 0) return JFM_OK;
  if (!masses || !far_acc || !near_force || !acc) return fail(JFM_ERR_ARG, "jfm_mutual_combine: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, {
    mutual_combine_kernel<T><<<nblocks(n, 256), 256, 0, s>>>(n, (const T*)masses, (const T*)far_acc,
                                                             (const T*)near_force, (T*)acc, (T*)far_force);
  });
  return check_launch("jfm_mutual_combine");
}

}  // extern "C"
