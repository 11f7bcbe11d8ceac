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
//      f = (c * inv_r^3) * G, F_i += f * (x_j - x_i),
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

// 1/sqrt(r2), 0 for r2 <= 0.  float32: v_rsq_f32 (about 1 ulp); float64:
// v_rsq_f64 refined by one third-order step of the Newton iteration for
// r^-1/2, y <- y + y e (1/2 + 3 e / 8) with e = 1 - r2 y^2 (the relative
// error is cubed).  A deterministic function of r2 alone, so the two ends of
// a pair get the same value.
template <typename T>
__device__ __forceinline__ T inv_sqrt(T r2) {
  if constexpr (sizeof(T) == 4) {
    return r2 > T(0) ? __builtin_amdgcn_rsqf(r2) : T(0);
  } else {
    const T y = __builtin_amdgcn_rsq(r2);
    const T e = fma(-(r2 * y), y, T(1));
    const T z = fma(y * e, fma(e, T(0.375), T(0.5)), y);
    return r2 > T(0) ? z : T(0);
  }
}

template <typename T>
__device__ __forceinline__ void kahan_add(T& sum, T& comp, T x) {
  const T y = x - comp;
  const T t = sum + y;
  comp = (t - sum) - y;
  sum = t;
}

// One workgroup per target leaf.  Thread t holds target slot t mod WT
// (WT = leaf width rounded up to a power of two) and the source partition
// t / WT; sources (self leaf, then the near leaves, leaf_size slots each)
// stream through a kTile-entry LDS tile (kLoads loads per thread; the
// next tile's positions and the near-list entries of the one after it are
// fetched into registers during the current one), partition q taking the
// tile entries q, q + NP, q + 2 NP, ... (NP = THREADS / WT partitions; the
// partitions of a wave read neighbouring entries, so the LDS reads hit
// distinct banks); the tile's memory then holds the partition sums.
// Empty slots hold mass 0 at the origin and the self pair has d = 0, so
// neither needs a branch: both add exactly zero.  Per-tile partials are
// folded with Kahan compensation and the partitions are added in fixed order.
template <typename T, bool WEIGHTED, int THREADS>
__global__ __launch_bounds__(THREADS) void mutual_p2p_kernel(
    const T* __restrict__ pos4, int32_t n, int32_t leaf_size, int32_t wt_log2, int32_t leaf_log2,
    const int32_t* __restrict__ near_off, const int32_t* __restrict__ near_src,
    const int32_t* __restrict__ rung_sorted, const T* __restrict__ level_w, int32_t nlev, T eps2, T G,
    const int32_t* __restrict__ out_index, T* __restrict__ force) {
  constexpr int kLoads = 2;
  constexpr int kTile = THREADS * kLoads;
  static_assert(kTile * 4 >= THREADS * 3, "the tile must hold the reduction");
  __shared__ V4<T> sp[kTile];
  __shared__ int32_t sr[WEIGHTED ? kTile : 1];
  __shared__ T lw[kMaxLevels];
  T(*red)[3] = reinterpret_cast<T(*)[3]>(sp);  // the tile, reused for the partition sums
  const int leaf = xcd_contiguous(blockIdx.x, gridDim.x);  // neighbouring leaves share an XCD's L2
  const int t = threadIdx.x;
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
    if constexpr (WEIGHTED) ri = rung_sorted[i];
  }
  if constexpr (WEIGHTED) {
    if (t < nlev) lw[t] = level_w[t];
  }
  const int b = near_off[leaf];
  const int nsrc = (near_off[leaf + 1] - b + 1) * leaf_size;
  T fx = T(0), fy = T(0), fz = T(0), cx = T(0), cy = T(0), cz = T(0);
  // two-stage fetch while a tile is in use: the positions of the next tile
  // (from source indices resolved one tile earlier) and the near-list
  // entries of the tile after it
  V4<T> nx[kLoads];
  int nr[kLoads];
  int jl[kLoads];
  auto resolve = [&](int base) {
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      const int s = base + q * THREADS + t;
      int j = -1;
      if (s < nsrc) {
        const int ql = leaf_log2 >= 0 ? s >> leaf_log2 : s / leaf_size;
        const int src_leaf = ql == 0 ? leaf : near_src[b + ql - 1];
        const int jj = src_leaf * leaf_size + (s - ql * leaf_size);
        if (jj < n) j = jj;
      }
      jl[q] = j;
    }
  };
  auto fetch = [&]() {
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      nx[q] = jl[q] >= 0 ? load4(pos4, jl[q]) : V4<T>{T(0), T(0), T(0), T(0)};
      if constexpr (WEIGHTED) nr[q] = jl[q] >= 0 ? rung_sorted[jl[q]] : 0;
    }
  };
  resolve(0);
  fetch();
  resolve(kTile);
  for (int base = 0; base < nsrc; base += kTile) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kLoads; ++q) {
      sp[q * THREADS + t] = nx[q];
      if constexpr (WEIGHTED) sr[q * THREADS + t] = nr[q];
    }
    __syncthreads();
    if (base + kTile < nsrc) {
      fetch();
      resolve(base + 2 * kTile);
    }
    T px = T(0), py = T(0), pz = T(0);
    const int lim = min(kTile, nsrc - base);
#pragma unroll 4
    for (int e = part; e < lim; e += np) {
      const V4<T> xj = sp[e];
      const T dx = xj.x - xi.x, dy = xj.y - xi.y, dz = xj.z - xi.z;
      const T r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, eps2)));
      const T inv = inv_sqrt(r2);
      const T inv3 = inv * inv * inv;
      T c = xi.w * xj.w;
      if constexpr (WEIGHTED) c = c * lw[min(max(max(ri, sr[e]), 0), nlev - 1)];
      const T f = (c * inv3) * G;
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
    const int64_t r = out_index ? out_index[i] : i;
    force[3 * r + 0] = sx;
    force[3 * r + 1] = sy;
    force[3 * r + 2] = sz;
  }
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

template <typename T, bool W, int TH>
void launch_mutual(const MutualArgs<T>& a, int num_leaves, hipStream_t s) {
  mutual_p2p_kernel<T, W, TH><<<num_leaves, TH, 0, s>>>(a.pos4, a.n, a.leaf_size, a.wt_log2, a.leaf_log2,
                                                        a.near_off, a.near_src, a.rung_sorted, a.level_w, a.nlev,
                                                        a.eps2, a.G, a.out_index, a.force);
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
  if (n < 1 || leaf_size < 1 || leaf_size > kMutualThreads || num_leaves != (n + leaf_size - 1) / leaf_size)
    return fail(JFM_ERR_ARG, "jfm_mutual_p2p: inconsistent sizes (leaf_size must be 1..256)");
  if (!pos4 || !near_off || !near_src || !force) return fail(JFM_ERR_ARG, "jfm_mutual_p2p: NULL buffer");
  const bool weighted = level_weights != nullptr;
  if (weighted && (!rung_sorted || num_levels < 1 || num_levels > kMaxLevels))
    return fail(JFM_ERR_ARG, "jfm_mutual_p2p: weights need rungs and 1..64 levels");
  int wt_log2 = 0;
  while ((1 << wt_log2) < leaf_size) ++wt_log2;
  const int leaf_log2 = (1 << wt_log2) == leaf_size ? wt_log2 : -1;
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, {
    const T* lw = weighted ? (const T*)level_weights : nullptr;
    const int32_t* rs = weighted ? rung_sorted : nullptr;
    const int nl = weighted ? num_levels : 1;
    const MutualArgs<T> a{(const T*)pos4, n, leaf_size, wt_log2, leaf_log2, near_off, near_src, rs, lw, nl,
                          static_cast<T>(eps2), static_cast<T>(G), out_index, (T*)force};
    if (weighted)
      launch_mutual<T, true, kMutualThreads>(a, num_leaves, s);
    else
      launch_mutual<T, false, kMutualThreads>(a, num_leaves, s);
  });
  return check_launch("jfm_mutual_p2p");
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
