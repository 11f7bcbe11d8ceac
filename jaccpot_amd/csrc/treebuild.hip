// Device construction of the fixed static-radix topology: the same spec as the
// host builder in topology.cpp (the role yggdrax's build_tree, tree geometry
// and build_interactions_and_neighbors play for the reference, called from
// jaccpot/runtime/fmm_state.py:630-800 and runtime/_interaction_cache.py:523,564),
// with bit-identical output:
//
//  * bounds by exact min/max reduction; 63-bit Morton codes on the cube they
//    span (floor((x - lo) * 2^21 / extent), clamped);
//  * stable radix sort of (code, input index) -> perm;
//  * L = ceil(N/leaf) buckets, binary radix tree over the leaf keys (code of
//    the leaf's first particle, leaf index as tie-break), one thread per
//    internal node (Karras 2012);
//  * depth by the parent chain; level schedules by sorting (level, node) keys;
//  * centres of mass per leaf in slot order, internal nodes level by level
//    from their children, plain mean for a massless node; radius = exact max
//    distance from the centre to the node's particles;
//  * dual-tree walk breadth first: every frontier item (a, b) is expanded in
//    parallel exactly as the host's depth-first walk expands it (self item ->
//    three child items; MAC (r_a + r_b)^2 <= theta^2 d^2, d > 0 -> far pair in
//    both directions; two leaves -> near pair in both directions; otherwise
//    split the larger-radius non-leaf node).  The walk's set of pairs does not
//    depend on the expansion order; sorting the packed (target << 32 | source)
//    keys gives the host's CSR (targets ascending, sources ascending).
//
// Floating-point expressions are evaluated without contraction so every sum,
// product, quotient and square root rounds exactly as the host's.
//
// Buffers are the caller's; the sorts and scans take a caller-provided
// scratch (jfm_sort_scratch_bytes, jfm_scan_scratch_bytes).  A walk step is
// classify (what each item emits), an exclusive scan of the counts, and emit
// (each item writes at its offsets), so the caller sizes every output exactly.

#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <hipcub/device/device_radix_sort.hpp>
#include <hipcub/device/device_scan.hpp>

#include <cmath>
#include <cstdint>

#include "launch.hpp"

namespace jfm {
namespace {

constexpr int kMortonBits = 21;
constexpr int kBoundBlocks = 1024;
constexpr int kMaxTreeHeight = 128;

__device__ inline uint64_t spread3_dev(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x001F00000000FFFFull;
  v = (v | (v << 16)) & 0x001F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// ---------------------------------------------------------------------------
// bounds and Morton codes
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) bounds_partial_kernel(const double* __restrict__ pos, int64_t n,
                                                             double* __restrict__ partial, int32_t* __restrict__ bad) {
  __shared__ double red[6][256];
  double lo[3] = {pos[0], pos[1], pos[2]}, hi[3] = {pos[0], pos[1], pos[2]};
  int nonfinite = 0;
  for (int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x; i < n; i += int64_t(gridDim.x) * 256) {
    for (int a = 0; a < 3; ++a) {
      const double v = pos[3 * i + a];
      nonfinite |= !isfinite(v);
      lo[a] = fmin(lo[a], v);
      hi[a] = fmax(hi[a], v);
    }
  }
  if (nonfinite) atomicOr(bad, 1);
  for (int a = 0; a < 3; ++a) {
    red[a][threadIdx.x] = lo[a];
    red[3 + a][threadIdx.x] = hi[a];
  }
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fmin(red[a][threadIdx.x], red[a][threadIdx.x + w]);
        red[3 + a][threadIdx.x] = fmax(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) bounds_final_kernel(const double* __restrict__ partial, int blocks,
                                                           double* __restrict__ bounds6) {
  __shared__ double red[6][256];
  for (int a = 0; a < 6; ++a) red[a][threadIdx.x] = partial[a];
  for (int b = threadIdx.x; b < blocks; b += 256)
    for (int a = 0; a < 3; ++a) {
      red[a][threadIdx.x] = fmin(red[a][threadIdx.x], partial[6 * b + a]);
      red[3 + a][threadIdx.x] = fmax(red[3 + a][threadIdx.x], partial[6 * b + 3 + a]);
    }
  __syncthreads();
  for (int w = 128; w >= 1; w >>= 1) {
    if (threadIdx.x < w)
      for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = fmin(red[a][threadIdx.x], red[a][threadIdx.x + w]);
        red[3 + a][threadIdx.x] = fmax(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + w]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) bounds6[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void __launch_bounds__(256) morton_kernel(const double* __restrict__ pos, int64_t n,
                                                     const double* __restrict__ bounds6, uint64_t* __restrict__ codes,
                                                     int32_t* __restrict__ index) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  const double lo[3] = {bounds6[0], bounds6[1], bounds6[2]};
  double extent = fmax(bounds6[3] - lo[0], fmax(bounds6[4] - lo[1], bounds6[5] - lo[2]));
  if (!(extent > 0.0)) extent = 1.0;
  const double scale = static_cast<double>(1 << kMortonBits) / extent;
  const double qmax = static_cast<double>((1 << kMortonBits) - 1);
  uint64_t q[3];
  for (int a = 0; a < 3; ++a) {
    double v = floor((pos[3 * i + a] - lo[a]) * scale);
    v = fmin(fmax(v, 0.0), qmax);
    q[a] = static_cast<uint64_t>(v);
  }
  codes[i] = (spread3_dev(q[0]) << 2) | (spread3_dev(q[1]) << 1) | spread3_dev(q[2]);
  index[i] = static_cast<int32_t>(i);
}

// ---------------------------------------------------------------------------
// radix tree over leaf keys
// ---------------------------------------------------------------------------

struct LeafKeys {
  const uint64_t* sorted;
  int32_t leaf, L;
  __device__ uint64_t key(int32_t l) const { return sorted[int64_t(l) * leaf]; }
  // common prefix of (code, index) keys: code bits, then index bits
  __device__ int delta(int32_t i, int32_t j) const {
    const uint64_t a = key(i), b = key(j);
    if (a != b) return __clzll(static_cast<long long>(a ^ b));
    return 64 + __clz(static_cast<int>(static_cast<uint32_t>(i ^ j)));
  }
  __device__ int delta_safe(int32_t i, int32_t j) const { return (j < 0 || j >= L) ? -1 : delta(i, j); }
};

__global__ void __launch_bounds__(256) radix_tree_kernel(LeafKeys keys, int64_t n, int32_t* __restrict__ parent,
                                                         int32_t* __restrict__ left, int32_t* __restrict__ right,
                                                         int32_t* __restrict__ range) {
  const int32_t L = keys.L, I = L - 1;
  const int32_t t = blockIdx.x * 256 + threadIdx.x;
  if (t < L) {  // leaf: its own bucket of slots
    const int32_t node = I + t;
    range[2 * node] = static_cast<int32_t>(int64_t(t) * keys.leaf);
    range[2 * node + 1] = static_cast<int32_t>(min<int64_t>((int64_t(t) + 1) * keys.leaf, n) - 1);
  }
  if (t == 0) parent[0] = -1;
  if (t >= I) return;
  const int32_t i = t;
  int32_t first, last;
  if (i == 0) {
    first = 0;
    last = L - 1;
  } else {
    const int d = (keys.delta_safe(i, i + 1) - keys.delta_safe(i, i - 1)) > 0 ? 1 : -1;
    const int dmin = keys.delta_safe(i, i - d);
    int32_t lmax = 2;
    while (keys.delta_safe(i, i + lmax * d) > dmin) lmax *= 2;
    int32_t len = 0;
    for (int32_t s = lmax / 2; s >= 1; s /= 2)
      if (keys.delta_safe(i, i + (len + s) * d) > dmin) len += s;
    const int32_t j = i + len * d;
    first = min(i, j);
    last = max(i, j);
  }
  const int common = keys.delta(first, last);
  int32_t split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    const int32_t nw = split + step;
    if (nw < last && keys.delta(first, nw) > common) split = nw;
  } while (step > 1);
  const int32_t lc = (split == first) ? I + split : split;
  const int32_t rc = (split + 1 == last) ? I + split + 1 : split + 1;
  left[i] = lc;
  right[i] = rc;
  parent[lc] = i;
  parent[rc] = i;
  range[2 * i] = static_cast<int32_t>(int64_t(first) * keys.leaf);
  range[2 * i + 1] = static_cast<int32_t>(min<int64_t>((int64_t(last) + 1) * keys.leaf, n) - 1);
}

// depth by the parent chain; a chain longer than kMaxTreeHeight flags an error
__global__ void __launch_bounds__(256) depth_kernel(int32_t nodes, const int32_t* __restrict__ parent,
                                                    int32_t* __restrict__ depth, int32_t* __restrict__ max_depth,
                                                    int32_t* __restrict__ bad) {
  const int32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nodes) return;
  int32_t d = 0, j = k;
  while (parent[j] >= 0 && d <= kMaxTreeHeight) {
    j = parent[j];
    ++d;
  }
  if (d > kMaxTreeHeight) atomicOr(bad, 2);
  depth[k] = d;
  atomicMax(max_depth, d);
}

// (level, node) keys: M2M order = internal nodes deepest first, L2L order =
// non-root nodes shallowest first; ascending node index within a level.
__global__ void __launch_bounds__(256) level_keys_kernel(int32_t nodes, int32_t I, const int32_t* __restrict__ depth,
                                                         const int32_t* __restrict__ max_depth,
                                                         uint64_t* __restrict__ m2m_keys,
                                                         uint64_t* __restrict__ l2l_keys) {
  const int32_t k = blockIdx.x * 256 + threadIdx.x;
  if (k >= nodes) return;
  const int32_t D = *max_depth;
  if (k < I) m2m_keys[k] = (uint64_t(D - depth[k]) << 32) | uint32_t(k);
  if (k >= 1) l2l_keys[k - 1] = (uint64_t(depth[k] - 1) << 32) | uint32_t(k);
}

// ---------------------------------------------------------------------------
// geometry
// ---------------------------------------------------------------------------

struct Particles {
  const double* pos;   // input order [n][3]
  const double* mass;  // input order [n]
  const int32_t* perm;
  __device__ double x(int64_t slot, int a) const { return pos[3 * int64_t(perm[slot]) + a]; }
  __device__ double m(int64_t slot) const { return mass[perm[slot]]; }
};

__device__ inline void mean_of(const Particles& P, int64_t a, int64_t b, double* c) {
  double s[3] = {0, 0, 0};
  for (int64_t i = a; i <= b; ++i)
    for (int k = 0; k < 3; ++k) s[k] += P.x(i, k);
  for (int k = 0; k < 3; ++k) c[k] = s[k] / static_cast<double>(b - a + 1);
}

__global__ void __launch_bounds__(256) leaf_geometry_kernel(Particles P, int32_t L, int32_t I,
                                                            const int32_t* __restrict__ range,
                                                            double* __restrict__ mass, double* __restrict__ center) {
  const int32_t l = blockIdx.x * 256 + threadIdx.x;
  if (l >= L) return;
  const int32_t node = I + l;
  const int64_t a = range[2 * node], b = range[2 * node + 1];
  double tot = 0.0, s[3] = {0, 0, 0};
  for (int64_t i = a; i <= b; ++i) tot += P.m(i);
  for (int64_t i = a; i <= b; ++i) {
    const double mi = P.m(i);
    for (int k = 0; k < 3; ++k) s[k] += mi * P.x(i, k);
  }
  mass[node] = tot;
  if (tot > 0.0) {
    for (int k = 0; k < 3; ++k) center[3 * node + k] = s[k] / tot;
  } else {
    mean_of(P, a, b, &center[3 * node]);
  }
}

__global__ void __launch_bounds__(256) internal_geometry_kernel(Particles P, int32_t count,
                                                                const int32_t* __restrict__ level_nodes,
                                                                const int32_t* __restrict__ left,
                                                                const int32_t* __restrict__ right,
                                                                const int32_t* __restrict__ range,
                                                                double* __restrict__ mass,
                                                                double* __restrict__ center) {
  const int32_t q = blockIdx.x * 256 + threadIdx.x;
  if (q >= count) return;
  const int32_t node = level_nodes[q];
  const int32_t lc = left[node], rc = right[node];
  const double ml = mass[lc], mr = mass[rc];
  const double tot = ml + mr;
  mass[node] = tot;
  if (tot > 0.0) {
    for (int k = 0; k < 3; ++k) center[3 * node + k] = (ml * center[3 * lc + k] + mr * center[3 * rc + k]) / tot;
  } else {
    mean_of(P, range[2 * node], range[2 * node + 1], &center[3 * node]);
  }
}

// tree-order copy of the positions, so the radius sweeps read contiguously
__global__ void __launch_bounds__(256) sorted_positions_kernel(Particles P, int64_t n, double* __restrict__ xs) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  for (int a = 0; a < 3; ++a) xs[3 * i + a] = P.x(i, a);
}

constexpr int kSmallRadius = 2048;  // nodes up to this many particles: one wave each

__device__ inline double dist2(const double* __restrict__ xs, int64_t i, double cx, double cy, double cz) {
  const double dx = xs[3 * i] - cx, dy = xs[3 * i + 1] - cy, dz = xs[3 * i + 2] - cz;
  return dx * dx + dy * dy + dz * dz;
}

// one wave per node of at most kSmallRadius particles: max squared distance
__global__ void __launch_bounds__(256) radius_small_kernel(const double* __restrict__ xs, int32_t nodes,
                                                           const int32_t* __restrict__ range,
                                                           const double* __restrict__ center,
                                                           double* __restrict__ radius) {
  const int32_t node = blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x % 64;
  if (node >= nodes) return;  // whole waves leave together
  const int64_t a = range[2 * node], b = range[2 * node + 1];
  if (b - a + 1 > kSmallRadius) return;
  const double cx = center[3 * node], cy = center[3 * node + 1], cz = center[3 * node + 2];
  double best = 0.0;
  for (int64_t i = a + lane; i <= b; i += 64) best = fmax(best, dist2(xs, i, cx, cy, cz));
  for (int off = 32; off >= 1; off >>= 1) best = fmax(best, __shfl_xor(best, off));
  if (lane == 0) radius[node] = sqrt(best);
}

// one 1024-thread block per larger node, four independent loads in flight per thread
__global__ void __launch_bounds__(1024) radius_big_kernel(const double* __restrict__ xs, int32_t nodes,
                                                          const int32_t* __restrict__ range,
                                                          const double* __restrict__ center,
                                                          double* __restrict__ radius) {
  __shared__ double red[16];
  const int32_t node = blockIdx.x;
  const int64_t a = range[2 * node], b = range[2 * node + 1];
  if (b - a + 1 <= kSmallRadius) return;  // uniform per block
  const double cx = center[3 * node], cy = center[3 * node + 1], cz = center[3 * node + 2];
  double best = 0.0;
  int64_t i = a + threadIdx.x;
  for (; i + 3 * 1024 <= b; i += 4 * 1024) {
    const double d0 = dist2(xs, i, cx, cy, cz), d1 = dist2(xs, i + 1024, cx, cy, cz);
    const double d2 = dist2(xs, i + 2048, cx, cy, cz), d3 = dist2(xs, i + 3072, cx, cy, cz);
    best = fmax(best, fmax(fmax(d0, d1), fmax(d2, d3)));
  }
  for (; i <= b; i += 1024) best = fmax(best, dist2(xs, i, cx, cy, cz));
  for (int off = 32; off >= 1; off >>= 1) best = fmax(best, __shfl_xor(best, off));
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) best = fmax(best, red[w]);
    radius[node] = sqrt(best);
  }
}

// ---------------------------------------------------------------------------
// dual-tree walk, one breadth-first step
// ---------------------------------------------------------------------------

// Every frontier item is classified first (what it emits), the per-item
// counts are exclusive-scanned, and a second pass writes each item's outputs
// at its scanned offsets: no atomics, and the output order is a function of
// the input order alone.
enum WalkCode : uint8_t { kNone = 0, kSelf = 1, kFar = 2, kNear = 3, kSplitA = 4, kSplitB = 5 };

__global__ void __launch_bounds__(256) walk_classify_kernel(const int2* __restrict__ in, int64_t n_in, int32_t I,
                                                            const double* __restrict__ C,
                                                            const double* __restrict__ R, double th2,
                                                            uint8_t* __restrict__ code,
                                                            int32_t* __restrict__ counts3) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= n_in) return;
  const int32_t a = in[t].x, b = in[t].y;
  uint8_t c;
  if (a == b) {
    c = a < I ? kSelf : kNone;
  } else {
    const double dx = C[3 * a] - C[3 * b], dy = C[3 * a + 1] - C[3 * b + 1], dz = C[3 * a + 2] - C[3 * b + 2];
    const double d2 = dx * dx + dy * dy + dz * dz;
    const double s = R[a] + R[b];
    const bool al = a >= I, bl = b >= I;
    if (d2 > 0.0 && s * s <= th2 * d2) c = kFar;
    else if (al && bl) c = kNear;
    else if (bl || (!al && R[a] >= R[b])) c = kSplitA;
    else c = kSplitB;
  }
  code[t] = c;
  counts3[t] = c == kSelf ? 3 : (c == kSplitA || c == kSplitB) ? 2 : 0;
  counts3[n_in + t] = c == kFar ? 2 : 0;
  counts3[2 * n_in + t] = c == kNear ? 2 : 0;
}

__global__ void __launch_bounds__(256) walk_emit_kernel(const int2* __restrict__ in, int64_t n_in, int32_t I,
                                                        const int32_t* __restrict__ left,
                                                        const int32_t* __restrict__ right,
                                                        const uint8_t* __restrict__ code,
                                                        const int32_t* __restrict__ offsets3, int64_t far_base,
                                                        int64_t near_base, int2* __restrict__ items_out,
                                                        uint64_t* __restrict__ far, uint64_t* __restrict__ near) {
  const int64_t t = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (t >= n_in) return;
  const int32_t a = in[t].x, b = in[t].y;
  const uint8_t c = code[t];
  const int64_t oi = offsets3[t];
  switch (c) {
    case kSelf: {
      const int32_t la = left[a], ra = right[a];
      items_out[oi] = make_int2(la, ra);
      items_out[oi + 1] = make_int2(ra, ra);
      items_out[oi + 2] = make_int2(la, la);
      break;
    }
    case kFar: {
      const int64_t o = far_base + offsets3[n_in + t];
      far[o] = (uint64_t(uint32_t(a)) << 32) | uint32_t(b);
      far[o + 1] = (uint64_t(uint32_t(b)) << 32) | uint32_t(a);
      break;
    }
    case kNear: {
      const int64_t o = near_base + offsets3[2 * n_in + t];
      const uint32_t ta = uint32_t(a - I), tb = uint32_t(b - I);
      near[o] = (uint64_t(ta) << 32) | tb;
      near[o + 1] = (uint64_t(tb) << 32) | ta;
      break;
    }
    case kSplitA:
      items_out[oi] = make_int2(right[a], b);
      items_out[oi + 1] = make_int2(left[a], b);
      break;
    case kSplitB:
      items_out[oi] = make_int2(a, right[b]);
      items_out[oi + 1] = make_int2(a, left[b]);
      break;
    default:
      break;
  }
}

// ---------------------------------------------------------------------------
// sorted (row << 32 | col) keys -> CSR
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(256) csr_offsets_kernel(const uint64_t* __restrict__ keys, int64_t K, int32_t rows,
                                                          int32_t* __restrict__ off) {
  const int32_t r = blockIdx.x * 256 + threadIdx.x;
  if (r > rows) return;
  const uint64_t want = uint64_t(r) << 32;
  int64_t lo = 0, hi = K;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (keys[mid] < want) lo = mid + 1;
    else hi = mid;
  }
  off[r] = static_cast<int32_t>(lo);
}

__global__ void __launch_bounds__(256) low_words_kernel(const uint64_t* __restrict__ keys, int64_t K,
                                                        int32_t* __restrict__ col) {
  const int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (k < K) col[k] = static_cast<int32_t>(keys[k] & 0xFFFFFFFFull);
}

inline unsigned grid_for(int64_t n, int per_block = 256) {
  return static_cast<unsigned>((n + per_block - 1) / per_block);
}

inline int key_bits(uint64_t max_key) {
  int bits = 0;
  while (bits < 64 && (max_key >> bits) != 0) ++bits;
  return bits;
}

}  // namespace
}  // namespace jfm

using jfm::fail;

extern "C" {

int32_t jfm_tree_bounds_scratch_doubles(void) { return 6 * jfm::kBoundBlocks; }

int32_t jfm_tree_bounds(void* stream, const double* pos, int64_t n, int32_t given, double* scratch, double* bounds6,
                        int32_t* flags) {
  if (!pos || !scratch || !bounds6 || !flags || n < 1) return fail(JFM_ERR_ARG, "jfm_tree_bounds: bad argument");
  const int blocks = static_cast<int>(std::min<int64_t>(jfm::kBoundBlocks, (n + 255) / 256));
  hipLaunchKernelGGL(jfm::bounds_partial_kernel, dim3(blocks), dim3(256), 0, jfm::as_stream(stream), pos, n, scratch,
                     flags);
  if (!given)
    hipLaunchKernelGGL(jfm::bounds_final_kernel, dim3(1), dim3(256), 0, jfm::as_stream(stream), scratch, blocks,
                       bounds6);
  return jfm::check_launch("jfm_tree_bounds");
}

int32_t jfm_tree_morton(void* stream, const double* pos, int64_t n, const double* bounds6, uint64_t* codes,
                        int32_t* index) {
  if (!pos || !bounds6 || !codes || !index || n < 1 || n > (int64_t(1) << 30))
    return fail(JFM_ERR_ARG, "jfm_tree_morton: bad argument");
  hipLaunchKernelGGL(jfm::morton_kernel, dim3(jfm::grid_for(n)), dim3(256), 0, jfm::as_stream(stream), pos, n,
                     bounds6, codes, index);
  return jfm::check_launch("jfm_tree_morton");
}

int32_t jfm_sort_scratch_bytes(int64_t n, int32_t with_values, int64_t* bytes) {
  if (!bytes || n < 0 || n > INT32_MAX) return fail(JFM_ERR_ARG, "jfm_sort_scratch_bytes: bad argument");
  size_t b = 0;
  hipError_t e;
  if (with_values)
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, b, static_cast<const uint64_t*>(nullptr),
                                           static_cast<uint64_t*>(nullptr), static_cast<const int32_t*>(nullptr),
                                           static_cast<int32_t*>(nullptr), static_cast<int>(n), 0, 64);
  else
    e = hipcub::DeviceRadixSort::SortKeys(nullptr, b, static_cast<const uint64_t*>(nullptr),
                                          static_cast<uint64_t*>(nullptr), static_cast<int>(n), 0, 64);
  if (e != hipSuccess) return fail(JFM_ERR_HIP, std::string("jfm_sort_scratch_bytes: ") + hipGetErrorString(e));
  *bytes = static_cast<int64_t>(b);
  return JFM_OK;
}

int32_t jfm_sort_u64(void* stream, int64_t n, int32_t end_bit, const uint64_t* keys_in, uint64_t* keys_out,
                     const int32_t* vals_in, int32_t* vals_out, void* scratch, int64_t scratch_bytes) {
  if (n < 0 || n > INT32_MAX || end_bit < 1 || end_bit > 64 || (n > 0 && (!keys_in || !keys_out)) ||
      ((vals_in == nullptr) != (vals_out == nullptr)))
    return fail(JFM_ERR_ARG, "jfm_sort_u64: bad argument");
  if (n == 0) return JFM_OK;
  int64_t need = 0;
  if (int32_t rc = jfm_sort_scratch_bytes(n, vals_in != nullptr, &need)) return rc;
  if (!scratch || scratch_bytes < need) return fail(JFM_ERR_ARG, "jfm_sort_u64: scratch too small");
  size_t b = static_cast<size_t>(scratch_bytes);
  hipError_t e;
  if (vals_in)
    e = hipcub::DeviceRadixSort::SortPairs(scratch, b, keys_in, keys_out, vals_in, vals_out, static_cast<int>(n), 0,
                                           end_bit, jfm::as_stream(stream));
  else
    e = hipcub::DeviceRadixSort::SortKeys(scratch, b, keys_in, keys_out, static_cast<int>(n), 0, end_bit,
                                          jfm::as_stream(stream));
  if (e != hipSuccess) return fail(JFM_ERR_HIP, std::string("jfm_sort_u64: ") + hipGetErrorString(e));
  return jfm::check_launch("jfm_sort_u64");
}

int32_t jfm_tree_radix(void* stream, int64_t n, int32_t leaf_size, const uint64_t* sorted_codes, int32_t* parent,
                       int32_t* left, int32_t* right, int32_t* node_range) {
  if (n < 1 || n > (int64_t(1) << 30) || leaf_size < 1 || !sorted_codes || !parent || !node_range)
    return fail(JFM_ERR_ARG, "jfm_tree_radix: bad argument");
  const int32_t L = static_cast<int32_t>((n + leaf_size - 1) / leaf_size);
  if (L > 1 && (!left || !right)) return fail(JFM_ERR_ARG, "jfm_tree_radix: left/right NULL");
  jfm::LeafKeys keys{sorted_codes, leaf_size, L};
  hipLaunchKernelGGL(jfm::radix_tree_kernel, dim3(jfm::grid_for(L)), dim3(256), 0, jfm::as_stream(stream), keys, n,
                     parent, left, right, node_range);
  return jfm::check_launch("jfm_tree_radix");
}

int32_t jfm_tree_levels(void* stream, int32_t nodes, int32_t I, const int32_t* parent, int32_t* depth,
                        int32_t* max_depth, int32_t* flags, uint64_t* m2m_keys, uint64_t* l2l_keys) {
  if (nodes < 1 || I < 0 || !parent || !depth || !max_depth || !flags || (I > 0 && (!m2m_keys || !l2l_keys)))
    return fail(JFM_ERR_ARG, "jfm_tree_levels: bad argument");
  hipStream_t s = jfm::as_stream(stream);
  hipLaunchKernelGGL(jfm::depth_kernel, dim3(jfm::grid_for(nodes)), dim3(256), 0, s, nodes, parent, depth, max_depth,
                     flags);
  if (I > 0)
    hipLaunchKernelGGL(jfm::level_keys_kernel, dim3(jfm::grid_for(nodes)), dim3(256), 0, s, nodes, I, depth,
                       max_depth, m2m_keys, l2l_keys);
  return jfm::check_launch("jfm_tree_levels");
}

int32_t jfm_tree_leaf_geometry(void* stream, int32_t L, int32_t I, const double* pos, const double* mass,
                               const int32_t* perm, const int32_t* node_range, double* node_mass, double* center) {
  if (L < 1 || !pos || !mass || !perm || !node_range || !node_mass || !center)
    return fail(JFM_ERR_ARG, "jfm_tree_leaf_geometry: bad argument");
  jfm::Particles P{pos, mass, perm};
  hipLaunchKernelGGL(jfm::leaf_geometry_kernel, dim3(jfm::grid_for(L)), dim3(256), 0, jfm::as_stream(stream), P, L, I,
                     node_range, node_mass, center);
  return jfm::check_launch("jfm_tree_leaf_geometry");
}

int32_t jfm_tree_level_geometry(void* stream, int32_t count, const int32_t* level_nodes, const double* pos,
                                const double* mass, const int32_t* perm, const int32_t* left, const int32_t* right,
                                const int32_t* node_range, double* node_mass, double* center) {
  if (count < 0 || (count > 0 && (!level_nodes || !pos || !mass || !perm || !left || !right || !node_range ||
                                  !node_mass || !center)))
    return fail(JFM_ERR_ARG, "jfm_tree_level_geometry: bad argument");
  if (count == 0) return JFM_OK;
  jfm::Particles P{pos, mass, perm};
  hipLaunchKernelGGL(jfm::internal_geometry_kernel, dim3(jfm::grid_for(count)), dim3(256), 0, jfm::as_stream(stream),
                     P, count, level_nodes, left, right, node_range, node_mass, center);
  return jfm::check_launch("jfm_tree_level_geometry");
}

int32_t jfm_tree_radius(void* stream, int32_t nodes, int64_t n, const double* pos, const int32_t* perm,
                        const int32_t* node_range, const double* center, double* scratch, double* radius) {
  if (nodes < 1 || n < 1 || !pos || !perm || !node_range || !center || !scratch || !radius)
    return fail(JFM_ERR_ARG, "jfm_tree_radius: bad argument");
  hipStream_t s = jfm::as_stream(stream);
  jfm::Particles P{pos, nullptr, perm};
  hipLaunchKernelGGL(jfm::sorted_positions_kernel, dim3(jfm::grid_for(n)), dim3(256), 0, s, P, n, scratch);
  hipLaunchKernelGGL(jfm::radius_small_kernel, dim3(jfm::grid_for(nodes, 4)), dim3(256), 0, s, scratch, nodes,
                     node_range, center, radius);
  hipLaunchKernelGGL(jfm::radius_big_kernel, dim3(nodes), dim3(1024), 0, s, scratch, nodes, node_range, center,
                     radius);
  return jfm::check_launch("jfm_tree_radius");
}

int32_t jfm_tree_walk_classify(void* stream, int32_t I, const int32_t* items_in, int64_t n_in, const double* center,
                               const double* radius, double theta, uint8_t* codes, int32_t* counts3) {
  if (n_in < 0 || n_in > INT32_MAX || I < 0 || !center || !radius || !(theta > 0.0) ||
      (n_in > 0 && (!items_in || !codes || !counts3)))
    return fail(JFM_ERR_ARG, "jfm_tree_walk_classify: bad argument");
  if (n_in == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::walk_classify_kernel, dim3(jfm::grid_for(n_in)), dim3(256), 0, jfm::as_stream(stream),
                     reinterpret_cast<const int2*>(items_in), n_in, I, center, radius, theta * theta, codes, counts3);
  return jfm::check_launch("jfm_tree_walk_classify");
}

int32_t jfm_tree_walk_emit(void* stream, int32_t I, const int32_t* items_in, int64_t n_in, const int32_t* left,
                           const int32_t* right, const uint8_t* codes, const int32_t* offsets3, int64_t far_base,
                           int64_t near_base, int32_t* items_out, uint64_t* far_keys, uint64_t* near_keys) {
  if (n_in < 0 || n_in > INT32_MAX || I < 0 || far_base < 0 || near_base < 0 ||
      (n_in > 0 && (!items_in || !codes || !offsets3 || !items_out || !far_keys || !near_keys)) ||
      (I > 0 && (!left || !right)))
    return fail(JFM_ERR_ARG, "jfm_tree_walk_emit: bad argument");
  if (n_in == 0) return JFM_OK;
  hipLaunchKernelGGL(jfm::walk_emit_kernel, dim3(jfm::grid_for(n_in)), dim3(256), 0, jfm::as_stream(stream),
                     reinterpret_cast<const int2*>(items_in), n_in, I, left, right, codes, offsets3, far_base,
                     near_base, reinterpret_cast<int2*>(items_out), far_keys, near_keys);
  return jfm::check_launch("jfm_tree_walk_emit");
}

int32_t jfm_scan_scratch_bytes(int64_t n, int64_t* bytes) {
  if (!bytes || n < 0 || n > INT32_MAX) return fail(JFM_ERR_ARG, "jfm_scan_scratch_bytes: bad argument");
  size_t b = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, b, static_cast<const int32_t*>(nullptr),
                                                  static_cast<int32_t*>(nullptr), static_cast<int>(n));
  if (e != hipSuccess) return fail(JFM_ERR_HIP, std::string("jfm_scan_scratch_bytes: ") + hipGetErrorString(e));
  *bytes = static_cast<int64_t>(b);
  return JFM_OK;
}

int32_t jfm_scan_i32(void* stream, int64_t n, const int32_t* in, int32_t* out, void* scratch, int64_t scratch_bytes) {
  if (n < 0 || n > INT32_MAX || (n > 0 && (!in || !out))) return fail(JFM_ERR_ARG, "jfm_scan_i32: bad argument");
  if (n == 0) return JFM_OK;
  int64_t need = 0;
  if (int32_t rc = jfm_scan_scratch_bytes(n, &need)) return rc;
  if (!scratch || scratch_bytes < need) return fail(JFM_ERR_ARG, "jfm_scan_i32: scratch too small");
  size_t b = static_cast<size_t>(scratch_bytes);
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(scratch, b, in, out, static_cast<int>(n), jfm::as_stream(stream));
  if (e != hipSuccess) return fail(JFM_ERR_HIP, std::string("jfm_scan_i32: ") + hipGetErrorString(e));
  return jfm::check_launch("jfm_scan_i32");
}

int32_t jfm_keys_to_csr(void* stream, const uint64_t* sorted_keys, int64_t K, int32_t rows, int32_t* off,
                        int32_t* col) {
  if (K < 0 || K > INT32_MAX || rows < 0 || !off || (K > 0 && (!sorted_keys || !col)))
    return fail(JFM_ERR_ARG, "jfm_keys_to_csr: bad argument");
  hipStream_t s = jfm::as_stream(stream);
  hipLaunchKernelGGL(jfm::csr_offsets_kernel, dim3(jfm::grid_for(int64_t(rows) + 1)), dim3(256), 0, s, sorted_keys, K,
                     rows, off);
  if (K > 0)
    hipLaunchKernelGGL(jfm::low_words_kernel, dim3(jfm::grid_for(K)), dim3(256), 0, s, sorted_keys, K, col);
  return jfm::check_launch("jfm_keys_to_csr");
}

int32_t jfm_key_bits(uint64_t max_key) { return jfm::key_bits(max_key); }

}  // extern "C"
