// Upward and downward tree sweeps: gather into tree order, leaf P2M, per-level
// M2M and L2L, and the leaf L2P far-field evaluation.
//
// Each sweep is HBM-bound (tens of bytes per coefficient row moved, a few
// hundred flops per row), so the kernels read whole 16-byte-padded rows with
// vector loads, keep the row in registers for the full rotate / shift /
// rotate-back cascade, and write it once.  Per-level launches give the
// level-synchronous order the recurrences need without atomics, so every sum
// is deterministic.

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "launch.hpp"
#include "rows.hpp"

namespace jfm {
namespace {

constexpr int kWave = 64;

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

template <typename TS, typename T>
__global__ __launch_bounds__(256) void gather_kernel(const TS* __restrict__ pos, const TS* __restrict__ mass,
                                                     const int32_t* __restrict__ perm, int32_t n,
                                                     T* __restrict__ pos4) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = perm[i];
  V4<T> q;
  q.x = static_cast<T>(pos[3 * static_cast<int64_t>(j) + 0]);
  q.y = static_cast<T>(pos[3 * static_cast<int64_t>(j) + 1]);
  q.z = static_cast<T>(pos[3 * static_cast<int64_t>(j) + 2]);
  q.w = static_cast<T>(mass[j]);
  store4(pos4, i, q);
}

// One wave per leaf: centre of mass, then M = sum_i m_i U(x_i - c).
template <typename T, int P>
__global__ __launch_bounds__(kWave) void p2m_kernel(const T* __restrict__ pos4, int32_t n, int32_t leaf_size,
                                                    int32_t num_internal, T* __restrict__ mult, int32_t cs,
                                                    T* __restrict__ centers4) {
  constexpr int C = ncoef(P);
  __shared__ T red[kWave][C + 1];
  const int leaf = blockIdx.x;
  const int lane = threadIdx.x;
  const int first = leaf * leaf_size;
  const int cnt = min(leaf_size, n - first);
  T ms = 0, mx = 0, my = 0, mz = 0, px = 0, py = 0, pz = 0;
  for (int i = lane; i < cnt; i += kWave) {
    const V4<T> q = load4(pos4, first + i);
    ms += q.w;
    mx += q.w * q.x;
    my += q.w * q.y;
    mz += q.w * q.z;
    px += q.x;
    py += q.y;
    pz += q.z;
  }
  ms = wave_sum(ms);
  mx = wave_sum(mx);
  my = wave_sum(my);
  mz = wave_sum(mz);
  px = wave_sum(px);
  py = wave_sum(py);
  pz = wave_sum(pz);
  T cx, cy, cz;
  if (ms > T(0)) {
    cx = mx / ms;
    cy = my / ms;
    cz = mz / ms;
  } else {
    const T inv = T(1) / static_cast<T>(cnt);
    cx = px * inv;
    cy = py * inv;
    cz = pz * inv;
  }
  T acc[C];
#pragma unroll
  for (int j = 0; j < C; ++j) acc[j] = T(0);
  for (int i = lane; i < cnt; i += kWave) {
    const V4<T> q = load4(pos4, first + i);
    T re[ntri(P)], im[ntri(P)];
    solid_harmonics<T, P>(q.x - cx, q.y - cy, q.z - cz, re, im);
#pragma unroll
    for (int nn = 0; nn <= P; ++nn) {
      acc[sh_idx(nn, 0)] = fma(q.w, re[tri(nn, 0)], acc[sh_idx(nn, 0)]);
#pragma unroll
      for (int m = 1; m <= nn; ++m) {
        acc[sh_idx(nn, m)] = fma(q.w, re[tri(nn, m)], acc[sh_idx(nn, m)]);
        acc[sh_idx(nn, -m)] = fma(q.w, im[tri(nn, m)], acc[sh_idx(nn, -m)]);
      }
    }
  }
#pragma unroll
  for (int j = 0; j < C; ++j) red[lane][j] = acc[j];
  __syncthreads();
  const int node = num_internal + leaf;
  for (int j = lane; j < cs; j += kWave) {
    T s = T(0);
    if (j < C)
      for (int l = 0; l < kWave; ++l) s += red[l][j];
    mult[static_cast<int64_t>(node) * cs + j] = s;
  }
  if (lane == 0) store4(centers4, node, V4<T>{cx, cy, cz, ms});
}

// The value of another lane of the same row of 16, picked by a DPP control
// (every lane of the row active): a VALU move instead of an LDS permute.
// 0xB1 = quad_perm [1,0,3,2] (lane ^ 1), 0x4E = quad_perm [2,3,0,1]
// (lane ^ 2), 0x141 = row_half_mirror (lane 7 - i of its 8), 0x140 =
// row_mirror (lane 15 - i of its 16).
template <int CTRL>
__device__ __forceinline__ float dpp_get(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_get(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}

// Narrow leaves (leaf_size <= G <= 32): G lanes per leaf, 64 / G leaves per
// wave.  The same per-leaf quantities as p2m_kernel, summed over the leaf's
// lanes by an xor tree inside each aligned group of G lanes (fixed order).
// After the step over 2^k lanes every lane of an aligned 2^k group holds the
// same sum (a + b and b + a round alike), so the partner of the next step may
// be any lane of the other half: the mirrors inside a row of 16 give the
// xor 4 and xor 8 sums bit for bit.
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
  static_assert(G == 8 || G == 16 || G == 32, "group of 8, 16 or 32 lanes");
  v += dpp_get<0xB1>(v);
  v += dpp_get<0x4E>(v);
  v += dpp_get<0x141>(v);
  if constexpr (G >= 16) v += dpp_get<0x140>(v);
  if constexpr (G >= 32) v += __shfl_xor(v, 16, kWave);
  return v;
}

template <typename T, int P, int G>
__global__ __launch_bounds__(kWave) void p2m_group_kernel(const T* __restrict__ pos4, int32_t n, int32_t leaf_size,
                                                          int32_t num_leaves, int32_t num_internal,
                                                          T* __restrict__ mult, int32_t cs, T* __restrict__ centers4) {
  constexpr int C = ncoef(P);
  const int lane = threadIdx.x;
  const int leaf = blockIdx.x * (kWave / G) + lane / G;
  const int sub = lane % G;
  const bool live = leaf < num_leaves;  // dead groups still join the shuffles
  const int first = leaf * leaf_size;
  const int cnt = live ? min(leaf_size, n - first) : 0;
  V4<T> q{T(0), T(0), T(0), T(0)};
  if (sub < cnt) q = load4(pos4, first + sub);
  const T ms = group_sum<G>(q.w);
  const T mx = group_sum<G>(q.w * q.x);
  const T my = group_sum<G>(q.w * q.y);
  const T mz = group_sum<G>(q.w * q.z);
  const T px = group_sum<G>(q.x);
  const T py = group_sum<G>(q.y);
  const T pz = group_sum<G>(q.z);
  T cx, cy, cz;
  if (ms > T(0)) {
    cx = mx / ms;
    cy = my / ms;
    cz = mz / ms;
  } else {
    const T inv = T(1) / static_cast<T>(cnt > 0 ? cnt : 1);
    cx = px * inv;
    cy = py * inv;
    cz = pz * inv;
  }
  T acc[C];
  {
    T re[ntri(P)], im[ntri(P)];
    solid_harmonics<T, P>(q.x - cx, q.y - cy, q.z - cz, re, im);
#pragma unroll
    for (int nn = 0; nn <= P; ++nn) {
      acc[sh_idx(nn, 0)] = q.w * re[tri(nn, 0)];
#pragma unroll
      for (int m = 1; m <= nn; ++m) {
        acc[sh_idx(nn, m)] = q.w * re[tri(nn, m)];
        acc[sh_idx(nn, -m)] = q.w * im[tri(nn, m)];
      }
    }
  }
#pragma unroll
  for (int j = 0; j < C; ++j) acc[j] = group_sum<G>(acc[j]);
  if (!live) return;
  const int node = num_internal + leaf;
  T* row = mult + static_cast<int64_t>(node) * cs;
#pragma unroll
  for (int j = 0; j < C; ++j)
    if (j % G == sub) row[j] = acc[j];
  for (int j = C + sub; j < cs; j += G) row[j] = T(0);
  if (sub == 0) store4(centers4, node, V4<T>{cx, cy, cz, ms});
}

// One thread per internal node of a level: centre of mass of the two children,
// M = m2m(M_left, c_left - c) + m2m(M_right, c_right - c).
// L2L into one node: L += l2l(L_parent, c_parent - c).
template <typename T, int P>
__device__ __forceinline__ void l2l_node(int node, const int32_t* __restrict__ parent,
                                         const T* __restrict__ centers4, T* __restrict__ locals, int32_t cs) {
  constexpr int C = ncoef(P);
  const int par = parent[node];
  const V4<T> cp = load4(centers4, par);
  const V4<T> cc = load4(centers4, node);
  T v[C];
  load_row<T, C>(locals + static_cast<int64_t>(par) * cs, v);
  Align<T, P> a;
  a.init(cp.x - cc.x, cp.y - cc.y, cp.z - cc.z);  // source(parent) - destination(child)
  rotate_all<T, P, Rot::ToZLocal>(v, a);
  zshift_l2l<T, P>(v, a.r);
  rotate_all<T, P, Rot::FromZLocal>(v, a);
  T w[C];
  load_row<T, C>(locals + static_cast<int64_t>(node) * cs, w);
#pragma unroll
  for (int j = 0; j < C; ++j) w[j] += v[j];
  store_row<T, C>(locals + static_cast<int64_t>(node) * cs, w);
}

// Two lanes per internal node of a level (children already done): lane 2i
// translates the left child's multipole to the node's centre of mass, lane
// 2i+1 the right one's; the sum is formed as left + right and the even lane
// writes the node.  Halves the dependent instruction chain of a level, which
// is what a level of the upward pass waits for.
template <typename T, int P>
__global__ __launch_bounds__(kWave) void m2m_kernel(const int32_t* __restrict__ nodes, int32_t count,
                                                    const int32_t* __restrict__ left,
                                                    const int32_t* __restrict__ right,
                                                    const int32_t* __restrict__ node_range, T* __restrict__ mult,
                                                    int32_t cs, T* __restrict__ centers4) {
  constexpr int C = ncoef(P);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int side = t & 1;
  // lanes past the level repeat its last node (results dropped) so that both
  // lanes of every pair reach the exchange below
  const bool live = (t >> 1) < count;
  const int node = nodes[live ? (t >> 1) : count - 1];
  const int lc = left[node], rc = right[node];
  const V4<T> cl = load4(centers4, lc);
  const V4<T> cr = load4(centers4, rc);
  const T tot = cl.w + cr.w;
  T cx, cy, cz;
  if (tot > T(0)) {
    cx = (cl.w * cl.x + cr.w * cr.x) / tot;
    cy = (cl.w * cl.y + cr.w * cr.y) / tot;
    cz = (cl.w * cl.z + cr.w * cr.z) / tot;
  } else {  // massless: particle-count-weighted mean of the children's means
    const T nl = static_cast<T>(node_range[2 * lc + 1] - node_range[2 * lc] + 1);
    const T nr = static_cast<T>(node_range[2 * rc + 1] - node_range[2 * rc] + 1);
    cx = (nl * cl.x + nr * cr.x) / (nl + nr);
    cy = (nl * cl.y + nr * cr.y) / (nl + nr);
    cz = (nl * cl.z + nr * cr.z) / (nl + nr);
  }
  const int ch = side == 0 ? lc : rc;
  const V4<T> cc = side == 0 ? cl : cr;
  T v[C];
  load_row<T, C>(mult + static_cast<int64_t>(ch) * cs, v);
  Align<T, P> a;
  a.init(cc.x - cx, cc.y - cy, cc.z - cz);  // source - destination
  rotate_all<T, P, Rot::ToZMultipole>(v, a);
  zshift_m2m<T, P>(v, a.r);
  rotate_all<T, P, Rot::FromZMultipole>(v, a);
#pragma unroll
  for (int j = 0; j < C; ++j) v[j] = v[j] + dpp_get<0xB1>(v[j]);  // left + right on both lanes
  if (side == 0 && live) {
    store_row<T, C>(mult + static_cast<int64_t>(node) * cs, v);
    store4(centers4, node, V4<T>{cx, cy, cz, tot});
  }
}

// One thread per node of a level: L += l2l(L_parent, c_parent - c).
template <typename T, int P>
__global__ __launch_bounds__(kWave) void l2l_kernel(const int32_t* __restrict__ nodes, int32_t count,
                                                    const int32_t* __restrict__ parent,
                                                    const T* __restrict__ centers4, T* __restrict__ locals,
                                                    int32_t cs) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  l2l_node<T, P>(nodes[t], parent, centers4, locals, cs);
}

// One thread per particle: far field from its leaf's local expansion.
// a = -G grad_delta phi, pot = -G phi, delta = centre - x.  Writes the output
// row, or adds to it (accumulate != 0, after the near field).
template <typename T, int P>
__global__ __launch_bounds__(256) void l2p_kernel(const T* __restrict__ pos4, int32_t n, int32_t leaf_size,
                                                  int32_t num_internal, const T* __restrict__ locals, int32_t cs,
                                                  const T* __restrict__ centers4, T G,
                                                  const int32_t* __restrict__ out_index, T* __restrict__ acc,
                                                  T* __restrict__ pot, int32_t accumulate) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int node = num_internal + i / leaf_size;
  const V4<T> q = load4(pos4, i);
  const V4<T> c = load4(centers4, node);
  const T* F = locals + static_cast<int64_t>(node) * cs;
  T phi, gx, gy, gz;
  l2p_eval<T, P>([F](int j) { return F[j]; }, c.x - q.x, c.y - q.y, c.z - q.z, phi, gx, gy, gz);
  const int64_t row = out_index ? out_index[i] : i;
  if (accumulate) {
    acc[3 * row + 0] += -G * gx;
    acc[3 * row + 1] += -G * gy;
    acc[3 * row + 2] += -G * gz;
    if (pot) pot[row] += -G * phi;
  } else {
    acc[3 * row + 0] = -G * gx;
    acc[3 * row + 1] = -G * gy;
    acc[3 * row + 2] = -G * gz;
    if (pot) pot[row] = -G * phi;
  }
}

// acc[k] += -G far[k] over 3 n entries (and pot[r] += -G far_pot[r]): the
// far field of an L2P run with G = -1 (unscaled gradient and potential rows),
// added to near-field rows with the same expression as the L2P accumulate,
// so the sum equals an L2P that accumulates onto those rows bit for bit.
template <typename T>
__global__ __launch_bounds__(256) void l2p_combine_kernel(int64_t n3, int32_t n, const T* __restrict__ far_acc,
                                                          const T* __restrict__ far_pot, T G, T* __restrict__ acc,
                                                          T* __restrict__ pot) {
  const int64_t k = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (k < n3) acc[k] += -G * far_acc[k];
  if (pot && k < n) pot[k] += -G * far_pot[k];
}

// Mixed precision: the internal nodes of the top three levels (depth <= 2,
// at most 7 nodes) keep float64 shadow rows.  One workgroup walks their M2M
// levels deepest first in float64, reading float32 rows for the children
// below, and writes both the shadow and the rounded float32 row.  Two lanes
// per node as in m2m_kernel: lane 2i translates the left child, lane 2i+1 the
// right one, and the even lane writes left + right, so a level waits for one
// child's translation instead of two.
template <int P>
__global__ __launch_bounds__(kWave) void m2m_top_kernel(const int32_t* __restrict__ nodes,
                                                        const int32_t* __restrict__ lev_off, int32_t nlev,
                                                        const int32_t* __restrict__ left,
                                                        const int32_t* __restrict__ right,
                                                        const int32_t* __restrict__ node_range,
                                                        const int32_t* __restrict__ shadow, float* __restrict__ mult,
                                                        int32_t cs, float* __restrict__ centers4,
                                                        double* __restrict__ mult64, int32_t cs64,
                                                        double* __restrict__ centers64) {
  constexpr int C = ncoef(P);
  const int t = threadIdx.x;
  const int side = t & 1;
  for (int lev = 0; lev < nlev; ++lev) {
    const int count = lev_off[lev + 1] - lev_off[lev];
    if ((t >> 1) < count) {  // both lanes of a pair take this branch together
      const int node = nodes[lev_off[lev] + (t >> 1)];
      const int lc = left[node], rc = right[node];
      V4<double> cl, cr;
      {
        const int sl = shadow[lc], sr = shadow[rc];
        if (sl >= 0) {
          cl = load4(centers64, sl);
        } else {
          const V4<float> f = load4(centers4, lc);
          cl = V4<double>{f.x, f.y, f.z, f.w};
        }
        if (sr >= 0) {
          cr = load4(centers64, sr);
        } else {
          const V4<float> f = load4(centers4, rc);
          cr = V4<double>{f.x, f.y, f.z, f.w};
        }
      }
      const double tot = cl.w + cr.w;
      double cx, cy, cz;
      if (tot > 0.0) {
        cx = (cl.w * cl.x + cr.w * cr.x) / tot;
        cy = (cl.w * cl.y + cr.w * cr.y) / tot;
        cz = (cl.w * cl.z + cr.w * cr.z) / tot;
      } else {
        const double nl = node_range[2 * lc + 1] - node_range[2 * lc] + 1;
        const double nr = node_range[2 * rc + 1] - node_range[2 * rc] + 1;
        cx = (nl * cl.x + nr * cr.x) / (nl + nr);
        cy = (nl * cl.y + nr * cr.y) / (nl + nr);
        cz = (nl * cl.z + nr * cr.z) / (nl + nr);
      }
      const int ch = side == 0 ? lc : rc;
      const V4<double> cc = side == 0 ? cl : cr;
      double v[C];
      const int sh = shadow[ch];
      if (sh >= 0) {
        load_row<double, C>(mult64 + static_cast<int64_t>(sh) * cs64, v);
      } else {
        float tmp[C];
        load_row<float, C>(mult + static_cast<int64_t>(ch) * cs, tmp);
#pragma unroll
        for (int j = 0; j < C; ++j) v[j] = tmp[j];
      }
      Align<double, P> a;
      a.init(cc.x - cx, cc.y - cy, cc.z - cz);  // source - destination
      rotate_all<double, P, Rot::ToZMultipole>(v, a);
      zshift_m2m<double, P>(v, a.r);
      rotate_all<double, P, Rot::FromZMultipole>(v, a);
#pragma unroll
      for (int j = 0; j < C; ++j) v[j] = v[j] + dpp_get<0xB1>(v[j]);  // left + right on both lanes
      if (side == 0) {
        const int sn = shadow[node];
        store_row<double, C>(mult64 + static_cast<int64_t>(sn) * cs64, v);
        store4(centers64, sn, V4<double>{cx, cy, cz, tot});
        float vf[C];
#pragma unroll
        for (int j = 0; j < C; ++j) vf[j] = static_cast<float>(v[j]);
        store_row<float, C>(mult + static_cast<int64_t>(node) * cs, vf);
        store4(centers4, node, V4<float>{static_cast<float>(cx), static_cast<float>(cy), static_cast<float>(cz),
                                         static_cast<float>(tot)});
      }
    }
    __syncthreads();
  }
}

// Mixed precision L2L of the top levels: loc64 starts from the float32 M2L
// rows of the shadowed nodes; then for the nodes at depth 1..3 (shallowest
// first) the parent's float64 local is shifted in float64 and added to the
// child's shadow (depth <= 2 internal nodes) or, rounded once, to its float32 row.
template <int P>
__global__ __launch_bounds__(kWave) void l2l_top_kernel(const int32_t* __restrict__ nodes,
                                                        const int32_t* __restrict__ lev_off, int32_t nlev,
                                                        const int32_t* __restrict__ top_nodes, int32_t ntop,
                                                        const int32_t* __restrict__ parent,
                                                        const int32_t* __restrict__ shadow,
                                                        const float* __restrict__ centers4,
                                                        const double* __restrict__ centers64,
                                                        float* __restrict__ locals, int32_t cs,
                                                        double* __restrict__ loc64, int32_t cs64) {
  constexpr int C = ncoef(P);
  const int t = threadIdx.x;
  if (t < ntop) {
    const int node = top_nodes[t];
    float tmp[C];
    double v[C];
    load_row<float, C>(locals + static_cast<int64_t>(node) * cs, tmp);
    for (int j = 0; j < C; ++j) v[j] = tmp[j];
    store_row<double, C>(loc64 + static_cast<int64_t>(shadow[node]) * cs64, v);
  }
  __syncthreads();
  for (int lev = 0; lev < nlev; ++lev) {
    const int count = lev_off[lev + 1] - lev_off[lev];
    if (t < count) {
      const int node = nodes[lev_off[lev] + t];
      const int par = parent[node];
      const int sp = shadow[par];
      const V4<double> cp = load4(centers64, sp);
      const int sn = shadow[node];
      V4<double> cn;
      if (sn >= 0) {
        cn = load4(centers64, sn);
      } else {
        const V4<float> f = load4(centers4, node);
        cn = V4<double>{f.x, f.y, f.z, f.w};
      }
      double v[C];
      load_row<double, C>(loc64 + static_cast<int64_t>(sp) * cs64, v);
      Align<double, P> a;
      a.init(cp.x - cn.x, cp.y - cn.y, cp.z - cn.z);
      rotate_all<double, P, Rot::ToZLocal>(v, a);
      zshift_l2l<double, P>(v, a.r);
      rotate_all<double, P, Rot::FromZLocal>(v, a);
      if (sn >= 0) {
        double w[C];
        load_row<double, C>(loc64 + static_cast<int64_t>(sn) * cs64, w);
        for (int j = 0; j < C; ++j) w[j] += v[j];
        store_row<double, C>(loc64 + static_cast<int64_t>(sn) * cs64, w);
        float wf[C];
        for (int j = 0; j < C; ++j) wf[j] = static_cast<float>(w[j]);
        store_row<float, C>(locals + static_cast<int64_t>(node) * cs, wf);
      } else {
        float w[C];
        load_row<float, C>(locals + static_cast<int64_t>(node) * cs, w);
        for (int j = 0; j < C; ++j) w[j] = static_cast<float>(static_cast<double>(w[j]) + v[j]);
        store_row<float, C>(locals + static_cast<int64_t>(node) * cs, w);
      }
    }
    __syncthreads();
  }
}

inline int blocks_for(int64_t n, int b) { return static_cast<int>((n + b - 1) / b); }

}  // namespace
}  // namespace jfm

using namespace jfm;

extern "C" {

int32_t jfm_coeff_stride(int32_t p, int32_t dtype) {
  if (p < 0 || p > JFM_MAX_ORDER || !valid_dtype(dtype)) return -1;
  return coef_stride(p, dtype == JFM_F32 ? 4 : 8);
}

int32_t jfm_order_supported(int32_t p, int32_t dtype) {
  return (p >= 1 && p <= JFM_MAX_ORDER && valid_dtype(dtype)) ? 1 : 0;
}

int32_t jfm_gather_sorted(void* stream, int32_t src_dtype, int32_t dtype, const void* positions, const void* masses,
                          const int32_t* perm, int32_t n, void* pos4) {
  if (n < 0 || (n > 0 && (!positions || !masses || !perm || !pos4)))
    return fail(JFM_ERR_ARG, "jfm_gather_sorted: bad arguments");
  if (!valid_dtype(src_dtype) || !valid_dtype(dtype)) return fail(JFM_ERR_ARG, "jfm_gather_sorted: bad dtype");
  if (n == 0) return JFM_OK;
  hipStream_t s = as_stream(stream);
  const int nb = blocks_for(n, 256);
  if (src_dtype == JFM_F32 && dtype == JFM_F32)
    gather_kernel<float, float><<<nb, 256, 0, s>>>((const float*)positions, (const float*)masses, perm, n, (float*)pos4);
  else if (src_dtype == JFM_F64 && dtype == JFM_F32)
    gather_kernel<double, float><<<nb, 256, 0, s>>>((const double*)positions, (const double*)masses, perm, n,
                                                   (float*)pos4);
  else if (src_dtype == JFM_F32 && dtype == JFM_F64)
    gather_kernel<float, double><<<nb, 256, 0, s>>>((const float*)positions, (const float*)masses, perm, n,
                                                   (double*)pos4);
  else
    gather_kernel<double, double><<<nb, 256, 0, s>>>((const double*)positions, (const double*)masses, perm, n,
                                                    (double*)pos4);
  return check_launch("jfm_gather_sorted");
}

int32_t jfm_p2m(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t num_leaves,
                int32_t leaf_size, int32_t num_internal, void* multipoles, void* centers4) {
  if (n < 1 || leaf_size < 1 || num_leaves != (n + leaf_size - 1) / leaf_size || num_internal != num_leaves - 1)
    return fail(JFM_ERR_ARG, "jfm_p2m: inconsistent sizes");
  if (!pos4 || !multipoles || !centers4) return fail(JFM_ERR_ARG, "jfm_p2m: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    const int cs = coef_stride(P, sizeof(T));
    if (leaf_size <= 8)
      p2m_group_kernel<T, P, 8><<<(num_leaves + 7) / 8, kWave, 0, s>>>(
          (const T*)pos4, n, leaf_size, num_leaves, num_internal, (T*)multipoles, cs, (T*)centers4);
    else if (leaf_size <= 16)
      p2m_group_kernel<T, P, 16><<<(num_leaves + 3) / 4, kWave, 0, s>>>(
          (const T*)pos4, n, leaf_size, num_leaves, num_internal, (T*)multipoles, cs, (T*)centers4);
    else if (leaf_size <= 32)
      p2m_group_kernel<T, P, 32><<<(num_leaves + 1) / 2, kWave, 0, s>>>(
          (const T*)pos4, n, leaf_size, num_leaves, num_internal, (T*)multipoles, cs, (T*)centers4);
    else
      p2m_kernel<T, P><<<num_leaves, kWave, 0, s>>>((const T*)pos4, n, leaf_size, num_internal, (T*)multipoles,
                                                    cs, (T*)centers4);
  }));
  return check_launch("jfm_p2m");
}

int32_t jfm_m2m_level(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t count,
                      const int32_t* left, const int32_t* right, const int32_t* node_range, void* multipoles,
                      void* centers4) {
  if (count < 0) return fail(JFM_ERR_ARG, "jfm_m2m_level: negative count");
  if (count == 0) return JFM_OK;
  if (!nodes || !left || !right || !node_range || !multipoles || !centers4)
    return fail(JFM_ERR_ARG, "jfm_m2m_level: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    const int cs = coef_stride(P, sizeof(T));
    m2m_kernel<T, P><<<blocks_for(2 * static_cast<int64_t>(count), kWave), kWave, 0, s>>>(nodes, count, left, right, node_range,
                                                                (T*)multipoles, cs, (T*)centers4);
  }));
  return check_launch("jfm_m2m_level");
}

int32_t jfm_l2l_level(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t count,
                      const int32_t* parent, const void* centers4, void* locals) {
  if (count < 0) return fail(JFM_ERR_ARG, "jfm_l2l_level: negative count");
  if (count == 0) return JFM_OK;
  if (!nodes || !parent || !centers4 || !locals) return fail(JFM_ERR_ARG, "jfm_l2l_level: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    const int cs = coef_stride(P, sizeof(T));
    l2l_kernel<T, P><<<blocks_for(count, kWave), kWave, 0, s>>>(nodes, count, parent, (const T*)centers4,
                                                                (T*)locals, cs);
  }));
  return check_launch("jfm_l2l_level");
}

int32_t jfm_m2m_top(void* stream, int32_t p, const int32_t* nodes, const int32_t* lev_off, int32_t nlev,
                    const int32_t* lev_off_host, const int32_t* left, const int32_t* right, const int32_t* node_range,
                    const int32_t* shadow, void* multipoles, void* centers4, void* mult64, void* centers64) {
  if (nlev < 0 || !lev_off_host) return fail(JFM_ERR_ARG, "jfm_m2m_top: bad level table");
  if (nlev == 0) return JFM_OK;
  for (int l = 0; l < nlev; ++l)
    if (lev_off_host[l + 1] - lev_off_host[l] > kWave / 2 || lev_off_host[l + 1] < lev_off_host[l])
      return fail(JFM_ERR_ARG, "jfm_m2m_top: a top level holds more than 32 nodes (two lanes per node)");
  if (!nodes || !lev_off || !left || !right || !node_range || !shadow || !multipoles || !centers4 || !mult64 ||
      !centers64)
    return fail(JFM_ERR_ARG, "jfm_m2m_top: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_ORDER(p, {
    m2m_top_kernel<P><<<1, kWave, 0, s>>>(nodes, lev_off, nlev, left, right, node_range, shadow, (float*)multipoles,
                                          coef_stride(P, 4), (float*)centers4, (double*)mult64, coef_stride(P, 8),
                                          (double*)centers64);
  });
  return check_launch("jfm_m2m_top");
}

int32_t jfm_l2l_top(void* stream, int32_t p, const int32_t* nodes, const int32_t* lev_off, int32_t nlev,
                    const int32_t* lev_off_host, const int32_t* top_nodes, int32_t ntop, const int32_t* parent,
                    const int32_t* shadow, const void* centers4, const void* centers64, void* locals, void* loc64) {
  if (nlev < 0 || ntop < 0 || ntop > kWave || !lev_off_host) return fail(JFM_ERR_ARG, "jfm_l2l_top: bad sizes");
  if (nlev == 0 && ntop == 0) return JFM_OK;
  for (int l = 0; l < nlev; ++l)
    if (lev_off_host[l + 1] - lev_off_host[l] > kWave || lev_off_host[l + 1] < lev_off_host[l])
      return fail(JFM_ERR_ARG, "jfm_l2l_top: a top level holds more than 64 nodes");
  if (!nodes || !lev_off || !top_nodes || !parent || !shadow || !centers4 || !centers64 || !locals || !loc64)
    return fail(JFM_ERR_ARG, "jfm_l2l_top: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_ORDER(p, {
    l2l_top_kernel<P><<<1, kWave, 0, s>>>(nodes, lev_off, nlev, top_nodes, ntop, parent, shadow,
                                          (const float*)centers4, (const double*)centers64, (float*)locals,
                                          coef_stride(P, 4), (double*)loc64, coef_stride(P, 8));
  });
  return check_launch("jfm_l2l_top");
}

int32_t jfm_l2p(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t leaf_size,
                int32_t num_internal, const void* locals, const void* centers4, double G, const int32_t* out_index,
                void* acc, void* pot, int32_t accumulate) {
  if (n < 1 || leaf_size < 1) return fail(JFM_ERR_ARG, "jfm_l2p: bad sizes");
  if (!pos4 || !locals || !centers4 || !acc) return fail(JFM_ERR_ARG, "jfm_l2p: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    const int cs = coef_stride(P, sizeof(T));
    l2p_kernel<T, P><<<blocks_for(n, 256), 256, 0, s>>>((const T*)pos4, n, leaf_size, num_internal,
                                                        (const T*)locals, cs, (const T*)centers4,
                                                        static_cast<T>(G), out_index, (T*)acc, (T*)pot,
                                                        accumulate);
  }));
  return check_launch("jfm_l2p");
}

int32_t jfm_l2p_combine(void* stream, int32_t dtype, int32_t n, double G, const void* far_acc, const void* far_pot,
                        void* acc, void* pot) {
  if (n < 1) return fail(JFM_ERR_ARG, "jfm_l2p_combine: bad size");
  if (!far_acc || !acc || (pot && !far_pot)) return fail(JFM_ERR_ARG, "jfm_l2p_combine: NULL buffer");
  hipStream_t s = as_stream(stream);
  const int64_t n3 = 3 * static_cast<int64_t>(n);
  const unsigned grid = static_cast<unsigned>((n3 + 255) / 256);
  if (dtype == JFM_F32)
    l2p_combine_kernel<float><<<grid, 256, 0, s>>>(n3, n, (const float*)far_acc, (const float*)far_pot,
                                                   static_cast<float>(G), (float*)acc, (float*)pot);
  else if (dtype == JFM_F64)
    l2p_combine_kernel<double><<<grid, 256, 0, s>>>(n3, n, (const double*)far_acc, (const double*)far_pot, G,
                                                    (double*)acc, (double*)pot);
  else
    return fail(JFM_ERR_ARG, "jfm_l2p_combine: bad dtype");
  return check_launch("jfm_l2p_combine");
}

}  // extern "C"
