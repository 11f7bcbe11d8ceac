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

// One thread per internal node of a level: centre of mass of the two children,
// M = m2m(M_left, c_left - c) + m2m(M_right, c_right - c).
template <typename T, int P>
__global__ __launch_bounds__(kWave) void m2m_kernel(const int32_t* __restrict__ nodes, int32_t count,
                                                    const int32_t* __restrict__ left,
                                                    const int32_t* __restrict__ right,
                                                    const int32_t* __restrict__ node_range, T* __restrict__ mult,
                                                    int32_t cs, T* __restrict__ centers4) {
  constexpr int C = ncoef(P);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int node = nodes[t];
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
  T acc[C];
#pragma unroll
  for (int j = 0; j < C; ++j) acc[j] = T(0);
#pragma unroll
  for (int side = 0; side < 2; ++side) {
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
    for (int j = 0; j < C; ++j) acc[j] += v[j];
  }
  store_row<T, C>(mult + static_cast<int64_t>(node) * cs, acc);
  store4(centers4, node, V4<T>{cx, cy, cz, tot});
}

// One thread per node of a level: L += l2l(L_parent, c_parent - c).
template <typename T, int P>
__global__ __launch_bounds__(kWave) void l2l_kernel(const int32_t* __restrict__ nodes, int32_t count,
                                                    const int32_t* __restrict__ parent,
                                                    const T* __restrict__ centers4, T* __restrict__ locals,
                                                    int32_t cs) {
  constexpr int C = ncoef(P);
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const int node = nodes[t];
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

// One thread per particle: far field from its leaf's local expansion.
// a = -G grad_delta phi, pot = -G phi, delta = centre - x.  Writes (not adds)
// the output row.
template <typename T, int P>
__global__ __launch_bounds__(256) void l2p_kernel(const T* __restrict__ pos4, int32_t n, int32_t leaf_size,
                                                  int32_t num_internal, const T* __restrict__ locals, int32_t cs,
                                                  const T* __restrict__ centers4, T G,
                                                  const int32_t* __restrict__ out_index, T* __restrict__ acc,
                                                  T* __restrict__ pot) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int node = num_internal + i / leaf_size;
  const V4<T> q = load4(pos4, i);
  const V4<T> c = load4(centers4, node);
  const T* F = locals + static_cast<int64_t>(node) * cs;
  T phi, gx, gy, gz;
  l2p_eval<T, P>([F](int j) { return F[j]; }, c.x - q.x, c.y - q.y, c.z - q.z, phi, gx, gy, gz);
  const int64_t row = out_index ? out_index[i] : i;
  acc[3 * row + 0] = -G * gx;
  acc[3 * row + 1] = -G * gy;
  acc[3 * row + 2] = -G * gz;
  if (pot) pot[row] = -G * phi;
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
    p2m_kernel<T, P><<<num_leaves, kWave, 0, s>>>((const T*)pos4, n, leaf_size, num_internal, (T*)multipoles, cs,
                                                  (T*)centers4);
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
    m2m_kernel<T, P><<<blocks_for(count, kWave), kWave, 0, s>>>(nodes, count, left, right, node_range,
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

int32_t jfm_l2p(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t leaf_size,
                int32_t num_internal, const void* locals, const void* centers4, double G, const int32_t* out_index,
                void* acc, void* pot) {
  if (n < 1 || leaf_size < 1) return fail(JFM_ERR_ARG, "jfm_l2p: bad sizes");
  if (!pos4 || !locals || !centers4 || !acc) return fail(JFM_ERR_ARG, "jfm_l2p: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ORDER(p, {
    const int cs = coef_stride(P, sizeof(T));
    l2p_kernel<T, P><<<blocks_for(n, 256), 256, 0, s>>>((const T*)pos4, n, leaf_size, num_internal,
                                                        (const T*)locals, cs, (const T*)centers4,
                                                        static_cast<T>(G), out_index, (T*)acc, (T*)pot);
  }));
  return check_launch("jfm_l2p");
}

}  // extern "C"
