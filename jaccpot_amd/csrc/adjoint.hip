// Reverse-mode (adjoint) kernels of the fixed-topology FMM: exact gradients of
// accelerations and potentials with respect to positions and masses.
//
// Structure.  M2M and L2L are exact re-expansions for the degrees kept, so
// the far field equals sum over far pairs (t <- s) of
//   L2P( M2L( P2M(particles of s, c_s), c_t - c_s ), c_t - x_i ),  i in t,
// and the centres matter only through the far pairs.  The reverse pass is:
//   leaves:    x_i += G (H a_i + grad phi phib_i) (L2P, Hessian-vector product
//              through the gradient-coefficient identities), Lbar_leaf =
//              -G sum_i (phib_i U(d_i) + abar_i . grad U(d_i)), d_i = c - x_i;
//   up:        Lbar_parent += l2l^T(Lbar_child) = m2m(...)   (forward M2M on
//              parity-flipped rows);
//   pairs:     Mbar_s = sum_t m2l^T(Lbar_t) = P m2l(P Lbar_t, c_s - c_t)
//              (forward M2L on flipped rows, m2l is symmetric);
//   centres:   cbar_t = Gr(L_t^own) . Lbar_t - Gr(Mbar_t^own) . M_t
//              + sum_s d/dd [Lbar_t . m2l(M_s, d)] - sum_s d/dd [Lbar_s . m2l(M_t, d)]
//              (this file: center_adjoint_kernel; d/dd [A . m2l_p(B, d)] =
//              -A . Gr(m2l_{p+1}(B, d)), exact and valid on the z axis);
//   down:      Mbar_child += m2m^T(Mbar_parent) = l2l(...) (forward L2L, flipped);
//   leaves:    x_j += m_j grad(Mbar . U)(x_j - c), m_j += Mbar . U(x_j - c)
//              (P2M), plus the centre-of-mass chain of every ancestor.
// P2P has its own kernel (p2p_adjoint_kernel).  Everything is deterministic.
//
// Identities used (packed real basis): op(w, -d) = P op(P w, d) for the three
// translations with P = (-1)^n, m2l^T = m2l, l2l^T = m2m at the same d.

#include <hip/hip_runtime.h>

#include "device_math.hpp"
#include "launch.hpp"
#include "rows.hpp"

// Orders with adjoint kernels (every order of the forward path; the centre
// pass translates at order p + 1, for which the B_U tables reach degree 9).
#define JFM_FOR_ADJ_ORDER(p, ...)                                                      \
  switch (p) {                                                                         \
    case 1: { constexpr int P = 1; __VA_ARGS__; } break;                               \
    case 2: { constexpr int P = 2; __VA_ARGS__; } break;                               \
    case 3: { constexpr int P = 3; __VA_ARGS__; } break;                               \
    case 4: { constexpr int P = 4; __VA_ARGS__; } break;                               \
    case 5: { constexpr int P = 5; __VA_ARGS__; } break;                               \
    case 6: { constexpr int P = 6; __VA_ARGS__; } break;                               \
    case 7: { constexpr int P = 7; __VA_ARGS__; } break;                               \
    case 8: { constexpr int P = 8; __VA_ARGS__; } break;                               \
    default:                                                                           \
      return ::jfm::fail(JFM_ERR_UNSUPPORTED, "adjoint kernels exist for orders 1..8"); \
  }

namespace jfm {
namespace {

constexpr int kWave = 64;
constexpr int kThreads = 256;

// Gradient-coefficient map: sum_j F_j dU_j/d(axis) == sum_j G_j U_j with G of degree <= P-1.
// Transpose of the degree-lowering identities of l2p_eval.
template <typename T, int P>
__host__ __device__ __forceinline__ void grad_coeffs(const T* F, T* Gx, T* Gy, T* Gz) {
  constexpr int C1 = ncoef(P - 1);
  for (int j = 0; j < C1; ++j) Gx[j] = Gy[j] = Gz[j] = T(0);
  for (int n = 1; n <= P; ++n) {
    // m = 0: dx U_n^0 = -U_{n-1}^1, dy U_n^0 = -U_{n-1}^{-1}, dz U_n^0 = U_{n-1}^0
    const T f0 = F[sh_idx(n, 0)];
    Gz[sh_idx(n - 1, 0)] += f0;
    if (n >= 2) {
      Gx[sh_idx(n - 1, 1)] -= f0;
      Gy[sh_idx(n - 1, -1)] -= f0;
    }
    for (int m = 1; m <= n; ++m) {
      const T fc = F[sh_idx(n, m)], fs = F[sh_idx(n, -m)];
      // dz
      if (m <= n - 1) {
        Gz[sh_idx(n - 1, m)] += fc;
        Gz[sh_idx(n - 1, -m)] += fs;
      }
      // dx U_n^m = (U_{n-1}^{m-1} - U_{n-1}^{m+1}) / 2 ; dx U_n^-m = (U_{n-1}^{-(m-1)} - U_{n-1}^{-(m+1)}) / 2
      Gx[sh_idx(n - 1, m - 1)] += T(0.5) * fc;
      if (m - 1 >= 1) Gx[sh_idx(n - 1, -(m - 1))] += T(0.5) * fs;
      if (m + 1 <= n - 1) {
        Gx[sh_idx(n - 1, m + 1)] -= T(0.5) * fc;
        Gx[sh_idx(n - 1, -(m + 1))] -= T(0.5) * fs;
      }
      // dy U_n^m = -(U_{n-1}^{-(m+1)} + U_{n-1}^{-(m-1)}) / 2 ; dy U_n^-m = (U_{n-1}^{m+1} + U_{n-1}^{m-1}) / 2
      if (m - 1 >= 1) Gy[sh_idx(n - 1, -(m - 1))] -= T(0.5) * fc;
      Gy[sh_idx(n - 1, m - 1)] += T(0.5) * fs;
      if (m + 1 <= n - 1) {
        Gy[sh_idx(n - 1, -(m + 1))] -= T(0.5) * fc;
        Gy[sh_idx(n - 1, m + 1)] += T(0.5) * fs;
      }
    }
  }
}

// out_j = w U_j(d) + sum_a abar_a dU_j/d(axis a) (d), all degrees <= P.
template <typename T, int P>
__device__ __forceinline__ void weighted_harmonics(T dx, T dy, T dz, T w, T bx, T by, T bz, T* out) {
  T re[ntri(P)], im[ntri(P)];
  solid_harmonics<T, P>(dx, dy, dz, re, im);
  for (int n = 0; n <= P; ++n) {
    {
      T v = w * re[tri(n, 0)];
      if (n >= 1) v = fma(bz, re[tri(n - 1, 0)], v);
      if (n >= 2) v = fma(-bx, re[tri(n - 1, 1)], fma(-by, im[tri(n - 1, 1)], v));
      out[sh_idx(n, 0)] = v;
    }
    for (int m = 1; m <= n; ++m) {
      const T ar = re[tri(n - 1, m - 1)], ai = im[tri(n - 1, m - 1)];
      const T br = (m + 1 <= n - 1) ? re[tri(n - 1, m + 1)] : T(0);
      const T bi = (m + 1 <= n - 1) ? im[tri(n - 1, m + 1)] : T(0);
      const T zr = (m <= n - 1) ? re[tri(n - 1, m)] : T(0);
      const T zi = (m <= n - 1) ? im[tri(n - 1, m)] : T(0);
      out[sh_idx(n, m)] = w * re[tri(n, m)] + bx * T(0.5) * (ar - br) - by * T(0.5) * (ai + bi) + bz * zr;
      out[sh_idx(n, -m)] = w * im[tri(n, m)] + bx * T(0.5) * (ai - bi) + by * T(0.5) * (ar + br) + bz * zi;
    }
  }
}

// ------------------------------------------------------------------ P2P
// xbar_i += G [ (I s3 - 3 d d^T s5)(m_i abar_j - m_j abar_i) - d s3 (m_j phib_i + m_i phib_j) ]
// mbar_i += G [ -d s3 . abar_j - s phib_j ],  d = x_j - x_i, over self (j != i) + near leaves.
//
// Work items as in the forward P2P: item k covers slots [s0, s1) of leaf
// `leaf`; part < 0 -> the item stores the leaf's rows of grad4, else its
// unscaled partial sums go to partials[(part * leaf_size + i) * 4 + c] and
// p2p_adjoint_reduce_kernel adds them in item order.  leaf == NULL: one item
// per leaf.
// 1/sqrt(x) for x > 0 (x == 0 is mapped to 0 by the SAFE guard).  float32:
// v_rsq_f32, as in the forward P2P; float64: the v_rsq_f64 estimate y with
// one third-order step, y (1 + e/2 + 3e^2/8) with e = 1 - x y^2 (the relative
// error is cubed).
template <typename T>
__device__ __forceinline__ T fast_rsqrt(T x) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_amdgcn_rsqf(x);
  } else {
    const T y = __builtin_amdgcn_rsq(x);
    const T e = fma(-(x * y), y, T(1));
    return fma(y * e, fma(e, T(0.375), T(0.5)), y);
  }
}

template <typename T>
struct AdjItems {
  const int32_t* leaf;
  const int32_t* s0;
  const int32_t* s1;
  const int32_t* part;
  T* partials;
  int32_t num_part;
};

// Block-step level weights of the mutual lane: pair (i, j) is scaled by
// level_w[clamp(max(rung[i], rung[j]), 0, nlev - 1)] (rungs in tree order for
// particles, per node for cells).
template <typename T>
struct PairLevels {
  const int32_t* rung;
  const T* level_w;
  int32_t nlev;
};

constexpr int kMaxLevelWeights = 64;

template <typename T, int WT, bool SAFE, bool WEIGHTED = false>
__global__ __launch_bounds__(kThreads) void p2p_adjoint_kernel(const T* __restrict__ pos4, const T* __restrict__ cot4,
                                                               int32_t n, int32_t leaf_size,
                                                               const int32_t* __restrict__ near_off,
                                                               const int32_t* __restrict__ near_src, T eps2, T G,
                                                               T* __restrict__ grad4, AdjItems<T> items,
                                                               PairLevels<T> lv = PairLevels<T>{}) {
  constexpr int NPART = kThreads / WT;
  constexpr int TILE = sizeof(T) == 4 ? 512 : 256;
  constexpr int LOADS = TILE / kThreads;
  static_assert(2 * kThreads <= 2 * TILE, "reduction buffer");
  __shared__ V4<T> tx[TILE], tc[TILE];
  __shared__ int32_t tr[WEIGHTED ? TILE : 1];
  __shared__ T lw[WEIGHTED ? kMaxLevelWeights : 1];
  const int leaf = items.leaf ? items.leaf[blockIdx.x] : blockIdx.x;
  const int i = threadIdx.x % WT, part = threadIdx.x / WT;
  const int first = leaf * leaf_size;
  const int cnt = min(leaf_size, n - first);
  const int me = first + (i < cnt ? i : 0);
  const V4<T> xi = load4(pos4, me);
  const V4<T> ci = load4(cot4, me);
  int ri = 0;
  if constexpr (WEIGHTED) {
    ri = lv.rung[me];
    if (threadIdx.x < lv.nlev) lw[threadIdx.x] = lv.level_w[threadIdx.x];
  }
  const int nb = near_off[leaf];
  const int start = items.leaf ? items.s0[blockIdx.x] : 0;
  const int total = items.leaf ? items.s1[blockIdx.x] : (1 + near_off[leaf + 1] - nb) * leaf_size;
  const int part_out = items.leaf ? items.part[blockIdx.x] : -1;
  T gx = 0, gy = 0, gz = 0, gm = 0;
  for (int base = start; base < total; base += TILE) {
    __syncthreads();
    for (int l = 0; l < LOADS; ++l) {
      const int slot = base + l * kThreads + threadIdx.x;
      V4<T> q{T(0), T(0), T(0), T(0)}, c{T(0), T(0), T(0), T(0)};
      int r = 0;
      if (slot < total) {
        const int qi = slot / leaf_size;
        const int o = slot - qi * leaf_size;
        const int sl = qi == 0 ? leaf : near_src[nb + qi - 1];
        const int g = sl * leaf_size + o;
        if (g < n) {
          q = load4(pos4, g);
          c = load4(cot4, g);
          if constexpr (WEIGHTED) r = lv.rung[g];
        }
      }
      tx[l * kThreads + threadIdx.x] = q;
      tc[l * kThreads + threadIdx.x] = c;
      if constexpr (WEIGHTED) tr[l * kThreads + threadIdx.x] = r;
    }
    __syncthreads();
    const int trips = (min(TILE, total - base) + NPART - 1) / NPART;
    // per pair: xbar += s3 u - (3 s5 (d.u) + s3 pw) d,  mbar -= s3 (d.abar_j) + s phib_j
    // (the radial terms share one factor; s = 1/r from fast_rsqrt)
#pragma unroll 4
    for (int k = 0; k < trips; ++k) {
      const V4<T> s = tx[part + k * NPART];
      const V4<T> c = tc[part + k * NPART];
      const T dx = s.x - xi.x, dy = s.y - xi.y, dz = s.z - xi.z;
      const T r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, eps2)));
      T sv = fast_rsqrt(r2);
      if (SAFE) sv = r2 > T(0) ? sv : T(0);
      const T sv2 = sv * sv;
      T s3 = sv2 * sv;
      T s5 = s3 * sv2;
      if constexpr (WEIGHTED) {  // every term is linear in (s, s3, s5)
        const T w = lw[min(max(max(ri, tr[part + k * NPART]), 0), lv.nlev - 1)];
        sv *= w;
        s3 *= w;
        s5 *= w;
      }
      // u = m_i abar_j - m_j abar_i
      const T ux = fma(xi.w, c.x, -s.w * ci.x), uy = fma(xi.w, c.y, -s.w * ci.y), uz = fma(xi.w, c.z, -s.w * ci.z);
      const T du = fma(dx, ux, fma(dy, uy, dz * uz));
      const T pw = fma(s.w, ci.w, xi.w * c.w);
      const T a = fma(T(3) * s5, du, s3 * pw);
      gx = fma(s3, ux, fma(-a, dx, gx));
      gy = fma(s3, uy, fma(-a, dy, gy));
      gz = fma(s3, uz, fma(-a, dz, gz));
      const T dc = fma(dx, c.x, fma(dy, c.y, dz * c.z));
      gm = fma(-s3, dc, fma(-sv, c.w, gm));
    }
  }
  __syncthreads();
  tx[threadIdx.x] = V4<T>{gx, gy, gz, gm};
  __syncthreads();
  if (part == 0 && i < cnt) {
    T ax = 0, ay = 0, az = 0, am = 0;
    for (int q = 0; q < NPART; ++q) {
      const V4<T> v = tx[q * WT + i];
      ax += v.x;
      ay += v.y;
      az += v.z;
      am += v.w;
    }
    if (part_out >= 0) {
      if (part_out < items.num_part)
        store4(items.partials, static_cast<int64_t>(part_out) * leaf_size + i, V4<T>{ax, ay, az, am});
      return;
    }
    if (eps2 > T(0)) am += ci.w / dev_sqrt(eps2);  // the self term is not in the forward potential
    store4(grad4, first + i, V4<T>{G * ax, G * ay, G * az, G * am});
  }
}

template <typename T>
__device__ __forceinline__ void kahan_add(T& s, T& c, T v) {
  const T y = v - c;
  const T t = s + y;
  c = (t - s) - y;
  s = t;
}

// One workgroup per split leaf: 256 / leaf_size lanes per target add strided
// item partials (compensated), lane 0 adds the lane sums in lane order.
template <typename T>
__global__ __launch_bounds__(kThreads) void p2p_adjoint_reduce_kernel(const T* __restrict__ cot4, int32_t n,
                                                                      int32_t leaf_size,
                                                                      const int32_t* __restrict__ split_leaf,
                                                                      const int32_t* __restrict__ p0,
                                                                      const int32_t* __restrict__ p1,
                                                                      const T* __restrict__ partials, T eps2, T G,
                                                                      T* __restrict__ grad4) {
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
  if (eps2 > T(0)) t[3] += cot4[4 * static_cast<int64_t>(first + i) + 3] / dev_sqrt(eps2);
  store4(grad4, first + i, V4<T>{G * t[0], G * t[1], G * t[2], G * t[3]});
}

// ------------------------------------------------------------- leaf adjoint A
// One wave per leaf: x_i += G (H abar_i + grad phi phib_i) into grad4, and the
// leaf's Lbar row (parity-flipped for the M2M upward pass) as the wave sum of
// -G (phib_i U(d_i) + abar_i . grad U(d_i)).
template <typename T, int P>
__global__ __launch_bounds__(kWave) void leaf_adjoint_a_kernel(const T* __restrict__ pos4,
                                                               const T* __restrict__ cot4, int32_t n,
                                                               int32_t leaf_size, int32_t num_internal,
                                                               const T* __restrict__ locals, int32_t cs,
                                                               const T* __restrict__ centers4, T G,
                                                               T* __restrict__ grad4, T* __restrict__ lbar_flip) {
  constexpr int C = ncoef(P);
  __shared__ T red[kWave][C + 1];
  const int leaf = blockIdx.x, lane = threadIdx.x;
  const int node = num_internal + leaf;
  const int first = leaf * leaf_size;
  const int cnt = min(leaf_size, n - first);
  const V4<T> c = load4(centers4, node);
  T F[C];
  load_row<T, C>(locals + static_cast<int64_t>(node) * cs, F);
  T Gx[ncoef(P - 1)], Gy[ncoef(P - 1)], Gz[ncoef(P - 1)];
  grad_coeffs<T, P>(F, Gx, Gy, Gz);
  T acc[C];
  for (int j = 0; j < C; ++j) acc[j] = T(0);
  for (int i = lane; i < cnt; i += kWave) {
    const V4<T> q = load4(pos4, first + i);
    const V4<T> cb = load4(cot4, first + i);
    const T dx = c.x - q.x, dy = c.y - q.y, dz = c.z - q.z;
    T phi, g1x, g1y, g1z;
    l2p_eval<T, P>([&F](int j) { return F[j]; }, dx, dy, dz, phi, g1x, g1y, g1z);
    // H abar = grad of (sum_a abar_a G^a) . U
    T V[ncoef(P - 1)];
    for (int j = 0; j < ncoef(P - 1); ++j) V[j] = cb.x * Gx[j] + cb.y * Gy[j] + cb.z * Gz[j];
    T phv, hx, hy, hz;
    l2p_eval<T, P - 1>([&V](int j) { return V[j]; }, dx, dy, dz, phv, hx, hy, hz);
    V4<T> g = load4(grad4, first + i);
    g.x += G * (hx + g1x * cb.w);
    g.y += G * (hy + g1y * cb.w);
    g.z += G * (hz + g1z * cb.w);
    store4(grad4, first + i, g);
    T w[C];
    weighted_harmonics<T, P>(dx, dy, dz, cb.w, cb.x, cb.y, cb.z, w);
    for (int j = 0; j < C; ++j) acc[j] -= G * w[j];
  }
  for (int j = 0; j < C; ++j) red[lane][j] = acc[j];
  __syncthreads();
  for (int j = lane; j < cs; j += kWave) {
    T s = T(0);
    if (j < C) {
      for (int l = 0; l < kWave; ++l) s += red[l][j];
      int deg = 0;
      while ((deg + 1) * (deg + 1) <= j) ++deg;
      if (deg & 1) s = -s;
    }
    lbar_flip[static_cast<int64_t>(node) * cs + j] = s;
  }
}

// --------------------------------------------------------- centre cotangents
// d/dd [ A . m2l(B, d) ] along `axis` by a dual-number pass of the forward cascade.
template <typename T, int P>
__device__ __forceinline__ void m2l_bilinear_grad(const T* A, const T* B, T dx, T dy, T dz, T& gx, T& gy, T& gz) {
  // d/dd [A . m2l_P(B, d)] = -A . Gr(m2l_{P+1}(B, d)), Gr the local-gradient
  // map: the separation derivative of a truncated translation is, term by
  // term, the gradient of the translation one order higher.  No derivative of
  // the alignment angles is involved, so it also holds for separations on the
  // z axis, where the angles are fixed by convention.
  constexpr int C = ncoef(P), C1 = ncoef(P + 1);
  T v[C1], w[C1];
  for (int j = 0; j < C1; ++j) v[j] = j < C ? B[j] : T(0);
  Align<T, P + 1> a;
  a.init(dx, dy, dz);
  rotate_all<T, P + 1, Rot::ToZMultipole>(v, a);
  zm2l<T, P + 1>(v, w, a.r);
  rotate_all<T, P + 1, Rot::FromZLocal>(w, a);
  T Gx[C], Gy[C], Gz[C];
  grad_coeffs<T, P + 1>(w, Gx, Gy, Gz);
  T sx = T(0), sy = T(0), sz = T(0);
  for (int j = 0; j < C; ++j) {
    sx = fma(A[j], Gx[j], sx);
    sy = fma(A[j], Gy[j], sy);
    sz = fma(A[j], Gz[j], sz);
  }
  gx = -sx;
  gy = -sy;
  gz = -sz;
}

// m2l_bilinear_grad at a run-time order g in [G, P] (rows hold at least
// ncoef(g) coefficients; only that prefix is read).
template <typename T, int G, int P>
__device__ __forceinline__ void m2l_bilinear_grad_at(int g, const T* A, const T* B, T dx, T dy, T dz, T& gx, T& gy,
                                                     T& gz) {
  if constexpr (G < P) {
    if (g <= G) {
      m2l_bilinear_grad<T, G>(A, B, dx, dy, dz, gx, gy, gz);
      return;
    }
    m2l_bilinear_grad_at<T, G + 1, P>(g, A, B, dx, dy, dz, gx, gy, gz);
  } else {
    m2l_bilinear_grad<T, P>(A, B, dx, dy, dz, gx, gy, gz);
  }
}

// One wave per node u with far partners: lanes over partners v, summing
//   + d/dd [Lbar_u . m2l(M_v, d)] at d = c_u - c_v      (pair u <- v)
//   - d/dd [Lbar_v . m2l(M_u, d)] at d = c_v - c_u      (pair v <- u)
// and lane 0 adds Gr(Lown_u) . Lbar_u - Gr(Mbar_own_u) . M_u.  Writes cbar4[u] (xyz).
// GEARS: pair k = (u <- v) of the CSR ran at order gear_fwd[k] and its mirror
// (v <- u) at gear_rev[k]; each bilinear term runs at its pair's order (the
// order-g operator keeps the terms n + k <= g, so it is not the order-P one
// with rows cut at degree g).
// WEIGHTED: both terms of pair (u, v) are scaled by its level weight (the
// mutual lane; L_own and Mbar_own then carry the weights already).
template <typename T, int P, bool GEARS, bool WEIGHTED = false>
__global__ __launch_bounds__(kWave) void center_adjoint_kernel(
    const int32_t* __restrict__ nodes, const int32_t* __restrict__ far_off, const int32_t* __restrict__ far_src,
    const int8_t* __restrict__ gear_fwd, const int8_t* __restrict__ gear_rev, const T* __restrict__ mult,
    const T* __restrict__ lown, const T* __restrict__ lbar_flip, const T* __restrict__ mbar_own_flip, int32_t cs,
    const T* __restrict__ centers4, T* __restrict__ cbar4, PairLevels<T> lv = PairLevels<T>{}) {
  constexpr int C = ncoef(P);
  __shared__ T red[kWave][4];
  const int u = nodes[blockIdx.x];
  const int lane = threadIdx.x;
  const V4<T> cu = load4(centers4, u);
  T Lu[C], Mu[C];
  load_row<T, C>(lbar_flip + static_cast<int64_t>(u) * cs, Lu);
  load_row<T, C>(mult + static_cast<int64_t>(u) * cs, Mu);
  for (int j = 0; j < C; ++j) {
    int deg = 0;
    while ((deg + 1) * (deg + 1) <= j) ++deg;
    if (deg & 1) Lu[j] = -Lu[j];
  }
  T gx = 0, gy = 0, gz = 0;
  const int b = far_off[u], e = far_off[u + 1];
  for (int k = b + lane; k < e; k += kWave) {
    const int v = far_src[k];
    const V4<T> cv = load4(centers4, v);
    T Mv[C], Lv[C];
    load_row<T, C>(mult + static_cast<int64_t>(v) * cs, Mv);
    load_row<T, C>(lbar_flip + static_cast<int64_t>(v) * cs, Lv);
    for (int j = 0; j < C; ++j) {
      int deg = 0;
      while ((deg + 1) * (deg + 1) <= j) ++deg;
      if (deg & 1) Lv[j] = -Lv[j];
    }
    const T dx = cu.x - cv.x, dy = cu.y - cv.y, dz = cu.z - cv.z;
    T ax, ay, az, bx, by, bz;
    if constexpr (GEARS) {
      m2l_bilinear_grad_at<T, 1, P>(gear_fwd[k], Lu, Mv, dx, dy, dz, ax, ay, az);
      m2l_bilinear_grad_at<T, 1, P>(gear_rev[k], Lv, Mu, -dx, -dy, -dz, bx, by, bz);
    } else {
      m2l_bilinear_grad<T, P>(Lu, Mv, dx, dy, dz, ax, ay, az);
      m2l_bilinear_grad<T, P>(Lv, Mu, -dx, -dy, -dz, bx, by, bz);
    }
    if constexpr (WEIGHTED) {
      const T w = lv.level_w[min(max(max(lv.rung[u], lv.rung[v]), 0), lv.nlev - 1)];
      gx = fma(w, ax - bx, gx);
      gy = fma(w, ay - by, gy);
      gz = fma(w, az - bz, gz);
    } else {
      gx += ax - bx;
      gy += ay - by;
      gz += az - bz;
    }
  }
  red[lane][0] = gx;
  red[lane][1] = gy;
  red[lane][2] = gz;
  __syncthreads();
  if (lane == 0) {
    T sx = 0, sy = 0, sz = 0;
    for (int l = 0; l < kWave; ++l) {
      sx += red[l][0];
      sy += red[l][1];
      sz += red[l][2];
    }
    // displacement terms of L2P (target side) and P2M (source side)
    T Lo[C], Mb[C];
    load_row<T, C>(lown + static_cast<int64_t>(u) * cs, Lo);
    load_row<T, C>(mbar_own_flip + static_cast<int64_t>(u) * cs, Mb);
    for (int j = 0; j < C; ++j) {
      int deg = 0;
      while ((deg + 1) * (deg + 1) <= j) ++deg;
      if (deg & 1) Mb[j] = -Mb[j];
    }
    T Ax[ncoef(P - 1)], Ay[ncoef(P - 1)], Az[ncoef(P - 1)];
    grad_coeffs<T, P>(Lo, Ax, Ay, Az);
    for (int j = 0; j < ncoef(P - 1); ++j) {
      sx = fma(Ax[j], Lu[j], sx);
      sy = fma(Ay[j], Lu[j], sy);
      sz = fma(Az[j], Lu[j], sz);
    }
    grad_coeffs<T, P>(Mb, Ax, Ay, Az);
    for (int j = 0; j < ncoef(P - 1); ++j) {
      sx = fma(-Ax[j], Mu[j], sx);
      sy = fma(-Ay[j], Mu[j], sy);
      sz = fma(-Az[j], Mu[j], sz);
    }
    store4(cbar4, u, V4<T>{sx, sy, sz, T(0)});
  }
}

// ------------------------------------------------- centre-of-mass chain per leaf
// For every ancestor a of the leaf (and the leaf): c_a = sum m_i x_i / m_a
// (count-weighted mean when m_a == 0).  Accumulates
//   um = sum cbar_a / m_a,  uc = sum cbar_a / n_a (massless),  vv = sum c_a . cbar_a / m_a.
template <typename T>
__global__ __launch_bounds__(kThreads) void com_chain_kernel(int32_t num_leaves, int32_t num_internal,
                                                             const int32_t* __restrict__ parent,
                                                             const int32_t* __restrict__ node_range,
                                                             const T* __restrict__ centers4,
                                                             const T* __restrict__ cbar4, T* __restrict__ chain) {
  const int l = blockIdx.x * kThreads + threadIdx.x;
  if (l >= num_leaves) return;
  T umx = 0, umy = 0, umz = 0, ucx = 0, ucy = 0, ucz = 0, vv = 0;
  for (int a = num_internal + l; a >= 0; a = parent[a]) {
    const V4<T> c = load4(centers4, a);
    const V4<T> g = load4(cbar4, a);
    if (c.w > T(0)) {
      const T inv = T(1) / c.w;
      umx += g.x * inv;
      umy += g.y * inv;
      umz += g.z * inv;
      vv += (c.x * g.x + c.y * g.y + c.z * g.z) * inv;
    } else {
      const T inv = T(1) / static_cast<T>(node_range[2 * a + 1] - node_range[2 * a] + 1);
      ucx += g.x * inv;
      ucy += g.y * inv;
      ucz += g.z * inv;
    }
  }
  store4(chain, 2 * l, V4<T>{umx, umy, umz, vv});
  store4(chain, 2 * l + 1, V4<T>{ucx, ucy, ucz, T(0)});
}

// ------------------------------------------------------------- leaf adjoint B
// Per particle: P2M adjoint with the pushed-down Mbar (flipped rows), the COM
// chain, then scatter (xbar, mbar) to input order.
template <typename T, int P>
__global__ __launch_bounds__(kThreads) void leaf_adjoint_b_kernel(const T* __restrict__ pos4, int32_t n,
                                                                  int32_t leaf_size, int32_t num_internal,
                                                                  const T* __restrict__ mbar_flip, int32_t cs,
                                                                  const T* __restrict__ centers4,
                                                                  const T* __restrict__ chain,
                                                                  const T* __restrict__ grad4,
                                                                  const int32_t* __restrict__ out_index,
                                                                  T* __restrict__ gpos, T* __restrict__ gmass) {
  constexpr int C = ncoef(P);
  const int i = blockIdx.x * kThreads + threadIdx.x;
  if (i >= n) return;
  const int l = i / leaf_size;
  const int node = num_internal + l;
  const V4<T> q = load4(pos4, i);
  const V4<T> c = load4(centers4, node);
  T Mb[C];
  load_row<T, C>(mbar_flip + static_cast<int64_t>(node) * cs, Mb);
  for (int j = 0; j < C; ++j) {
    int deg = 0;
    while ((deg + 1) * (deg + 1) <= j) ++deg;
    if (deg & 1) Mb[j] = -Mb[j];
  }
  T phi, gx, gy, gz;
  l2p_eval<T, P>([&Mb](int j) { return Mb[j]; }, q.x - c.x, q.y - c.y, q.z - c.z, phi, gx, gy, gz);
  const V4<T> um = load4(chain, 2 * l);
  const V4<T> uc = load4(chain, 2 * l + 1);
  const V4<T> g = load4(grad4, i);
  const T xb = g.x + q.w * gx + q.w * um.x + uc.x;
  const T yb = g.y + q.w * gy + q.w * um.y + uc.y;
  const T zb = g.z + q.w * gz + q.w * um.z + uc.z;
  const T mb = g.w + phi + q.x * um.x + q.y * um.y + q.z * um.z - um.w;
  const int64_t row = out_index ? out_index[i] : i;
  gpos[3 * row + 0] = xb;
  gpos[3 * row + 1] = yb;
  gpos[3 * row + 2] = zb;
  gmass[row] = mb;
}

template <typename T>
int32_t launch_p2p_adj(hipStream_t s, const T* pos4, const T* cot4, int32_t n, int32_t leaf_size,
                       int32_t num_blocks, const int32_t* near_off, const int32_t* near_src, T eps2, T G, T* grad4,
                       AdjItems<T> items = AdjItems<T>{}, PairLevels<T> lv = PairLevels<T>{}) {
  const bool safe = !(eps2 > T(0));
  const bool weighted = lv.level_w != nullptr;
#define JFM_ADJ_W(WT, SAFE, W)                                                                                   \
  p2p_adjoint_kernel<T, WT, SAFE, W><<<num_blocks, kThreads, 0, s>>>(pos4, cot4, n, leaf_size, near_off, near_src, \
                                                                     eps2, G, grad4, items, lv)
#define JFM_ADJ(WT)                 \
  {                                 \
    if (weighted && safe)           \
      JFM_ADJ_W(WT, true, true);    \
    else if (weighted)              \
      JFM_ADJ_W(WT, false, true);   \
    else if (safe)                  \
      JFM_ADJ_W(WT, true, false);   \
    else                            \
      JFM_ADJ_W(WT, false, false);  \
    return check_launch("jfm_p2p_adjoint"); \
  }
  if (leaf_size <= 16) JFM_ADJ(16)
  if (leaf_size <= 32) JFM_ADJ(32)
  if (leaf_size <= 64) JFM_ADJ(64)
  if (leaf_size <= 128) JFM_ADJ(128)
  JFM_ADJ(256)
#undef JFM_ADJ
#undef JFM_ADJ_W
}

inline int nblk(int64_t n, int b) { return static_cast<int>((n + b - 1) / b); }

}  // namespace
}  // namespace jfm

using namespace jfm;

extern "C" {

int32_t jfm_p2p_adjoint(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                        int32_t num_leaves, int32_t leaf_size, const int32_t* near_off, const int32_t* near_src,
                        double eps2, double G, void* grad4) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_leaves != (n + leaf_size - 1) / leaf_size)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint: inconsistent sizes");
  if (!pos4 || !cot4 || !near_off || !near_src || !grad4) return fail(JFM_ERR_ARG, "jfm_p2p_adjoint: NULL buffer");
  hipStream_t s = as_stream(stream);
  if (dtype == JFM_F32)
    return launch_p2p_adj<float>(s, (const float*)pos4, (const float*)cot4, n, leaf_size, num_leaves, near_off,
                                 near_src, static_cast<float>(eps2), static_cast<float>(G), (float*)grad4);
  if (dtype == JFM_F64)
    return launch_p2p_adj<double>(s, (const double*)pos4, (const double*)cot4, n, leaf_size, num_leaves, near_off,
                                  near_src, eps2, G, (double*)grad4);
  return fail(JFM_ERR_ARG, "jfm_p2p_adjoint: bad dtype");
}

int32_t jfm_p2p_adjoint_items(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                              int32_t num_leaves, int32_t leaf_size, int32_t num_items, const int32_t* item_leaf,
                              const int32_t* item_s0, const int32_t* item_s1, const int32_t* item_part,
                              int32_t num_part, void* partials, const int32_t* near_off, const int32_t* near_src,
                              double eps2, double G, void* grad4) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_leaves != (n + leaf_size - 1) / leaf_size || num_items < 1 ||
      num_part < 0)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_items: inconsistent sizes");
  if (!pos4 || !cot4 || !near_off || !near_src || !grad4 || !item_leaf || !item_s0 || !item_s1 || !item_part ||
      (num_part > 0 && !partials))
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_items: NULL buffer");
  hipStream_t s = as_stream(stream);
  if (dtype == JFM_F32)
    return launch_p2p_adj<float>(s, (const float*)pos4, (const float*)cot4, n, leaf_size, num_items, near_off,
                                 near_src, static_cast<float>(eps2), static_cast<float>(G), (float*)grad4,
                                 AdjItems<float>{item_leaf, item_s0, item_s1, item_part, (float*)partials, num_part});
  if (dtype == JFM_F64)
    return launch_p2p_adj<double>(s, (const double*)pos4, (const double*)cot4, n, leaf_size, num_items, near_off,
                                  near_src, eps2, G, (double*)grad4,
                                  AdjItems<double>{item_leaf, item_s0, item_s1, item_part, (double*)partials,
                                                   num_part});
  return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_items: bad dtype");
}

int32_t jfm_p2p_adjoint_reduce(void* stream, int32_t dtype, const void* cot4, int32_t n, int32_t leaf_size,
                               int32_t num_split, const int32_t* split_leaf, const int32_t* split_p0,
                               const int32_t* split_p1, const void* partials, double eps2, double G, void* grad4) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_split < 0)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_reduce: sizes");
  if (num_split == 0) return JFM_OK;
  if (!cot4 || !split_leaf || !split_p0 || !split_p1 || !partials || !grad4)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_reduce: NULL buffer");
  hipStream_t s = as_stream(stream);
  if (dtype == JFM_F32)
    p2p_adjoint_reduce_kernel<float><<<num_split, kThreads, 0, s>>>(
        (const float*)cot4, n, leaf_size, split_leaf, split_p0, split_p1, (const float*)partials,
        static_cast<float>(eps2), static_cast<float>(G), (float*)grad4);
  else if (dtype == JFM_F64)
    p2p_adjoint_reduce_kernel<double><<<num_split, kThreads, 0, s>>>(
        (const double*)cot4, n, leaf_size, split_leaf, split_p0, split_p1, (const double*)partials, eps2, G,
        (double*)grad4);
  else
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_reduce: bad dtype");
  return check_launch("jfm_p2p_adjoint_reduce");
}

int32_t jfm_leaf_adjoint_a(void* stream, int32_t dtype, int32_t p, const void* pos4, const void* cot4, int32_t n,
                           int32_t leaf_size, int32_t num_internal, const void* locals, const void* centers4,
                           double G, void* grad4, void* lbar_flip) {
  if (n < 1 || leaf_size < 1) return fail(JFM_ERR_ARG, "jfm_leaf_adjoint_a: bad sizes");
  if (!pos4 || !cot4 || !locals || !centers4 || !grad4 || !lbar_flip)
    return fail(JFM_ERR_ARG, "jfm_leaf_adjoint_a: NULL buffer");
  const int L = (n + leaf_size - 1) / leaf_size;
  if (num_internal != L - 1) return fail(JFM_ERR_ARG, "jfm_leaf_adjoint_a: num_internal must be L-1");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ADJ_ORDER(p, {
    leaf_adjoint_a_kernel<T, P><<<L, kWave, 0, s>>>((const T*)pos4, (const T*)cot4, n, leaf_size, num_internal,
                                                    (const T*)locals, coef_stride(P, sizeof(T)),
                                                    (const T*)centers4, static_cast<T>(G), (T*)grad4,
                                                    (T*)lbar_flip);
  }));
  return check_launch("jfm_leaf_adjoint_a");
}

int32_t jfm_center_adjoint(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t num_nodes,
                           const int32_t* far_off, const int32_t* far_src, const void* multipoles,
                           const void* locals_own, const void* lbar_flip, const void* mbar_own_flip,
                           const void* centers4, void* cbar4) {
  if (num_nodes < 0) return fail(JFM_ERR_ARG, "jfm_center_adjoint: bad sizes");
  if (num_nodes == 0) return JFM_OK;
  if (!nodes || !far_off || !far_src || !multipoles || !locals_own || !lbar_flip || !mbar_own_flip || !centers4 ||
      !cbar4)
    return fail(JFM_ERR_ARG, "jfm_center_adjoint: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ADJ_ORDER(p, {
    center_adjoint_kernel<T, P, false><<<num_nodes, kWave, 0, s>>>(
        nodes, far_off, far_src, nullptr, nullptr, (const T*)multipoles, (const T*)locals_own,
        (const T*)lbar_flip, (const T*)mbar_own_flip, coef_stride(P, sizeof(T)), (const T*)centers4, (T*)cbar4);
  }));
  return check_launch("jfm_center_adjoint");
}

int32_t jfm_center_adjoint_weighted(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t num_nodes,
                                    const int32_t* far_off, const int32_t* far_src, const int32_t* node_rung,
                                    const void* level_weights, int32_t num_levels, const void* multipoles,
                                    const void* locals_own, const void* lbar_flip, const void* mbar_own_flip,
                                    const void* centers4, void* cbar4) {
  if (num_nodes < 0 || num_levels < 1) return fail(JFM_ERR_ARG, "jfm_center_adjoint_weighted: bad sizes");
  if (num_nodes == 0) return JFM_OK;
  if (!nodes || !far_off || !far_src || !node_rung || !level_weights || !multipoles || !locals_own || !lbar_flip ||
      !mbar_own_flip || !centers4 || !cbar4)
    return fail(JFM_ERR_ARG, "jfm_center_adjoint_weighted: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ADJ_ORDER(p, {
    const PairLevels<T> lv{node_rung, (const T*)level_weights, num_levels};
    center_adjoint_kernel<T, P, false, true><<<num_nodes, kWave, 0, s>>>(
        nodes, far_off, far_src, nullptr, nullptr, (const T*)multipoles, (const T*)locals_own, (const T*)lbar_flip,
        (const T*)mbar_own_flip, coef_stride(P, sizeof(T)), (const T*)centers4, (T*)cbar4, lv);
  }));
  return check_launch("jfm_center_adjoint_weighted");
}

int32_t jfm_p2p_adjoint_weighted(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                                 int32_t num_leaves, int32_t leaf_size, const int32_t* near_off,
                                 const int32_t* near_src, const int32_t* rung_sorted, const void* level_weights,
                                 int32_t num_levels, double eps2, double G, void* grad4) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_leaves != (n + leaf_size - 1) / leaf_size)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted: inconsistent sizes");
  if (num_levels < 1 || num_levels > kMaxLevelWeights)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted: 1..64 level weights");
  if (!pos4 || !cot4 || !near_off || !near_src || !rung_sorted || !level_weights || !grad4)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted: NULL buffer");
  hipStream_t s = as_stream(stream);
  if (dtype == JFM_F32)
    return launch_p2p_adj<float>(s, (const float*)pos4, (const float*)cot4, n, leaf_size, num_leaves, near_off,
                                 near_src, static_cast<float>(eps2), static_cast<float>(G), (float*)grad4,
                                 AdjItems<float>{}, PairLevels<float>{rung_sorted, (const float*)level_weights,
                                                                      num_levels});
  if (dtype == JFM_F64)
    return launch_p2p_adj<double>(s, (const double*)pos4, (const double*)cot4, n, leaf_size, num_leaves, near_off,
                                  near_src, eps2, G, (double*)grad4, AdjItems<double>{},
                                  PairLevels<double>{rung_sorted, (const double*)level_weights, num_levels});
  return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted: bad dtype");
}

int32_t jfm_p2p_adjoint_weighted_items(void* stream, int32_t dtype, const void* pos4, const void* cot4, int32_t n,
                                       int32_t num_leaves, int32_t leaf_size, int32_t num_items,
                                       const int32_t* item_leaf, const int32_t* item_s0, const int32_t* item_s1,
                                       const int32_t* item_part, int32_t num_part, void* partials,
                                       const int32_t* near_off, const int32_t* near_src, const int32_t* rung_sorted,
                                       const void* level_weights, int32_t num_levels, double eps2, double G,
                                       void* grad4) {
  if (n < 1 || leaf_size < 1 || leaf_size > 256 || num_leaves != (n + leaf_size - 1) / leaf_size || num_items < 1 ||
      num_part < 0)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted_items: inconsistent sizes");
  if (num_levels < 1 || num_levels > kMaxLevelWeights)
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted_items: 1..64 level weights");
  if (!pos4 || !cot4 || !near_off || !near_src || !rung_sorted || !level_weights || !grad4 || !item_leaf ||
      !item_s0 || !item_s1 || !item_part || (num_part > 0 && !partials))
    return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted_items: NULL buffer");
  hipStream_t s = as_stream(stream);
  if (dtype == JFM_F32)
    return launch_p2p_adj<float>(s, (const float*)pos4, (const float*)cot4, n, leaf_size, num_items, near_off,
                                 near_src, static_cast<float>(eps2), static_cast<float>(G), (float*)grad4,
                                 AdjItems<float>{item_leaf, item_s0, item_s1, item_part, (float*)partials, num_part},
                                 PairLevels<float>{rung_sorted, (const float*)level_weights, num_levels});
  if (dtype == JFM_F64)
    return launch_p2p_adj<double>(s, (const double*)pos4, (const double*)cot4, n, leaf_size, num_items, near_off,
                                  near_src, eps2, G, (double*)grad4,
                                  AdjItems<double>{item_leaf, item_s0, item_s1, item_part, (double*)partials,
                                                   num_part},
                                  PairLevels<double>{rung_sorted, (const double*)level_weights, num_levels});
  return fail(JFM_ERR_ARG, "jfm_p2p_adjoint_weighted_items: bad dtype");
}

int32_t jfm_center_adjoint_gears(void* stream, int32_t dtype, int32_t p, const int32_t* nodes, int32_t num_nodes,
                                 const int32_t* far_off, const int32_t* far_src, const int8_t* gear_fwd,
                                 const int8_t* gear_rev, const void* multipoles, const void* locals_own,
                                 const void* lbar_flip, const void* mbar_own_flip, const void* centers4,
                                 void* cbar4) {
  if (num_nodes < 0) return fail(JFM_ERR_ARG, "jfm_center_adjoint_gears: bad sizes");
  if (num_nodes == 0) return JFM_OK;
  if (!nodes || !far_off || !far_src || !gear_fwd || !gear_rev || !multipoles || !locals_own || !lbar_flip ||
      !mbar_own_flip || !centers4 || !cbar4)
    return fail(JFM_ERR_ARG, "jfm_center_adjoint_gears: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ADJ_ORDER(p, {
    center_adjoint_kernel<T, P, true><<<num_nodes, kWave, 0, s>>>(
        nodes, far_off, far_src, gear_fwd, gear_rev, (const T*)multipoles, (const T*)locals_own,
        (const T*)lbar_flip, (const T*)mbar_own_flip, coef_stride(P, sizeof(T)), (const T*)centers4, (T*)cbar4);
  }));
  return check_launch("jfm_center_adjoint_gears");
}

int32_t jfm_com_chain(void* stream, int32_t dtype, int32_t num_leaves, int32_t num_internal, const int32_t* parent,
                      const int32_t* node_range, const void* centers4, const void* cbar4, void* chain) {
  if (num_leaves < 1 || num_internal != num_leaves - 1) return fail(JFM_ERR_ARG, "jfm_com_chain: bad sizes");
  if (!parent || !node_range || !centers4 || !cbar4 || !chain) return fail(JFM_ERR_ARG, "jfm_com_chain: NULL");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, {
    com_chain_kernel<T><<<nblk(num_leaves, kThreads), kThreads, 0, s>>>(num_leaves, num_internal, parent,
                                                                        node_range, (const T*)centers4,
                                                                        (const T*)cbar4, (T*)chain);
  });
  return check_launch("jfm_com_chain");
}

int32_t jfm_leaf_adjoint_b(void* stream, int32_t dtype, int32_t p, const void* pos4, int32_t n, int32_t leaf_size,
                           int32_t num_internal, const void* mbar_flip, const void* centers4, const void* chain,
                           const void* grad4, const int32_t* out_index, void* gpos, void* gmass) {
  if (n < 1 || leaf_size < 1) return fail(JFM_ERR_ARG, "jfm_leaf_adjoint_b: bad sizes");
  if (!pos4 || !mbar_flip || !centers4 || !chain || !grad4 || !gpos || !gmass)
    return fail(JFM_ERR_ARG, "jfm_leaf_adjoint_b: NULL buffer");
  hipStream_t s = as_stream(stream);
  JFM_FOR_DTYPE(dtype, JFM_FOR_ADJ_ORDER(p, {
    leaf_adjoint_b_kernel<T, P><<<nblk(n, kThreads), kThreads, 0, s>>>(
        (const T*)pos4, n, leaf_size, num_internal, (const T*)mbar_flip, coef_stride(P, sizeof(T)),
        (const T*)centers4, (const T*)chain, (const T*)grad4, out_index, (T*)gpos, (T*)gmass);
  }));
  return check_launch("jfm_leaf_adjoint_b");
}

}  // extern "C"
