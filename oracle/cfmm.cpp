// CPU restatement of the fixed-order real-basis FMM numerics, C++17 + OpenMP.
//
// Test infrastructure and CPU baseline only: it checks the HIP path at sizes
// the NumPy oracle cannot reach and times the reference's plain CPU pipeline
// on host cores.  It is never linked into, loaded by or called from the
// product library.
//
// The algorithm is the one oracle/fmm.py and oracle/harmonics.py state (each
// step cites the reference file it follows there):
//   gather      x, m into tree (Morton) order, SoA
//   P2M         M_leaf = sum_i m_i U(x_i - c_leaf)            real_p2m_l2p.py:51-232
//   M2M         deepest level first, M_par = sum_children m2m  real_translations.py:528-598
//   M2L         per target node over its far list, list order  real_translations.py:381-449
//   L2L         shallowest level first, L_child += l2l(L_par)  real_translations.py:601-667
//   L2P         a = -G grad_delta phi, delta = c - x           real_p2m_l2p.py:241-547
//   P2P         self leaf (i != j) + near leaves, softened     nearfield/_kernels.py:40-338
// with the rotation-based translations D_to^M = B_U Dz(-ax) B_U Dz(az),
// D_from^M = Dz(-az) B_U Dz(ax) B_U and their transposes for locals
// (real_rotations.py:417-649), az = atan2(x, y), ax = atan2(rho, z) taken
// from their sines and cosines (no trigonometric calls), the z-axis
// translations of real_translations.py:72-367, and the z-M2L evaluated as
// r^-(n+1) sum_k (n+k)! (M_k r^-k) (same value as (n+k)!/r^(n+k+1), without
// forming r^-(2p+1), which underflows float32 for deep nodes).
//
// B_U comes in from the caller (the NumPy oracle's Dehnen recursion), as
// dense per-degree blocks; factorials are exp(lgamma(n+1)) as in the oracle.
// Pinned against oracle/fmm.py in float64 (tests/test_cfmm.py).

#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include <omp.h>

namespace {

constexpr int kPMax = 8;
constexpr int kCMax = (kPMax + 1) * (kPMax + 1);
constexpr int kLanes = 8;  // P2P targets per SIMD block

inline int sh(int n, int m) { return n * n + n + m; }

// ---------------------------------------------------------------------------
// tables
// ---------------------------------------------------------------------------

template <class T>
struct Tables {
  int p = 0;
  T fact[2 * kPMax + 2];
  T dfact[kPMax + 1];  // (2m-1)!! = (2m)! / (2^m m!)
  // B_U per degree as sparse rows: for degree l, row i of the block lists
  // (col, val) for the nonzero entries; bt_* is the transpose.
  std::vector<int> b_off, b_col, bt_off, bt_col;
  std::vector<T> b_val, bt_val;
  int b_row0[kPMax + 1];  // first row slot of degree l in b_off / bt_off
};

template <class T>
void make_tables(Tables<T>& tb, int p, const double* bu) {
  tb.p = p;
  for (int n = 0; n < 2 * kPMax + 2; ++n) tb.fact[n] = T(std::exp(std::lgamma(n + 1.0)));
  for (int m = 0; m <= kPMax; ++m) {
    const double f2m = std::exp(std::lgamma(2.0 * m + 1.0)), fm = std::exp(std::lgamma(m + 1.0));
    tb.dfact[m] = T(f2m / (std::pow(2.0, m) * fm));
  }
  tb.b_off.assign(1, 0);
  tb.bt_off.assign(1, 0);
  tb.b_col.clear(); tb.b_val.clear(); tb.bt_col.clear(); tb.bt_val.clear();
  const double* blk = bu;
  for (int l = 0; l <= p; ++l) {
    const int w = 2 * l + 1;
    tb.b_row0[l] = int(tb.b_off.size()) - 1;
    for (int i = 0; i < w; ++i) {
      for (int j = 0; j < w; ++j)
        if (blk[i * w + j] != 0.0) { tb.b_col.push_back(j); tb.b_val.push_back(T(blk[i * w + j])); }
      tb.b_off.push_back(int(tb.b_col.size()));
      for (int j = 0; j < w; ++j)
        if (blk[j * w + i] != 0.0) { tb.bt_col.push_back(j); tb.bt_val.push_back(T(blk[j * w + i])); }
      tb.bt_off.push_back(int(tb.bt_col.size()));
    }
    blk += w * w;
  }
}

// ---------------------------------------------------------------------------
// solid harmonics U_n^m (value only, and value + gradient by dual numbers),
// oracle/harmonics.py _solid_harmonics_generic
// ---------------------------------------------------------------------------

template <class T>
struct Dual {
  T v, d[3];
};
template <class T> inline Dual<T> dconst(T c) { return {c, {0, 0, 0}}; }
template <class T> inline Dual<T> operator+(Dual<T> a, Dual<T> b) {
  return {a.v + b.v, {a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]}};
}
template <class T> inline Dual<T> operator-(Dual<T> a, Dual<T> b) {
  return {a.v - b.v, {a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]}};
}
template <class T> inline Dual<T> operator*(Dual<T> a, Dual<T> b) {
  return {a.v * b.v, {a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]}};
}
template <class T> inline Dual<T> operator*(T s, Dual<T> a) { return {s * a.v, {s * a.d[0], s * a.d[1], s * a.d[2]}}; }
template <class T> inline Dual<T> operator/(Dual<T> a, Dual<T> b) {
  const T q = a.v / b.v;
  return {q, {(a.d[0] - q * b.d[0]) / b.v, (a.d[1] - q * b.d[1]) / b.v, (a.d[2] - q * b.d[2]) / b.v}};
}
template <class T> inline Dual<T> dsqrt(Dual<T> a) {
  const T s = std::sqrt(a.v);
  const T h = T(0.5) / s;
  return {s, {a.d[0] * h, a.d[1] * h, a.d[2] * h}};
}
template <class T> inline Dual<T> dfloor(Dual<T> a, T f) { return a.v < f ? dconst(f) : a; }
template <class T> inline T sqrt_(T a) { return std::sqrt(a); }
template <class T> inline Dual<T> sqrt_(Dual<T> a) { return dsqrt(a); }
template <class T> inline T floor_(T a, T f) { return a < f ? f : a; }
template <class T> inline Dual<T> floor_(Dual<T> a, T f) { return dfloor(a, f); }
template <class T> inline T cst(T, T c) { return c; }
template <class T> inline Dual<T> cst(Dual<T>, T c) { return dconst(c); }
template <class T> inline T mul(T s, T a) { return s * a; }
template <class T> inline Dual<T> mul(T s, Dual<T> a) { return s * a; }

template <class T>
inline T radius_floor() {
  return std::max(T(1e-60) > T(0) ? T(1e-60) : T(0), std::numeric_limits<T>::min());
}

// U[sh(n,m)] for n <= p; S is T or Dual<T>.
template <class T, class S>
void solid_harmonics(const Tables<T>& tb, int p, S x, S y, S z, S* U) {
  const T fl = radius_floor<T>();
  const S r = sqrt_(floor_(x * x + y * y + z * z, fl));
  const S rho = sqrt_(floor_(x * x + y * y, fl));
  const S cos_t = z / r, sin_t = rho / r, cos_p = x / rho, sin_p = y / rho;
  S cm[kPMax + 1], sm[kPMax + 1], rn[kPMax + 1], st[kPMax + 1];
  cm[0] = cst(x, T(1));
  sm[0] = cst(x, T(0));
  if (p >= 1) { cm[1] = cos_p; sm[1] = sin_p; }
  for (int m = 2; m <= p; ++m) {
    cm[m] = mul(T(2), cos_p * cm[m - 1]) - cm[m - 2];
    sm[m] = mul(T(2), cos_p * sm[m - 1]) - sm[m - 2];
  }
  rn[0] = cst(x, T(1));
  st[0] = cst(x, T(1));
  for (int n = 1; n <= p; ++n) { rn[n] = rn[n - 1] * r; st[n] = st[n - 1] * sin_t; }
  for (int ma = 0; ma <= p; ++ma) {
    const S pmm = ma == 0 ? cst(x, T(1)) : mul(tb.dfact[ma], st[ma]);
    S p2 = pmm, p1 = pmm;
    for (int n = ma; n <= p; ++n) {
      S pnm;
      if (n == ma) {
        pnm = pmm;
      } else if (n == ma + 1) {
        pnm = mul(T(2 * ma + 1), cos_t * pmm);
        p2 = pmm;
        p1 = pnm;
      } else {
        pnm = (mul(T(2 * n - 1), cos_t * p1) - mul(T(n + ma - 1), p2)) / cst(x, T(n - ma));
        p2 = p1;
        p1 = pnm;
      }
      const T inv = T(1) / tb.fact[n + ma];
      U[sh(n, ma)] = mul(inv, rn[n] * pnm * cm[ma]);
      if (ma > 0) U[sh(n, -ma)] = mul(inv, rn[n] * pnm * sm[ma]);
    }
  }
}

// ---------------------------------------------------------------------------
// rotations and z-translations
// ---------------------------------------------------------------------------

// cos / sin of m*a for m = 0..p from (cos a, sin a) by angle addition.
template <class T>
inline void trig_ladder(int p, T c1, T s1, T* c, T* s) {
  c[0] = T(1);
  s[0] = T(0);
  for (int m = 1; m <= p; ++m) {
    c[m] = c[m - 1] * c1 - s[m - 1] * s1;
    s[m] = s[m - 1] * c1 + c[m - 1] * s1;
  }
}

// v <- Dz(a) v for every degree (sgn = -1 applies Dz(-a)).
template <class T>
inline void apply_dz(int p, const T* c, const T* s, T sgn, T* v) {
  for (int l = 1; l <= p; ++l) {
    T* b = v + l * l + l;
    for (int m = 1; m <= l; ++m) {
      const T cc = c[m], ss = sgn * s[m];
      const T vp = b[m], vm = b[-m];
      b[m] = cc * vp - ss * vm;
      b[-m] = ss * vp + cc * vm;
    }
  }
}

// v <- B_U v (transpose = false) or B_U^T v.
template <class T>
inline void apply_b(const Tables<T>& tb, int p, bool transpose, T* v) {
  const std::vector<int>& off = transpose ? tb.bt_off : tb.b_off;
  const std::vector<int>& col = transpose ? tb.bt_col : tb.b_col;
  const std::vector<T>& vv = transpose ? tb.bt_val : tb.b_val;
  T tmp[2 * kPMax + 1];
  for (int l = 1; l <= p; ++l) {
    const int w = 2 * l + 1;
    T* b = v + l * l;
    for (int i = 0; i < w; ++i) {
      const int row = tb.b_row0[l] + i;
      T acc = T(0);
      for (int e = off[row]; e < off[row + 1]; ++e) acc += vv[e] * b[col[e]];
      tmp[i] = acc;
    }
    std::memcpy(b, tmp, sizeof(T) * w);
  }
}

template <class T>
struct Frame {
  T ca[kPMax + 1], sa[kPMax + 1];  // cos/sin(m az)
  T cx[kPMax + 1], sx[kPMax + 1];  // cos/sin(m ax)
};

// Angles of real_rotations.py:340-414: az = atan2(x, y) (0 at rho = 0),
// ax = atan2(rho, z) (0 at r = 0), as cos/sin ladders.
template <class T>
inline void make_frame(int p, T x, T y, T z, Frame<T>& f) {
  const T rho2 = x * x + y * y;
  const T r2 = rho2 + z * z;
  T c1 = T(1), s1 = T(0), c2 = T(1), s2 = T(0);
  if (rho2 > T(0)) {
    const T rho = std::sqrt(rho2);
    c1 = y / rho;
    s1 = x / rho;
  }
  if (r2 > T(0)) {
    const T r = std::sqrt(r2);
    c2 = z / r;
    s2 = std::sqrt(rho2) / r;
  }
  trig_ladder(p, c1, s1, f.ca, f.sa);
  trig_ladder(p, c2, s2, f.cx, f.sx);
}

// D_to^M = B_U Dz(-ax) B_U Dz(az)
template <class T>
inline void to_z_multipole(const Tables<T>& tb, int p, const Frame<T>& f, T* v) {
  apply_dz(p, f.ca, f.sa, T(1), v);
  apply_b(tb, p, false, v);
  apply_dz(p, f.cx, f.sx, T(-1), v);
  apply_b(tb, p, false, v);
}
// D_from^M = Dz(-az) B_U Dz(ax) B_U
template <class T>
inline void from_z_multipole(const Tables<T>& tb, int p, const Frame<T>& f, T* v) {
  apply_b(tb, p, false, v);
  apply_dz(p, f.cx, f.sx, T(1), v);
  apply_b(tb, p, false, v);
  apply_dz(p, f.ca, f.sa, T(-1), v);
}
// local world -> z = (D_from^M)^T = B_U^T Dz(-ax) B_U^T Dz(az)
template <class T>
inline void to_z_local(const Tables<T>& tb, int p, const Frame<T>& f, T* v) {
  apply_dz(p, f.ca, f.sa, T(1), v);
  apply_b(tb, p, true, v);
  apply_dz(p, f.cx, f.sx, T(-1), v);
  apply_b(tb, p, true, v);
}
// local z -> world = (D_to^M)^T = Dz(-az) B_U^T Dz(ax) B_U^T
template <class T>
inline void from_z_local(const Tables<T>& tb, int p, const Frame<T>& f, T* v) {
  apply_b(tb, p, true, v);
  apply_dz(p, f.cx, f.sx, T(1), v);
  apply_b(tb, p, true, v);
  apply_dz(p, f.ca, f.sa, T(-1), v);
}

template <class T>
inline T floored_radius(T x, T y, T z) {
  return std::sqrt(std::max(x * x + y * y + z * z, radius_floor<T>()));
}

// child multipole -> contribution to the parent; delta = child - parent
template <class T>
void m2m(const Tables<T>& tb, int p, const T* M, T dx, T dy, T dz, T* out) {
  const int C = (p + 1) * (p + 1);
  T v[kCMax], w[kCMax];
  std::memcpy(v, M, sizeof(T) * C);
  Frame<T> f;
  make_frame(p, dx, dy, dz, f);
  to_z_multipole(tb, p, f, v);
  const T r = floored_radius(dx, dy, dz);
  const T d = r < T(1e-30) ? T(0) : r;
  T ck[kPMax + 1];
  ck[0] = T(1);
  for (int k = 1; k <= p; ++k) ck[k] = ck[k - 1] * d;
  for (int k = 0; k <= p; ++k) ck[k] /= tb.fact[k];
  for (int n = 0; n <= p; ++n)
    for (int m = -n; m <= n; ++m) {
      T acc = T(0);
      for (int k = 0; k <= n - std::abs(m); ++k) acc += v[sh(n - k, m)] * ck[k];
      w[sh(n, m)] = acc;
    }
  from_z_multipole(tb, p, f, w);
  for (int c = 0; c < C; ++c) out[c] += w[c];
}

// parent local -> contribution to the child; delta = parent - child
template <class T>
void l2l(const Tables<T>& tb, int p, const T* L, T dx, T dy, T dz, T* out) {
  const int C = (p + 1) * (p + 1);
  T v[kCMax], w[kCMax];
  std::memcpy(v, L, sizeof(T) * C);
  Frame<T> f;
  make_frame(p, dx, dy, dz, f);
  to_z_local(tb, p, f, v);
  const T r = floored_radius(dx, dy, dz);
  const T d = r < T(1e-30) ? T(0) : r;
  T ck[kPMax + 1];
  ck[0] = T(1);
  for (int k = 1; k <= p; ++k) ck[k] = ck[k - 1] * d;
  for (int k = 0; k <= p; ++k) ck[k] /= tb.fact[k];
  for (int n = 0; n <= p; ++n)
    for (int m = -n; m <= n; ++m) {
      T acc = T(0);
      for (int k = 0; k <= p - n; ++k)
        if (n + k >= std::abs(m)) acc += v[sh(n + k, m)] * ck[k];
      w[sh(n, m)] = acc;
    }
  from_z_local(tb, p, f, w);
  for (int c = 0; c < C; ++c) out[c] += w[c];
}

// source multipole -> target local; delta = target - source
template <class T>
void m2l(const Tables<T>& tb, int p, const T* M, T dx, T dy, T dz, T* out) {
  const int C = (p + 1) * (p + 1);
  T v[kCMax], w[kCMax];
  std::memcpy(v, M, sizeof(T) * C);
  Frame<T> f;
  make_frame(p, dx, dy, dz, f);
  to_z_multipole(tb, p, f, v);
  const T r2 = dx * dx + dy * dy + dz * dz;
  const T r = std::max(r2 > T(0) ? std::sqrt(r2) : T(0), T(1e-30));
  const T ir = T(1) / r;
  T irp[2 * kPMax + 2];
  irp[0] = T(1);
  for (int k = 1; k <= 2 * p + 1; ++k) irp[k] = irp[k - 1] * ir;
  for (int n = 0; n <= p; ++n)
    for (int m = -n; m <= n; ++m) {
      const int ma = std::abs(m);
      const T s = T((ma & 1) ? -1 : 1) * T(m != 0 ? 2 : 1);
      T acc = T(0);
      for (int k = ma; k <= p - n; ++k) acc += tb.fact[n + k] * (v[sh(k, m)] * irp[k]);
      w[sh(n, m)] = s * irp[n + 1] * acc;
    }
  from_z_local(tb, p, f, w);
  for (int c = 0; c < C; ++c) out[c] += w[c];
}

// ---------------------------------------------------------------------------
// the pipeline
// ---------------------------------------------------------------------------

struct Topo {
  int n, I, L, nodes, p;
  const int32_t *perm, *node_range, *left, *right, *parent, *depth, *far_off, *far_src, *near_off, *near_src;
  const int8_t* orders;  // per far pair expansion order (NULL: p for every pair)
  const double* center;
};

template <class T>
int run(const Topo& tp, const double* bu, const double* x, const double* mass, double G, double eps,
        double* acc_out, double* pot_out, double* seconds) {
  const int p = tp.p, C = (p + 1) * (p + 1), n = tp.n, I = tp.I, L = tp.L, nodes = tp.nodes;
  for (int l = 0; l < L; ++l)
    if (tp.node_range[2 * (I + l) + 1] - tp.node_range[2 * (I + l)] + 1 > 256) return 3;  // P2P target buffers
  Tables<T> tb;
  make_tables(tb, p, bu);
  double t[8];
  int ti = 0;
  t[ti++] = omp_get_wtime();

  // gather into tree order, SoA
  std::vector<T> xs(n), ys(n), zs(n), ms(n), cen(3 * size_t(nodes));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const int k = tp.perm[i];
    xs[i] = T(x[3 * size_t(k)]);
    ys[i] = T(x[3 * size_t(k) + 1]);
    zs[i] = T(x[3 * size_t(k) + 2]);
    ms[i] = T(mass[k]);
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < 3 * nodes; ++k) cen[k] = T(tp.center[k]);
  t[ti++] = omp_get_wtime();

  // P2M
  std::vector<T> M(size_t(nodes) * C, T(0)), Lc(size_t(nodes) * C, T(0));
#pragma omp parallel for schedule(dynamic, 64)
  for (int l = 0; l < L; ++l) {
    const int node = I + l;
    const int a = tp.node_range[2 * node], b = tp.node_range[2 * node + 1];
    T* out = &M[size_t(node) * C];
    T U[kCMax];
    for (int i = a; i <= b; ++i) {
      solid_harmonics<T, T>(tb, p, xs[i] - cen[3 * node], ys[i] - cen[3 * node + 1], zs[i] - cen[3 * node + 2], U);
      for (int c = 0; c < C; ++c) out[c] += ms[i] * U[c];
    }
  }
  t[ti++] = omp_get_wtime();

  // M2M by level, deepest first
  int max_depth = 0;
  for (int k = 0; k < nodes; ++k) max_depth = std::max(max_depth, int(tp.depth[k]));
  std::vector<std::vector<int>> by_depth(max_depth + 1);
  for (int k = 0; k < nodes; ++k) by_depth[tp.depth[k]].push_back(k);
  for (int d = max_depth; d >= 0; --d) {
    const std::vector<int>& lv = by_depth[d];
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t q = 0; q < lv.size(); ++q) {
      const int node = lv[q];
      if (node >= I) continue;
      T* out = &M[size_t(node) * C];
      for (int side = 0; side < 2; ++side) {
        const int ch = side == 0 ? tp.left[node] : tp.right[node];
        m2m(tb, p, &M[size_t(ch) * C], cen[3 * ch] - cen[3 * node], cen[3 * ch + 1] - cen[3 * node + 1],
            cen[3 * ch + 2] - cen[3 * node + 2], out);
      }
    }
  }
  t[ti++] = omp_get_wtime();

  // M2L per target node, far list order
#pragma omp parallel for schedule(dynamic, 32)
  for (int tgt = 0; tgt < nodes; ++tgt) {
    const int e0 = tp.far_off[tgt], e1 = tp.far_off[tgt + 1];
    if (e0 == e1) continue;
    T* out = &Lc[size_t(tgt) * C];
    for (int e = e0; e < e1; ++e) {
      const int s = tp.far_src[e];
      // a per-pair order g translates the degrees <= g of the source at order g
      // and adds into the degrees <= g of the target (oracle/fmm.py downward)
      const int g = tp.orders ? int(tp.orders[e]) : p;
      m2l(tb, g, &M[size_t(s) * C], cen[3 * tgt] - cen[3 * s], cen[3 * tgt + 1] - cen[3 * s + 1],
          cen[3 * tgt + 2] - cen[3 * s + 2], out);
    }
  }
  t[ti++] = omp_get_wtime();

  // L2L by level, shallowest first
  for (int d = 1; d <= max_depth; ++d) {
    const std::vector<int>& lv = by_depth[d];
#pragma omp parallel for schedule(dynamic, 16)
    for (size_t q = 0; q < lv.size(); ++q) {
      const int node = lv[q], par = tp.parent[node];
      l2l(tb, p, &Lc[size_t(par) * C], cen[3 * par] - cen[3 * node], cen[3 * par + 1] - cen[3 * node + 1],
          cen[3 * par + 2] - cen[3 * node + 2], &Lc[size_t(node) * C]);
    }
  }
  t[ti++] = omp_get_wtime();

  // L2P + P2P per leaf, outputs scattered to input order
  const T eps2 = T(eps) * T(eps), g = T(G);
#pragma omp parallel for schedule(dynamic, 8)
  for (int l = 0; l < L; ++l) {
    const int node = I + l;
    const int a = tp.node_range[2 * node], b = tp.node_range[2 * node + 1], cnt = b - a + 1;
    // Targets in blocks of kLanes (SIMD lanes, accumulators in registers),
    // sources streamed one at a time: self leaf (i != j) then the near leaves.
    T ax[256], ay[256], az[256], ph[256];
    for (int i0 = 0; i0 < cnt; i0 += kLanes) {
      T xi[kLanes], yi[kLanes], zi[kLanes], sx[kLanes], sy[kLanes], sz[kLanes], sp[kLanes];
      int id[kLanes];
      for (int u = 0; u < kLanes; ++u) {
        const int i = std::min(i0 + u, cnt - 1);
        xi[u] = xs[a + i];
        yi[u] = ys[a + i];
        zi[u] = zs[a + i];
        id[u] = a + i0 + u;
        sx[u] = sy[u] = sz[u] = sp[u] = T(0);
      }
      for (int j = a; j <= b; ++j) {
        const T xj = xs[j], yj = ys[j], zj = zs[j], mj = ms[j];
#pragma omp simd
        for (int u = 0; u < kLanes; ++u) {
          const T dx = xi[u] - xj, dy = yi[u] - yj, dz = zi[u] - zj;
          const T r2 = dx * dx + dy * dy + dz * dz + eps2;
          const T inv = id[u] == j ? T(0) : T(1) / std::sqrt(r2);
          const T w = mj * inv, w3 = w * inv * inv;
          sp[u] += w;
          sx[u] += w3 * dx;
          sy[u] += w3 * dy;
          sz[u] += w3 * dz;
        }
      }
      for (int e = tp.near_off[l]; e < tp.near_off[l + 1]; ++e) {
        const int sn = I + tp.near_src[e];
        const int sa = tp.node_range[2 * sn], sb = tp.node_range[2 * sn + 1];
        for (int j = sa; j <= sb; ++j) {
          const T xj = xs[j], yj = ys[j], zj = zs[j], mj = ms[j];
#pragma omp simd
          for (int u = 0; u < kLanes; ++u) {
            const T dx = xi[u] - xj, dy = yi[u] - yj, dz = zi[u] - zj;
            const T r2 = dx * dx + dy * dy + dz * dz + eps2;
            const T inv = T(1) / std::sqrt(r2);
            const T w = mj * inv, w3 = w * inv * inv;
            sp[u] += w;
            sx[u] += w3 * dx;
            sy[u] += w3 * dy;
            sz[u] += w3 * dz;
          }
        }
      }
      for (int u = 0; u < kLanes && i0 + u < cnt; ++u) {
        ax[i0 + u] = sx[u];
        ay[i0 + u] = sy[u];
        az[i0 + u] = sz[u];
        ph[i0 + u] = sp[u];
      }
    }
    const T* F = &Lc[size_t(node) * C];
    for (int i = 0; i < cnt; ++i) {
      Dual<T> dx{cen[3 * node] - xs[a + i], {1, 0, 0}}, dy{cen[3 * node + 1] - ys[a + i], {0, 1, 0}},
          dz{cen[3 * node + 2] - zs[a + i], {0, 0, 1}};
      Dual<T> U[kCMax];
      solid_harmonics<T, Dual<T>>(tb, p, dx, dy, dz, U);
      T phi = T(0), gx = T(0), gy = T(0), gz = T(0);
      for (int c = 0; c < C; ++c) {
        phi += F[c] * U[c].v;
        gx += F[c] * U[c].d[0];
        gy += F[c] * U[c].d[1];
        gz += F[c] * U[c].d[2];
      }
      const int k = tp.perm[a + i];
      acc_out[3 * size_t(k)] = double(-g * ax[i] + -g * gx);
      acc_out[3 * size_t(k) + 1] = double(-g * ay[i] + -g * gy);
      acc_out[3 * size_t(k) + 2] = double(-g * az[i] + -g * gz);
      if (pot_out) pot_out[k] = double(-g * ph[i] + -g * phi);
    }
  }
  t[ti++] = omp_get_wtime();
  if (seconds)
    for (int k = 0; k + 1 < ti; ++k) seconds[k] = t[k + 1] - t[k];
  return 0;
}

}  // namespace

extern "C" {

int cfmm_version(void) { return 1; }

int cfmm_max_threads(void) { return omp_get_max_threads(); }

// Full fixed-order FMM on a given topology.  dtype 0 = float64, 1 = float32
// (arithmetic type; inputs and outputs are float64 arrays).  x [n][3] and
// m [n] in input order; perm [n] tree slot -> input index; node arrays as
// the topology holds them (leaves are nodes I..I+L-1, near_src are leaf
// indices); orders [K] (optional) per far pair order in 0..p; bu: dense B_U
// blocks for degrees 0..p, row-major, concatenated.
// acc [n][3] (and pot [n] unless NULL) in input order.  seconds[6] (optional):
// gather, P2M, M2M, M2L, L2L, L2P+P2P wall times.  threads <= 0: OpenMP default.
int cfmm_evaluate(int dtype, int p, int n, const double* x, const double* m, const int32_t* perm, int I, int L,
                  const int32_t* node_range, const int32_t* left, const int32_t* right, const int32_t* parent,
                  const int32_t* depth, const double* center, const int32_t* far_off, const int32_t* far_src,
                  const int32_t* near_off, const int32_t* near_src, const int8_t* orders, const double* bu,
                  double G, double eps,
                  int threads, double* acc, double* pot, double* seconds) {
  if (p < 0 || p > kPMax || n < 1 || L < 1 || I < 0 || !x || !m || !perm || !node_range || !depth || !center ||
      !far_off || !near_off || !bu || !acc)
    return 1;
  if (I > 0 && (!left || !right || !parent)) return 1;
  if (threads > 0) omp_set_num_threads(threads);
  if (orders)
    for (int e = 0; e < far_off[I + L]; ++e)
      if (orders[e] < 0 || orders[e] > p) return 4;
  Topo tp{n, I, L, I + L, p, perm, node_range, left, right, parent, depth, far_off, far_src, near_off, near_src,
          orders, center};
  if (dtype == 0) return run<double>(tp, bu, x, m, G, eps, acc, pot, seconds);
  if (dtype == 1) return run<float>(tp, bu, x, m, G, eps, acc, pot, seconds);
  return 2;
}

}  // extern "C"
