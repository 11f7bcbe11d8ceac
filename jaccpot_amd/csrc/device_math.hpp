// Device math of the real (Dehnen no-sqrt2) basis, shared by every kernel.
//
// All routines act on register arrays and are fully unrolled for a
// compile-time order P; the B_U swap matrices are folded in as literals
// (tables_gen.h) so the structural zeros of B_U (about half its entries)
// cost nothing.
//
// Conventions (packed index n*n + n + m):
//   U_n^m  = Re R_n^|m| (m >= 0), Im R_n^|m| (m < 0), with the regular solid
//            harmonic R_n^m = r^n P_n^m(cos t) e^{i m phi} / (n+m)!, P without the
//            Condon-Shortley phase (operators/real_p2m_l2p.py:51-232).  Evaluated
//            as polynomials: R_m^m = (x+iy)/(2m) R_{m-1}^{m-1}, R_{m+1}^m = z R_m^m,
//            R_n^m = ((2n-1) z R_{n-1}^m - r^2 R_{n-2}^m) / (n^2 - m^2).
//   Dz(a)  block rotation of (v_{+m}, v_{-m}) by m*a (operators/real_rotations.py:75-147).
//   D_to^M = B_U Dz(-ax) B_U Dz(az),  D_from^M = Dz(-az) B_U Dz(ax) B_U,
//   D_to^L = (D_from^M)^T,            D_from^L = (D_to^M)^T
//            with az = atan2(x, y), ax = atan2(rho, z) (real_rotations.py:340-649).
//   z-M2L  F_n^m = s(m) sum_{k=|m|}^{P-n} (n+k)! / r^{n+k+1} M_k^m,
//            s(m) = (-1)^m (2 if m != 0) (operators/real_translations.py:213-334),
//            evaluated in the scaled form r^-(n+1) sum_k (n+k)! (M_k^m r^-k) so
//            float32 never forms r^-(2P+1).
//   z-shift M2M: out_n^m = sum_{k=0}^{n-|m|} dz^k/k! in_{n-k}^m;
//           L2L: out_n^m = sum_{k=0}^{P-n} dz^k/k! in_{n+k}^m (real_translations.py:72-177).
#pragma once

#include <hip/hip_runtime.h>

#include "tables_gen.h"

namespace jfm {

__host__ __device__ constexpr int ncoef(int p) { return (p + 1) * (p + 1); }
__host__ __device__ constexpr int sh_idx(int n, int m) { return n * n + n + m; }
__host__ __device__ constexpr int bu_off(int l) { return l * (2 * l - 1) * (2 * l + 1) / 3; }
__host__ __device__ constexpr int tri(int n, int m) { return n * (n + 1) / 2 + m; }
__host__ __device__ constexpr int ntri(int p) { return (p + 1) * (p + 2) / 2; }
// Row stride of coefficient arrays: rows padded to 16 bytes.
__host__ __device__ constexpr int coef_stride(int p, int bytes) {
  return bytes == 4 ? ((ncoef(p) + 3) / 4) * 4 : ((ncoef(p) + 1) / 2) * 2;
}

constexpr double kFact[] = {1.0,
                            1.0,
                            2.0,
                            6.0,
                            24.0,
                            120.0,
                            720.0,
                            5040.0,
                            40320.0,
                            362880.0,
                            3628800.0,
                            39916800.0,
                            479001600.0,
                            6227020800.0,
                            87178291200.0,
                            1307674368000.0,
                            20922789888000.0,
                            355687428096000.0};

template <typename T>
__host__ __device__ __forceinline__ T dev_sqrt(T x) {
  return sqrt(x);
}

// Per-pair alignment trig: cos/sin of m*az and m*ax for m = 0..P.
template <typename T, int P>
struct Align {
  T ca[P + 1], sa[P + 1];  // az
  T cx[P + 1], sx[P + 1];  // ax
  T r;                     // |delta|

  __host__ __device__ __forceinline__ void init(T x, T y, T z) {
    const T rho2 = x * x + y * y;
    const T r2 = rho2 + z * z;
    T c1a = T(1), s1a = T(0), c1x = T(1), s1x = T(0);
    T rho = T(0);
    r = T(0);
#if defined(__HIP_DEVICE_COMPILE__)
    // float32 on the GPU: the hardware reciprocal square root (about 1 ulp)
    // replaces the correctly rounded square roots and divisions (about 60
    // VALU instructions per pair) for arguments above the denormal range
    constexpr T kNormMin = T(1.17549435e-38);
    if (sizeof(T) == 4 && rho2 >= kNormMin) {
      const T ri = __builtin_amdgcn_rsqf(static_cast<float>(rho2));
      rho = rho2 * ri;
      c1a = y * ri;  // az = atan2(x, y)
      s1a = x * ri;
      const T rri = __builtin_amdgcn_rsqf(static_cast<float>(r2));
      r = r2 * rri;
      c1x = z * rri;  // ax = atan2(rho, z)
      s1x = rho * rri;
    } else if (sizeof(T) == 4 && r2 >= kNormMin) {  // on the z axis (rho = 0 or denormal)
      if (rho2 > T(0)) {
        rho = dev_sqrt(rho2);
        c1a = y / rho;
        s1a = x / rho;
      }
      const T rri = __builtin_amdgcn_rsqf(static_cast<float>(r2));
      r = r2 * rri;
      c1x = z * rri;
      s1x = rho * rri;
    } else
#endif
    {
      if (rho2 > T(0)) {
        rho = dev_sqrt(rho2);
        c1a = y / rho;  // az = atan2(x, y)
        s1a = x / rho;
      }
      if (r2 > T(0)) {
        r = dev_sqrt(r2);
        c1x = z / r;  // ax = atan2(rho, z)
        s1x = rho / r;
      }
    }
    ca[0] = T(1);
    sa[0] = T(0);
    cx[0] = T(1);
    sx[0] = T(0);
#pragma unroll
    for (int m = 1; m <= P; ++m) {
      ca[m] = c1a * ca[m - 1] - s1a * sa[m - 1];
      sa[m] = s1a * ca[m - 1] + c1a * sa[m - 1];
      cx[m] = c1x * cx[m - 1] - s1x * sx[m - 1];
      sx[m] = s1x * cx[m - 1] + c1x * sx[m - 1];
    }
  }
};

// blk <- Dz(+/-a) blk for one degree; c[m], s[m] = cos/sin(m a).
template <typename T, int L, bool NEG>
__host__ __device__ __forceinline__ void apply_dz(T* blk, const T* c, const T* s) {
#pragma unroll
  for (int m = 1; m <= L; ++m) {
    const T vp = blk[L + m], vm = blk[L - m];
    const T sm = NEG ? -s[m] : s[m];
    blk[L + m] = c[m] * vp - sm * vm;
    blk[L - m] = sm * vp + c[m] * vm;
  }
}

// blk <- B_U blk (TRANS: B_U^T blk) for one degree, literals folded in.
template <typename T, int L, bool TRANS>
__host__ __device__ __forceinline__ void apply_bu(T* blk) {
  constexpr int W = 2 * L + 1;
  constexpr int O = bu_off(L);
  T in[W];
#pragma unroll
  for (int j = 0; j < W; ++j) in[j] = blk[j];
#pragma unroll
  for (int i = 0; i < W; ++i) {
    T acc = T(0);
#pragma unroll
    for (int j = 0; j < W; ++j) {
      const double b = TRANS ? kBU[O + j * W + i] : kBU[O + i * W + j];
      if (b != 0.0) acc = fma(static_cast<T>(b), in[j], acc);
    }
    blk[i] = acc;
  }
}

enum class Rot { ToZMultipole, FromZMultipole, ToZLocal, FromZLocal };

template <typename T, int P, int L, Rot R>
__host__ __device__ __forceinline__ void rotate_block(T* v, const Align<T, P>& a) {
  T* blk = v + L * L;
  if constexpr (R == Rot::ToZMultipole) {  // B Dz(-ax) B Dz(az)
    apply_dz<T, L, false>(blk, a.ca, a.sa);
    apply_bu<T, L, false>(blk);
    apply_dz<T, L, true>(blk, a.cx, a.sx);
    apply_bu<T, L, false>(blk);
  } else if constexpr (R == Rot::FromZMultipole) {  // Dz(-az) B Dz(ax) B
    apply_bu<T, L, false>(blk);
    apply_dz<T, L, false>(blk, a.cx, a.sx);
    apply_bu<T, L, false>(blk);
    apply_dz<T, L, true>(blk, a.ca, a.sa);
  } else if constexpr (R == Rot::ToZLocal) {  // B^T Dz(-ax) B^T Dz(az)
    apply_dz<T, L, false>(blk, a.ca, a.sa);
    apply_bu<T, L, true>(blk);
    apply_dz<T, L, true>(blk, a.cx, a.sx);
    apply_bu<T, L, true>(blk);
  } else {  // Dz(-az) B^T Dz(ax) B^T
    apply_bu<T, L, true>(blk);
    apply_dz<T, L, false>(blk, a.cx, a.sx);
    apply_bu<T, L, true>(blk);
    apply_dz<T, L, true>(blk, a.ca, a.sa);
  }
}

template <typename T, int P, Rot R, int L = 0>
__host__ __device__ __forceinline__ void rotate_all(T* v, const Align<T, P>& a) {
  if constexpr (L == 0) {
    // degree 0 is invariant under every rotation
    rotate_all<T, P, R, 1>(v, a);
  } else if constexpr (L <= P) {
    rotate_block<T, P, L, R>(v, a);
    rotate_all<T, P, R, L + 1>(v, a);
  }
}

// z-axis M2M shift, in place (n descending keeps the inputs intact).
template <typename T, int P>
__host__ __device__ __forceinline__ void zshift_m2m(T* v, T dz) {
  T dk[P + 1];
  dk[0] = T(1);
#pragma unroll
  for (int k = 1; k <= P; ++k) dk[k] = dk[k - 1] * dz * static_cast<T>(1.0 / k);
#pragma unroll
  for (int n = P; n >= 0; --n) {
#pragma unroll
    for (int m = -n; m <= n; ++m) {
      const int am = m < 0 ? -m : m;
      T acc = T(0);
#pragma unroll
      for (int k = 0; k <= n - am; ++k) acc = fma(dk[k], v[sh_idx(n - k, m)], acc);
      v[sh_idx(n, m)] = acc;
    }
  }
}

// z-axis L2L shift, in place (n ascending keeps the inputs intact).
template <typename T, int P>
__host__ __device__ __forceinline__ void zshift_l2l(T* v, T dz) {
  T dk[P + 1];
  dk[0] = T(1);
#pragma unroll
  for (int k = 1; k <= P; ++k) dk[k] = dk[k - 1] * dz * static_cast<T>(1.0 / k);
#pragma unroll
  for (int n = 0; n <= P; ++n) {
#pragma unroll
    for (int m = -n; m <= n; ++m) {
      T acc = T(0);
#pragma unroll
      for (int k = 0; k <= P - n; ++k) acc = fma(dk[k], v[sh_idx(n + k, m)], acc);
      v[sh_idx(n, m)] = acc;
    }
  }
}

// z-axis M2L core: out = Zcore(r) * in (in: z-frame multipole, out: z-frame local).
template <typename T, int P>
__host__ __device__ __forceinline__ void zm2l(const T* in, T* out, T r) {
  const T rr = r > T(1e-30) ? r : T(1e-30);
#if defined(__HIP_DEVICE_COMPILE__)
  // float32 on the GPU: hardware reciprocal (about 1 ulp); rr >= 1e-30 is normal
  const T rinv = sizeof(T) == 4 ? T(__builtin_amdgcn_rcpf(static_cast<float>(rr))) : T(1) / rr;
#else
  const T rinv = T(1) / rr;
#endif
  T rp[P + 2];
  rp[0] = T(1);
#pragma unroll
  for (int k = 1; k <= P + 1; ++k) rp[k] = rp[k - 1] * rinv;
  T sc[ncoef(P)];
#pragma unroll
  for (int k = 0; k <= P; ++k)
#pragma unroll
    for (int m = -k; m <= k; ++m) sc[sh_idx(k, m)] = in[sh_idx(k, m)] * rp[k];
#pragma unroll
  for (int n = 0; n <= P; ++n) {
#pragma unroll
    for (int m = -n; m <= n; ++m) {
      const int am = m < 0 ? -m : m;
      const double sgn = ((am & 1) ? -1.0 : 1.0) * (am ? 2.0 : 1.0);
      T acc = T(0);
#pragma unroll
      for (int k = am; k <= P - n; ++k) acc = fma(static_cast<T>(sgn * kFact[n + k]), sc[sh_idx(k, m)], acc);
      out[sh_idx(n, m)] = acc * rp[n + 1];
    }
  }
}

// Regular solid harmonics: re/im of R_n^m for 0 <= m <= n <= P (triangular index).
template <typename T, int P>
__host__ __device__ __forceinline__ void solid_harmonics(T x, T y, T z, T* re, T* im) {
  const T r2 = x * x + y * y + z * z;
  re[0] = T(1);
  im[0] = T(0);
#pragma unroll
  for (int m = 1; m <= P; ++m) {  // diagonal
    const T f = static_cast<T>(0.5 / m);
    const T pr = re[tri(m - 1, m - 1)], pi = im[tri(m - 1, m - 1)];
    re[tri(m, m)] = f * (x * pr - y * pi);
    im[tri(m, m)] = f * (x * pi + y * pr);
  }
#pragma unroll
  for (int m = 0; m < P; ++m) {  // first off-diagonal
    re[tri(m + 1, m)] = z * re[tri(m, m)];
    im[tri(m + 1, m)] = z * im[tri(m, m)];
  }
#pragma unroll
  for (int m = 0; m <= P; ++m)
#pragma unroll
    for (int n = m + 2; n <= P; ++n) {
      const T a = static_cast<T>((2.0 * n - 1.0) / double((n - m) * (n + m)));
      const T b = static_cast<T>(1.0 / double((n - m) * (n + m)));
      re[tri(n, m)] = a * z * re[tri(n - 1, m)] - b * r2 * re[tri(n - 2, m)];
      im[tri(n, m)] = a * z * im[tri(n - 1, m)] - b * r2 * im[tri(n - 2, m)];
    }
}

// phi = sum F U(d) and grad_d phi, with F packed (stride-free pointer to C values).
template <typename T, int P, typename FLoad>
__host__ __device__ __forceinline__ void l2p_eval(FLoad F, T dx, T dy, T dz, T& phi, T& gx, T& gy, T& gz) {
  T re[ntri(P)], im[ntri(P)];
  solid_harmonics<T, P>(dx, dy, dz, re, im);
  T p = T(0), ax = T(0), ay = T(0), az = T(0);
#pragma unroll
  for (int n = 0; n <= P; ++n) {
    // m = 0
    {
      const T f = F(sh_idx(n, 0));
      p = fma(f, re[tri(n, 0)], p);
      if (n >= 1) {
        az = fma(f, re[tri(n - 1, 0)], az);
        if (n >= 2) {
          ax = fma(-f, re[tri(n - 1, 1)], ax);
          ay = fma(-f, im[tri(n - 1, 1)], ay);
        }
      }
    }
#pragma unroll
    for (int m = 1; m <= n; ++m) {
      const T fc = F(sh_idx(n, m)), fs = F(sh_idx(n, -m));
      p = fma(fc, re[tri(n, m)], fma(fs, im[tri(n, m)], p));
      // A = R_{n-1}^{m-1}, B = R_{n-1}^{m+1}, Zc = R_{n-1}^m
      const T ar = re[tri(n - 1, m - 1)], ai = im[tri(n - 1, m - 1)];
      const T br = (m + 1 <= n - 1) ? re[tri(n - 1, m + 1)] : T(0);
      const T bi = (m + 1 <= n - 1) ? im[tri(n - 1, m + 1)] : T(0);
      const T zr = (m <= n - 1) ? re[tri(n - 1, m)] : T(0);
      const T zi = (m <= n - 1) ? im[tri(n - 1, m)] : T(0);
      // d/dx R = (A - B)/2 ; d/dy R = i (A + B)/2 ; d/dz R = Zc
      ax = fma(fc, T(0.5) * (ar - br), fma(fs, T(0.5) * (ai - bi), ax));
      ay = fma(fc, T(-0.5) * (ai + bi), fma(fs, T(0.5) * (ar + br), ay));
      az = fma(fc, zr, fma(fs, zi, az));
    }
  }
  phi = p;
  gx = ax;
  gy = ay;
  gz = az;
}

}  // namespace jfm
