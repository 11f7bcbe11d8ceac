// Float64 host generation of the pair-independent constants of the real
// (Dehnen no-sqrt2) basis: the multipole swap matrices B_U for every degree.
//
// Algorithm (jaccpot/operators/real_dehnen_q.py:49-145, 146-204, 343-412):
//   complex B_l by the Dehnen (2014) eq. 63 recursion from B_0 = [[1]]
//   (interior rows 63a, top edge 63b, bottom edge 63c), then
//   B_U = Re(Q B^T Q^-1) with Q the no-sqrt2 complex -> real map
//   T_{+m} = (Th_m + (-1)^m Th_{-m})/2, T_{-m} = -i (Th_m - (-1)^m Th_{-m})/2,
//   whose inverse is Th_m = T_{+m} + i T_{-m}, Th_{-m} = (-1)^m (T_{+m} - i T_{-m}).
// Used at build time to bake the matrices into the kernels as literals and at
// run time by jfm_tables_bu().
#pragma once

#include <complex>
#include <vector>

namespace jfm {

// Complex B_l, (2l+1)^2 row-major, index [m + l][k + l].
inline std::vector<double> dehnen_b_complex(int ell) {
  std::vector<double> B(1, 1.0);
  for (int n = 0; n < ell; ++n) {
    const int nn = n + 1, w = 2 * nn + 1, wp = 2 * n + 1;
    std::vector<double> Bn(static_cast<size_t>(w) * w, 0.0);
    auto get = [&](int m, int k) -> double {
      if (m < -n || m > n || k < -n || k > n) return 0.0;
      return B[static_cast<size_t>(m + n) * wp + (k + n)];
    };
    for (int mt = -nn; mt <= nn; ++mt)
      for (int lt = -nn; lt <= nn; ++lt) {
        double v;
        if (mt >= -n && mt <= n)
          v = 0.5 * (get(mt, lt - 1) - get(mt, lt + 1));
        else if (mt == n + 1)
          v = 0.5 * (get(n, lt - 1) + get(n, lt + 1) + 2.0 * get(n, lt));
        else
          v = 0.5 * (get(-n, lt - 1) + get(-n, lt + 1) - 2.0 * get(-n, lt));
        Bn[static_cast<size_t>(mt + nn) * w + (lt + nn)] = v;
      }
    B.swap(Bn);
  }
  return B;
}

// Real multipole swap matrix B_U for degree l, (2l+1)^2 row-major.
inline std::vector<double> dehnen_bu(int ell) {
  using cd = std::complex<double>;
  const int w = 2 * ell + 1;
  std::vector<double> Bc = dehnen_b_complex(ell);
  std::vector<cd> Q(static_cast<size_t>(w) * w, cd(0, 0)), Qi(static_cast<size_t>(w) * w, cd(0, 0));
  auto at = [w](std::vector<cd>& M, int r, int c) -> cd& { return M[static_cast<size_t>(r) * w + c]; };
  at(Q, ell, ell) = 1.0;
  at(Qi, ell, ell) = 1.0;
  for (int m = 1; m <= ell; ++m) {
    const double ph = (m % 2) ? -1.0 : 1.0;
    at(Q, ell + m, ell + m) = 0.5;
    at(Q, ell + m, ell - m) = 0.5 * ph;
    at(Q, ell - m, ell + m) = cd(0, -0.5);
    at(Q, ell - m, ell - m) = cd(0, 0.5 * ph);
    // inverse: columns are real rows (T_{+m}, T_{-m}), rows are complex (Th_m, Th_{-m})
    at(Qi, ell + m, ell + m) = 1.0;
    at(Qi, ell + m, ell - m) = cd(0, 1.0);
    at(Qi, ell - m, ell + m) = ph;
    at(Qi, ell - m, ell - m) = cd(0, -ph);
  }
  // B_U = Q B^T Qi
  std::vector<cd> tmp(static_cast<size_t>(w) * w, cd(0, 0));
  for (int r = 0; r < w; ++r)
    for (int c = 0; c < w; ++c) {
      cd s(0, 0);
      for (int k = 0; k < w; ++k) s += at(Q, r, k) * Bc[static_cast<size_t>(c) * w + k];  // B^T[k][c] = B[c][k]
      at(tmp, r, c) = s;
    }
  std::vector<double> out(static_cast<size_t>(w) * w, 0.0);
  for (int r = 0; r < w; ++r)
    for (int c = 0; c < w; ++c) {
      cd s(0, 0);
      for (int k = 0; k < w; ++k) s += at(tmp, r, k) * at(Qi, k, c);
      out[static_cast<size_t>(r) * w + c] = s.real();
    }
  return out;
}

}  // namespace jfm
