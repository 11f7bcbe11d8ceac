// Build-time generator: writes the B_U matrices (degrees 0..JFM_MAX_ORDER) as
// float64 literals into a header that the HIP kernels include, so every
// rotation is compiled with its constants folded in and the checkerboard zeros
// of B_U eliminated at compile time.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <vector>

#include "jfm.h"
#include "tables.hpp"

// MFMA schedule of the B_U rotation blocks for v_mfma_f32_4x4x1_16b_f32.
// Every degree block of M = B_U (t = 0) or B_U^T (t = 1) splits into
// connected components (index sets closed under M's nonzeros; four per degree,
// at most five indices).  A component's output rows are cut into groups of
// four; a group takes one instruction per column of the component:
//   out[rows[i]] += A[inst][i] * in[cols[k]],  A[inst][i] = M[rows[i]][cols[k]],
// with inst = inst0 + k numbered per matrix over degrees 1..JFM_MAX_ORDER.
namespace {
struct Group {
  int deg, rows[4], ncols, cols[5], inst0;
};

void emit_mfma_schedule(FILE* f) {
  std::vector<Group> groups[2];
  std::vector<std::array<double, 4>> avals[2];
  int goff[2][JFM_MAX_ORDER + 2], ioff[2][JFM_MAX_ORDER + 2];
  for (int t = 0; t < 2; ++t) {
    goff[t][0] = goff[t][1] = 0;
    ioff[t][0] = ioff[t][1] = 0;
    for (int ell = 1; ell <= JFM_MAX_ORDER; ++ell) {
      const std::vector<double> bu = jfm::dehnen_bu(ell);
      const int w = 2 * ell + 1;
      auto M = [&](int r, int c) {
        const double v = t ? bu[static_cast<size_t>(c) * w + r] : bu[static_cast<size_t>(r) * w + c];
        return std::fabs(v) < 1e-15 ? 0.0 : v;
      };
      // components of the bipartite graph rows <-> cols (union-find over 2w nodes)
      std::vector<int> par(2 * w);
      for (int i = 0; i < 2 * w; ++i) par[i] = i;
      std::function<int(int)> find = [&](int x) { return par[x] == x ? x : par[x] = find(par[x]); };
      for (int r = 0; r < w; ++r)
        for (int c = 0; c < w; ++c)
          if (M(r, c) != 0.0) par[find(r)] = find(w + c);
      std::vector<int> seen;
      for (int r = 0; r < w; ++r) {
        const int root = find(r);
        bool done = false;
        for (int q : seen) done |= q == root;
        if (done) continue;
        seen.push_back(root);
        std::vector<int> rows, cols;
        for (int i = 0; i < w; ++i) {
          if (find(i) == root) rows.push_back(i);
          if (find(w + i) == root) cols.push_back(i);
        }
        if (cols.size() > 5) {
          std::fprintf(stderr, "gen_tables: component of %zu columns at degree %d\n", cols.size(), ell);
          std::exit(3);
        }
        for (size_t g0 = 0; g0 < rows.size(); g0 += 4) {
          Group g{};
          g.deg = ell;
          for (int i = 0; i < 4; ++i) g.rows[i] = g0 + i < rows.size() ? rows[g0 + i] : -1;
          g.ncols = static_cast<int>(cols.size());
          for (int k = 0; k < 5; ++k) g.cols[k] = k < g.ncols ? cols[k] : -1;
          g.inst0 = static_cast<int>(avals[t].size());
          for (int k = 0; k < g.ncols; ++k) {
            std::array<double, 4> a{};
            for (int i = 0; i < 4; ++i) a[i] = g.rows[i] >= 0 ? M(g.rows[i], cols[k]) : 0.0;
            avals[t].push_back(a);
          }
          groups[t].push_back(g);
        }
      }
      goff[t][ell + 1] = static_cast<int>(groups[t].size());
      ioff[t][ell + 1] = static_cast<int>(avals[t].size());
    }
  }
  std::fprintf(f, "// MFMA schedule of B_U (matrix 0) and B_U^T (matrix 1) for v_mfma_f32_4x4x1_16b_f32\n");
  std::fprintf(f, "// (gen_tables.cpp): groups of degree l of matrix t are kBuGroups[t][kBuGroupOff[t][l] ..\n");
  std::fprintf(f, "// kBuGroupOff[t][l+1]); degrees 1..l use instructions [0, kBuInstOff[t][l+1]).\n");
  std::fprintf(f, "struct BuGroup {\n  int rows[4];\n  int ncols;\n  int cols[5];\n  int inst0;\n};\n");
  for (const char* name : {"kBuGroupOff", "kBuInstOff"}) {
    std::fprintf(f, "constexpr int %s[2][%d] = {", name, JFM_MAX_ORDER + 2);
    for (int t = 0; t < 2; ++t) {
      std::fprintf(f, "{");
      for (int l = 0; l < JFM_MAX_ORDER + 2; ++l)
        std::fprintf(f, "%d%s", name[3] == 'G' ? goff[t][l] : ioff[t][l], l + 1 < JFM_MAX_ORDER + 2 ? ", " : "");
      std::fprintf(f, "}%s", t ? "" : ", ");
    }
    std::fprintf(f, "};\n");
  }
  const size_t ng = std::max(groups[0].size(), groups[1].size());
  const size_t ni = std::max(avals[0].size(), avals[1].size());
  std::fprintf(f, "constexpr BuGroup kBuGroups[2][%zu] = {\n", ng);
  for (int t = 0; t < 2; ++t) {
    std::fprintf(f, "  {\n");
    for (size_t q = 0; q < ng; ++q) {
      Group g = q < groups[t].size() ? groups[t][q] : Group{0, {-1, -1, -1, -1}, 0, {-1, -1, -1, -1, -1}, 0};
      std::fprintf(f, "    {{%d, %d, %d, %d}, %d, {%d, %d, %d, %d, %d}, %d},  // degree %d\n", g.rows[0], g.rows[1],
                   g.rows[2], g.rows[3], g.ncols, g.cols[0], g.cols[1], g.cols[2], g.cols[3], g.cols[4], g.inst0, g.deg);
    }
    std::fprintf(f, "  },\n");
  }
  std::fprintf(f, "};\n");
  std::fprintf(f, "constexpr int kBuInstMax = %zu;\n", ni);
  std::fprintf(f, "// A operand of instruction q, rows 0..3 (float32, rounded from the float64 B_U)\n");
  std::fprintf(f, "constexpr float kBuMfmaA[2][%zu][4] = {\n", ni);
  for (int t = 0; t < 2; ++t) {
    std::fprintf(f, "  {\n");
    for (size_t q = 0; q < ni; ++q) {
      std::array<double, 4> a = q < avals[t].size() ? avals[t][q] : std::array<double, 4>{};
      std::fprintf(f, "    {%.9g, %.9g, %.9g, %.9g},\n", static_cast<float>(a[0]), static_cast<float>(a[1]),
                   static_cast<float>(a[2]), static_cast<float>(a[3]));
    }
    std::fprintf(f, "  },\n");
  }
  std::fprintf(f, "};\n\n");
}
}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: gen_tables <out.h>\n");
    return 2;
  }
  FILE* f = std::fopen(argv[1], "w");
  if (!f) return 1;
  std::fprintf(f, "// generated by gen_tables.cpp -- do not edit\n#pragma once\n\nnamespace jfm {\n\n");
  // one degree past JFM_MAX_ORDER: the centre adjoint translates at order p + 1
  std::fprintf(f, "// B_U blocks for degrees 0..%d, row-major, block l at offset l(2l-1)(2l+1)/3\n",
               JFM_MAX_ORDER + 1);
  std::fprintf(f, "constexpr double kBU[] = {\n");
  for (int ell = 0; ell <= JFM_MAX_ORDER + 1; ++ell) {
    std::vector<double> bu = jfm::dehnen_bu(ell);
    std::fprintf(f, "  // degree %d\n", ell);
    for (size_t k = 0; k < bu.size(); ++k) {
      double v = bu[k];
      if (std::fabs(v) < 1e-15) v = 0.0;  // exact zeros of the checkerboard
      std::fprintf(f, "  %.17g,\n", v);
    }
  }
  std::fprintf(f, "};\n\n");
  emit_mfma_schedule(f);
  std::fprintf(f, "}  // namespace jfm\n");
  std::fclose(f);
  return 0;
}
