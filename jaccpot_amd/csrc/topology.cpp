// Host construction of the fixed static-radix topology and its interaction
// lists -- the role yggdrax plays for the reference (build_tree,
// compute_tree_mass_moments, compute_tree_geometry and
// build_interactions_and_neighbors, called from jaccpot/runtime/fmm_state.py:630-800,
// upward/real_tree_expansions.py:439-440, upward/tree_geometry.py:65-127 and
// runtime/_interaction_cache.py:523,564).  yggdrax is not vendored, so this is
// this library's own documented spec:
//
//  * 63-bit Morton codes on the cube spanned by the bounds, stable sort by
//    (code, input index) (runtime/fmm_prepare.py:380-383);
//  * L = ceil(N/leaf) consecutive buckets as leaves, nodes = 2L-1
//    (fmm_prepare.py:384-397); internal nodes [0, L-1) root 0, leaf l = node L-1+l;
//  * binary radix tree over the leaf keys (code of the leaf's first particle,
//    leaf index as tie-break), built per node in parallel (Karras 2012);
//  * centres of mass (leaves from their particles, internal nodes from their
//    children, plain mean for a massless node) and the exact max distance from
//    the centre to the node's particles (mutual/device_topology.py:50-160);
//  * dual-tree walk: MAC (r_a + r_b)^2 <= theta^2 d^2 with d > 0
//    (runtime/_interaction_cache.py:1252-1258), split the larger-radius
//    non-leaf node (mutual/topology.py:300-395), both directions emitted, self
//    leaf not listed.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "common.hpp"

namespace jfm {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int32_t fail(int32_t code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace jfm

struct jfm_topology {
  int64_t n = 0;
  int32_t leaf_size = 0;
  int32_t L = 0, I = 0, nodes = 0;
  int32_t max_depth = 0;
  std::vector<int32_t> perm;
  std::vector<int32_t> parent, left, right, range, depth;
  std::vector<double> center, radius, mass;
  std::vector<int32_t> far_off, far_src, near_off, near_src;
};

namespace {

using jfm::fail;

constexpr int kMortonBits = 21;

inline uint64_t spread3(uint64_t v) {
  v &= 0x1FFFFFull;
  uint64_t out = 0;
  for (int b = 0; b < kMortonBits; ++b) out |= ((v >> b) & 1ull) << (3 * b);
  return out;
}

// Common prefix of leaf keys (code, index): code bits, then index bits.
inline int delta_key(const uint64_t* keys, int32_t i, int32_t j) {
  uint64_t a = keys[i], b = keys[j];
  if (a != b) return __builtin_clzll(a ^ b);
  return 64 + __builtin_clz(static_cast<uint32_t>(i ^ j));
}

inline int delta_safe(const uint64_t* keys, int32_t L, int32_t i, int32_t j) {
  if (j < 0 || j >= L) return -1;
  return delta_key(keys, i, j);
}

int32_t find_split(const uint64_t* keys, int32_t first, int32_t last) {
  int common = delta_key(keys, first, last);
  int32_t split = first, step = last - first;
  do {
    step = (step + 1) >> 1;
    int32_t nw = split + step;
    if (nw < last && delta_key(keys, first, nw) > common) split = nw;
  } while (step > 1);
  return split;
}

void parallel_sort_perm(const std::vector<uint64_t>& codes, std::vector<int32_t>& perm, int threads) {
  const int64_t n = static_cast<int64_t>(codes.size());
  perm.resize(n);
  std::iota(perm.begin(), perm.end(), 0);
  auto less = [&codes](int32_t a, int32_t b) {
    return codes[a] < codes[b] || (codes[a] == codes[b] && a < b);
  };
  int t = std::max(1, std::min<int>(threads, static_cast<int>(n / 65536) + 1));
  // round down to a power of two so the merge tree is balanced
  int tp = 1;
  while (tp * 2 <= t) tp *= 2;
  std::vector<int64_t> bounds(tp + 1);
  for (int k = 0; k <= tp; ++k) bounds[k] = n * k / tp;
  {
    std::vector<std::thread> pool;
    for (int k = 0; k < tp; ++k)
      pool.emplace_back([&, k] { std::sort(perm.begin() + bounds[k], perm.begin() + bounds[k + 1], less); });
    for (auto& th : pool) th.join();
  }
  for (int width = 1; width < tp; width *= 2) {
    std::vector<std::thread> pool;
    for (int k = 0; k + width < tp; k += 2 * width) {
      int64_t b = bounds[k], m = bounds[k + width], e = bounds[std::min(k + 2 * width, tp)];
      pool.emplace_back([&, b, m, e] { std::inplace_merge(perm.begin() + b, perm.begin() + m, perm.begin() + e, less); });
    }
    for (auto& th : pool) th.join();
  }
}

struct WalkItem {
  int32_t a, b;  // a == b means "self"
};

}  // namespace

extern "C" {

int32_t jfm_abi_version(void) { return 13; }

const char* jfm_last_error(void) { return jfm::g_last_error.c_str(); }

int32_t jfm_topology_build(const double* positions, const double* masses, int64_t n, int32_t leaf_size,
                           double theta, const double* bounds, int32_t num_threads, jfm_topology** out) {
  if (!out) return fail(JFM_ERR_ARG, "jfm_topology_build: out is NULL");
  *out = nullptr;
  if (!positions || !masses) return fail(JFM_ERR_ARG, "jfm_topology_build: positions/masses is NULL");
  if (n < 1) return fail(JFM_ERR_ARG, "jfm_topology_build: need at least one particle");
  if (n > (int64_t(1) << 30)) return fail(JFM_ERR_ARG, "jfm_topology_build: N exceeds 2^30");
  if (leaf_size < 1) return fail(JFM_ERR_ARG, "jfm_topology_build: leaf_size must be >= 1");
  if (!(theta > 0.0) || !std::isfinite(theta)) return fail(JFM_ERR_ARG, "jfm_topology_build: theta must be > 0");
  const int threads = jfm::resolve_threads(num_threads);
  try {
    auto topo = std::make_unique<jfm_topology>();
    topo->n = n;
    topo->leaf_size = leaf_size;

    // ---- Morton codes ------------------------------------------------------
    double lo[3], hi[3];
    if (bounds) {
      for (int a = 0; a < 3; ++a) {
        lo[a] = bounds[a];
        hi[a] = bounds[3 + a];
      }
    } else {
      for (int a = 0; a < 3; ++a) {
        lo[a] = positions[a];
        hi[a] = positions[a];
      }
      for (int64_t i = 1; i < n; ++i)
        for (int a = 0; a < 3; ++a) {
          double v = positions[3 * i + a];
          lo[a] = std::min(lo[a], v);
          hi[a] = std::max(hi[a], v);
        }
    }
    for (int64_t i = 0; i < 3 * n; ++i)
      if (!std::isfinite(positions[i])) return fail(JFM_ERR_ARG, "jfm_topology_build: non-finite position");
    double extent = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    if (!(extent > 0.0)) extent = 1.0;
    const double scale = static_cast<double>(1 << kMortonBits) / extent;
    const double qmax = static_cast<double>((1 << kMortonBits) - 1);
    std::vector<uint64_t> codes(n);
    jfm::parallel_chunks(n, threads, [&](int64_t b, int64_t e) {
      for (int64_t i = b; i < e; ++i) {
        uint64_t q[3];
        for (int a = 0; a < 3; ++a) {
          double v = std::floor((positions[3 * i + a] - lo[a]) * scale);
          v = std::min(std::max(v, 0.0), qmax);
          q[a] = static_cast<uint64_t>(v);
        }
        codes[i] = (spread3(q[0]) << 2) | (spread3(q[1]) << 1) | spread3(q[2]);
      }
    });
    parallel_sort_perm(codes, topo->perm, threads);

    // ---- leaves and radix tree ---------------------------------------------
    const int32_t L = static_cast<int32_t>((n + leaf_size - 1) / leaf_size);
    const int32_t I = L - 1;
    const int32_t nodes = 2 * L - 1;
    topo->L = L;
    topo->I = I;
    topo->nodes = nodes;
    std::vector<uint64_t> keys(L);
    for (int32_t l = 0; l < L; ++l) keys[l] = codes[topo->perm[static_cast<int64_t>(l) * leaf_size]];
    topo->parent.assign(nodes, -1);
    topo->left.assign(std::max(I, 0), 0);
    topo->right.assign(std::max(I, 0), 0);
    std::vector<int32_t> lrange(2 * static_cast<size_t>(nodes));
    for (int32_t l = 0; l < L; ++l) {
      lrange[2 * (I + l)] = l;
      lrange[2 * (I + l) + 1] = l;
    }
    const uint64_t* kp = keys.data();
    jfm::parallel_chunks(I, threads, [&](int64_t b, int64_t e) {
      for (int64_t ii = b; ii < e; ++ii) {
        const int32_t i = static_cast<int32_t>(ii);
        int32_t first, last;
        if (i == 0) {
          first = 0;
          last = L - 1;
        } else {
          int d = (delta_safe(kp, L, i, i + 1) - delta_safe(kp, L, i, i - 1)) > 0 ? 1 : -1;
          int dmin = delta_safe(kp, L, i, i - d);
          int32_t lmax = 2;
          while (delta_safe(kp, L, i, i + lmax * d) > dmin) lmax *= 2;
          int32_t len = 0;
          for (int32_t t = lmax / 2; t >= 1; t /= 2)
            if (delta_safe(kp, L, i, i + (len + t) * d) > dmin) len += t;
          int32_t j = i + len * d;
          first = std::min(i, j);
          last = std::max(i, j);
        }
        int32_t s = find_split(kp, first, last);
        int32_t lc = (s == first) ? I + s : s;
        int32_t rc = (s + 1 == last) ? I + s + 1 : s + 1;
        topo->left[i] = lc;
        topo->right[i] = rc;
        topo->parent[lc] = i;
        topo->parent[rc] = i;
        lrange[2 * i] = first;
        lrange[2 * i + 1] = last;
      }
    });
    topo->range.resize(2 * static_cast<size_t>(nodes));
    for (int32_t k = 0; k < nodes; ++k) {
      int64_t a = static_cast<int64_t>(lrange[2 * k]) * leaf_size;
      int64_t b = std::min<int64_t>((static_cast<int64_t>(lrange[2 * k + 1]) + 1) * leaf_size, n) - 1;
      topo->range[2 * k] = static_cast<int32_t>(a);
      topo->range[2 * k + 1] = static_cast<int32_t>(b);
    }

    // ---- depths (breadth-first from the root) ------------------------------
    topo->depth.assign(nodes, 0);
    std::vector<std::vector<int32_t>> levels;
    levels.push_back({0});
    while (true) {
      std::vector<int32_t> next;
      for (int32_t k : levels.back())
        if (k < I) {
          for (int32_t c : {topo->left[k], topo->right[k]}) {
            topo->depth[c] = topo->depth[k] + 1;
            next.push_back(c);
          }
        }
      if (next.empty()) break;
      std::sort(next.begin(), next.end());
      levels.push_back(std::move(next));
    }
    topo->max_depth = static_cast<int32_t>(levels.size()) - 1;

    // ---- geometry ------------------------------------------------------------
    topo->mass.assign(nodes, 0.0);
    topo->center.assign(3 * static_cast<size_t>(nodes), 0.0);
    topo->radius.assign(nodes, 0.0);
    auto xs = [&](int64_t slot, int a) { return positions[3 * static_cast<int64_t>(topo->perm[slot]) + a]; };
    auto ms = [&](int64_t slot) { return masses[topo->perm[slot]]; };
    auto mean_of = [&](int32_t node, double* c) {
      int64_t a = topo->range[2 * node], b = topo->range[2 * node + 1];
      double s[3] = {0, 0, 0};
      for (int64_t i = a; i <= b; ++i)
        for (int k = 0; k < 3; ++k) s[k] += xs(i, k);
      for (int k = 0; k < 3; ++k) c[k] = s[k] / static_cast<double>(b - a + 1);
    };
    jfm::parallel_chunks(L, threads, [&](int64_t b0, int64_t e0) {
      for (int64_t l = b0; l < e0; ++l) {
        int32_t node = I + static_cast<int32_t>(l);
        int64_t a = topo->range[2 * node], b = topo->range[2 * node + 1];
        double tot = 0.0, s[3] = {0, 0, 0};
        for (int64_t i = a; i <= b; ++i) tot += ms(i);
        for (int64_t i = a; i <= b; ++i)
          for (int k = 0; k < 3; ++k) s[k] += ms(i) * xs(i, k);
        topo->mass[node] = tot;
        if (tot > 0.0) {
          for (int k = 0; k < 3; ++k) topo->center[3 * node + k] = s[k] / tot;
        } else {
          mean_of(node, &topo->center[3 * node]);
        }
      }
    });
    for (int lev = static_cast<int>(levels.size()) - 1; lev >= 0; --lev) {
      const auto& lv = levels[lev];
      jfm::parallel_chunks(static_cast<int64_t>(lv.size()), threads, [&](int64_t b0, int64_t e0) {
        for (int64_t q = b0; q < e0; ++q) {
          int32_t node = lv[q];
          if (node >= I) continue;
          int32_t lc = topo->left[node], rc = topo->right[node];
          double tot = topo->mass[lc] + topo->mass[rc];
          topo->mass[node] = tot;
          if (tot > 0.0) {
            for (int k = 0; k < 3; ++k)
              topo->center[3 * node + k] =
                  (topo->mass[lc] * topo->center[3 * lc + k] + topo->mass[rc] * topo->center[3 * rc + k]) / tot;
          } else {
            mean_of(node, &topo->center[3 * node]);
          }
        }
      });
    }
    jfm::parallel_dynamic(nodes, threads, 64, [&](int64_t node) {
      int64_t a = topo->range[2 * node], b = topo->range[2 * node + 1];
      const double* c = &topo->center[3 * node];
      double best = 0.0;
      for (int64_t i = a; i <= b; ++i) {
        double dx = xs(i, 0) - c[0], dy = xs(i, 1) - c[1], dz = xs(i, 2) - c[2];
        best = std::max(best, dx * dx + dy * dy + dz * dz);
      }
      topo->radius[node] = std::sqrt(best);
    });

    // ---- dual-tree walk -----------------------------------------------------
    const double th2 = theta * theta;
    const double* C = topo->center.data();
    const double* R = topo->radius.data();
    auto mac = [&](int32_t a, int32_t b) {
      double dx = C[3 * a] - C[3 * b], dy = C[3 * a + 1] - C[3 * b + 1], dz = C[3 * a + 2] - C[3 * b + 2];
      double d2 = dx * dx + dy * dy + dz * dz;
      double s = R[a] + R[b];
      return d2 > 0.0 && s * s <= th2 * d2;
    };
    // Expands one item; appends children to `stack`, emits pairs.
    auto expand = [&](const WalkItem& it, std::vector<WalkItem>& stack, std::vector<uint64_t>& far,
                      std::vector<uint64_t>& near) {
      const int32_t a = it.a, b = it.b;
      if (a == b) {
        if (a >= I) return;
        int32_t la = topo->left[a], ra = topo->right[a];
        stack.push_back({la, ra});
        stack.push_back({ra, ra});
        stack.push_back({la, la});
        return;
      }
      if (mac(a, b)) {
        far.push_back((static_cast<uint64_t>(a) << 32) | static_cast<uint32_t>(b));
        far.push_back((static_cast<uint64_t>(b) << 32) | static_cast<uint32_t>(a));
        return;
      }
      bool al = a >= I, bl = b >= I;
      if (al && bl) {
        uint64_t ta = static_cast<uint32_t>(a - I), tb = static_cast<uint32_t>(b - I);
        near.push_back((ta << 32) | tb);
        near.push_back((tb << 32) | ta);
        return;
      }
      bool split_a = bl || (!al && R[a] >= R[b]);
      if (split_a) {
        stack.push_back({topo->right[a], b});
        stack.push_back({topo->left[a], b});
      } else {
        stack.push_back({a, topo->right[b]});
        stack.push_back({a, topo->left[b]});
      }
    };
    std::vector<uint64_t> far_all, near_all;
    {
      // breadth-first frontier so the parallel work is balanced
      std::vector<WalkItem> frontier{{0, 0}};
      std::vector<uint64_t> far0, near0;
      const size_t want = static_cast<size_t>(threads) * 256;
      while (!frontier.empty() && frontier.size() < want) {
        std::vector<WalkItem> nxt;
        for (const auto& it : frontier) expand(it, nxt, far0, near0);
        frontier.swap(nxt);
      }
      std::vector<std::vector<uint64_t>> far_t(threads), near_t(threads);
      std::atomic<int64_t> next{0};
      const int64_t nf = static_cast<int64_t>(frontier.size());
      auto worker = [&](int tid) {
        std::vector<WalkItem> stack;
        for (;;) {
          int64_t q = next.fetch_add(1);
          if (q >= nf) break;
          stack.push_back(frontier[q]);
          while (!stack.empty()) {
            WalkItem it = stack.back();
            stack.pop_back();
            expand(it, stack, far_t[tid], near_t[tid]);
          }
        }
      };
      std::vector<std::thread> pool;
      for (int k = 0; k < threads; ++k) pool.emplace_back(worker, k);
      for (auto& th : pool) th.join();
      far_all = std::move(far0);
      near_all = std::move(near0);
      for (int k = 0; k < threads; ++k) {
        far_all.insert(far_all.end(), far_t[k].begin(), far_t[k].end());
        near_all.insert(near_all.end(), near_t[k].begin(), near_t[k].end());
      }
    }
    // CSR by target with ascending sources: sorting the packed (tgt << 32 | src) keys.
    auto to_csr = [&](std::vector<uint64_t>& pairs, int32_t rows, std::vector<int32_t>& off,
                      std::vector<int32_t>& src) {
      off.assign(static_cast<size_t>(rows) + 1, 0);
      for (uint64_t v : pairs) off[(v >> 32) + 1]++;
      for (int32_t r = 0; r < rows; ++r) off[r + 1] += off[r];
      src.resize(pairs.size());
      std::vector<int32_t> cur(off.begin(), off.end() - 1);
      for (uint64_t v : pairs) src[cur[v >> 32]++] = static_cast<int32_t>(v & 0xFFFFFFFFull);
      jfm::parallel_dynamic(rows, threads, 256, [&](int64_t r) {
        std::sort(src.begin() + off[r], src.begin() + off[r + 1]);
      });
    };
    to_csr(far_all, nodes, topo->far_off, topo->far_src);
    to_csr(near_all, L, topo->near_off, topo->near_src);
    *out = topo.release();
    return JFM_OK;
  } catch (const std::bad_alloc&) {
    return fail(JFM_ERR_ALLOC, "jfm_topology_build: out of host memory");
  } catch (const std::exception& e) {
    return fail(JFM_ERR_ARG, std::string("jfm_topology_build: ") + e.what());
  }
}

int32_t jfm_topology_info(const jfm_topology* t, int64_t* info) {
  if (!t || !info) return fail(JFM_ERR_ARG, "jfm_topology_info: NULL argument");
  int64_t targets = 0;
  for (int32_t k = 0; k < t->nodes; ++k) targets += (t->far_off[k + 1] > t->far_off[k]) ? 1 : 0;
  info[0] = t->n;
  info[1] = t->leaf_size;
  info[2] = t->L;
  info[3] = t->I;
  info[4] = t->nodes;
  info[5] = static_cast<int64_t>(t->far_src.size());
  info[6] = static_cast<int64_t>(t->near_src.size());
  info[7] = t->max_depth;
  info[8] = targets;
  return JFM_OK;
}

int32_t jfm_topology_export(const jfm_topology* t, int32_t* perm, int32_t* parent, int32_t* left, int32_t* right,
                            int32_t* node_range, int32_t* depth, double* center, double* radius, double* mass,
                            int32_t* far_off, int32_t* far_src, int32_t* near_off, int32_t* near_src) {
  if (!t) return fail(JFM_ERR_ARG, "jfm_topology_export: NULL topology");
  auto cp = [](auto* dst, const auto& v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
  };
  cp(perm, t->perm);
  cp(parent, t->parent);
  cp(left, t->left);
  cp(right, t->right);
  cp(node_range, t->range);
  cp(depth, t->depth);
  cp(center, t->center);
  cp(radius, t->radius);
  cp(mass, t->mass);
  cp(far_off, t->far_off);
  cp(far_src, t->far_src);
  cp(near_off, t->near_off);
  cp(near_src, t->near_src);
  return JFM_OK;
}

int32_t jfm_topology_schedules(const jfm_topology* t, int32_t* m2m_nodes, int32_t* m2m_off, int32_t* l2l_nodes,
                               int32_t* l2l_off, int32_t* m2l_targets) {
  if (!t) return fail(JFM_ERR_ARG, "jfm_topology_schedules: NULL topology");
  const int32_t D = t->max_depth;
  std::vector<std::vector<int32_t>> by_depth(D + 1);
  for (int32_t k = 0; k < t->nodes; ++k) by_depth[t->depth[k]].push_back(k);
  if (m2m_nodes && m2m_off) {
    int32_t pos = 0, lev = 0;
    for (int32_t d = D; d >= 0; --d) {
      m2m_off[lev++] = pos;
      for (int32_t k : by_depth[d])
        if (k < t->I) m2m_nodes[pos++] = k;
    }
    m2m_off[lev] = pos;
  }
  if (l2l_nodes && l2l_off) {
    int32_t pos = 0, lev = 0;
    for (int32_t d = 1; d <= D; ++d) {
      l2l_off[lev++] = pos;
      for (int32_t k : by_depth[d]) l2l_nodes[pos++] = k;
    }
    l2l_off[lev] = pos;
    // pad the unused slot so the array has D+1 offsets even when D == 0
    if (D == 0) l2l_off[0] = 0;
  }
  if (m2l_targets) {
    int32_t pos = 0;
    for (int32_t k = 0; k < t->nodes; ++k)
      if (t->far_off[k + 1] > t->far_off[k]) m2l_targets[pos++] = k;
  }
  return JFM_OK;
}

void jfm_topology_free(jfm_topology* t) { delete t; }

}  // extern "C"
