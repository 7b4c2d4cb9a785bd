// Test infrastructure (oracle). CPU restatement of X-TED_CPU's parallel
// Zhang-Shasha: outermost-right leaves, keyroots, keyroot depth, the per-table
// forest recurrence and the level-synchronous thread schedule. Only tests, the
// smoke check and the bench's CPU-baseline leg use it.
//
// Parity pins (see tests/test_oracle.py):
//   oracle_orl            closed form orl[i] = i + |subtree(i)| - 1; the
//                         compiled reference (oracle/_ref) on dataset trees.
//   oracle_keyroots       closed form {0} U {not-last-child}, descending;
//                         the compiled reference on dataset trees.
//   oracle_keyroot_depth  keyroot nesting height computed independently
//                         (tree recursion) on random trees.
//   oracle_levels         invariant: every TD cell a table reads is written at
//                         a strictly lower level.
//   oracle_ted            golden fixtures written by the compiled reference,
//                         the dataset known answers, the textbook oracle
//                         (zs_textbook.c), brute force over edit mappings on
//                         trees of <= 6 nodes, Levenshtein closed forms for
//                         path trees and stars, metric properties.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <thread>
#include <vector>

#include "oracle.h"

namespace {

std::vector<std::vector<int>> children_of(const int32_t* parent, int32_t n) {
  std::vector<std::vector<int>> adj(n);
  for (int32_t i = 1; i < n; ++i) adj[parent[i]].push_back(i);
  return adj;
}

bool valid_parent(const int32_t* parent, int32_t n) {
  if (n == 0) return true;
  if (parent[0] != -1) return false;
  for (int32_t i = 1; i < n; ++i)
    if (parent[i] < 0 || parent[i] >= i) return false;
  return true;
}

// X-TED_CPU/TED.cpp:21-42: walk the last-child chain, memoising.
std::vector<int> orl_of(const std::vector<std::vector<int>>& adj) {
  const int m = (int)adj.size();
  std::vector<int> orl(m, -1);
  for (int i = 0; i < m; ++i) {
    int r = i;
    for (;;) {
      if (adj[r].empty()) {
        orl[i] = r;
        break;
      }
      if (orl[r] >= 0) {
        orl[i] = orl[r];
        break;
      }
      r = adj[r].back();
    }
  }
  return orl;
}

// X-TED_CPU/TED.cpp:46-77: first node per distinct orl, descending order.
std::vector<int> keyroots_of(const std::vector<int>& orl) {
  const int m = (int)orl.size();
  std::vector<int> first(m, -1);
  for (int i = 0; i < m; ++i)
    if (first[orl[i]] < 0) first[orl[i]] = i;
  std::vector<int> kr;
  for (int r = 0; r < m; ++r)
    if (first[r] >= 0) kr.push_back(first[r]);
  std::sort(kr.begin(), kr.end(), [](int a, int b) { return a > b; });
  return kr;
}

// X-TED_CPU/parallel_computing.cpp:93-119: O(K^2) nesting depth, keyroots
// taken in descending preorder so nested keyroots come first.
std::vector<int> depth_of(const std::vector<int>& orl, const std::vector<int>& kr) {
  const int K = (int)kr.size();
  std::vector<int> d(K, 0);
  for (int i = 0; i < K; ++i) {
    const int node = kr[i];
    if (node == orl[node] || node == orl[node] - 1) continue;
    for (int j = 0; j < i; ++j) {
      const int node2 = kr[j];
      if (node <= node2 && orl[node] >= node2) d[i] = std::max(d[i], d[j] + 1);
    }
  }
  return d;
}

struct Problem {
  int m, n;
  const int32_t* xl;
  const int32_t* yl;
  std::vector<int> xo, yo, xk, yk;
  int32_t* td;  // m*n
  const oracle_costs* costs;  // NULL: unit costs
  int del(int i) const { return costs ? costs->del[xl[i]] : 1; }
  int ins(int j) const { return costs ? costs->ins[yl[j]] : 1; }
  int ren(int i, int j) const {
    return costs ? costs->ren[(size_t)xl[i] * costs->alphabet + yl[j]] : (xl[i] != yl[j] ? 1 : 0);
  }
};

// X-TED_CPU/parallel_computing.cpp:4-45: one keyroot-pair table, forest F is
// an (m+1) x (n+1) scratch addressed by absolute (i, j).
void compute_table(const Problem& P, int k, int l, int32_t* F) {
  const int W = P.n + 1;
  const int i0 = P.xk[k], j0 = P.yk[l];
  const int imax = P.xo[i0] + 1, jmax = P.yo[j0] + 1;
  F[imax * W + jmax] = 0;
  for (int i = imax - 1; i >= i0; --i) F[i * W + jmax] = P.del(i) + F[(i + 1) * W + jmax];
  for (int j = jmax - 1; j >= j0; --j) F[imax * W + j] = P.ins(j) + F[imax * W + j + 1];
  for (int i = imax - 1; i >= i0; --i) {
    for (int j = jmax - 1; j >= j0; --j) {
      int v;
      const int dv = P.del(i) + F[(i + 1) * W + j], iv = P.ins(j) + F[i * W + j + 1];
      if (P.xo[i] == P.xo[i0] && P.yo[j] == P.yo[j0]) {
        v = std::min({P.ren(i, j) + F[(i + 1) * W + j + 1], dv, iv});
        P.td[(size_t)i * P.n + j] = v;
      } else {
        v = std::min({P.td[(size_t)i * P.n + j] + F[(P.xo[i] + 1) * W + P.yo[j] + 1], dv, iv});
      }
      F[i * W + j] = v;
    }
  }
}

}  // namespace

extern "C" {

int oracle_orl(const int32_t* parent, int32_t n, int32_t* orl) {
  if (!valid_parent(parent, n)) return -1;
  auto o = orl_of(children_of(parent, n));
  std::copy(o.begin(), o.end(), orl);
  return 0;
}

int oracle_keyroots(const int32_t* orl, int32_t n, int32_t* kr, int32_t* K) {
  std::vector<int> o(orl, orl + n);
  auto k = keyroots_of(o);
  std::copy(k.begin(), k.end(), kr);
  *K = (int32_t)k.size();
  return 0;
}

int oracle_keyroot_depth(const int32_t* orl, const int32_t* kr, int32_t K, int32_t* depth) {
  int32_t mx = 0;
  for (int32_t i = 0; i < K; ++i) mx = std::max(mx, orl[kr[i]]);
  std::vector<int> o(orl, orl + mx + 1), k(kr, kr + K);
  auto d = depth_of(o, k);
  std::copy(d.begin(), d.end(), depth);
  return 0;
}

int oracle_levels(const int32_t* x_parent, int32_t m, const int32_t* y_parent, int32_t n,
                  int32_t* levels, int32_t* K, int32_t* L) {
  if (!valid_parent(x_parent, m) || !valid_parent(y_parent, n)) return -1;
  auto xo = orl_of(children_of(x_parent, m)), yo = orl_of(children_of(y_parent, n));
  auto xk = keyroots_of(xo), yk = keyroots_of(yo);
  auto dx = depth_of(xo, xk), dy = depth_of(yo, yk);
  // X-TED_CPU/parallel_computing.cpp:121-123
  for (size_t a = 0; a < xk.size(); ++a)
    for (size_t b = 0; b < yk.size(); ++b) levels[a * yk.size() + b] = dx[a] + dy[b];
  *K = (int32_t)xk.size();
  *L = (int32_t)yk.size();
  return 0;
}

// X-TED_CPU/TED.cpp:145-190 and X-TED_CPU/parallel_computing.cpp:66-187.
int oracle_ted(const int32_t* x_parent, const int32_t* x_label, int32_t m,
               const int32_t* y_parent, const int32_t* y_label, int32_t n, int32_t threads,
               int32_t* td, double* compute_ms) {
  return oracle_ted_costs(x_parent, x_label, m, y_parent, y_label, n, nullptr, threads, td,
                          compute_ms);
}

int oracle_ted_costs(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                     const int32_t* y_parent, const int32_t* y_label, int32_t n,
                     const oracle_costs* costs, int32_t threads, int32_t* td,
                     double* compute_ms) {
  if (compute_ms) *compute_ms = 0.0;
  if (m < 0 || n < 0) return -1;
  // X-TED_CPU/TED.cpp:181-186 (with costs: every node of the other tree
  // inserted, or deleted)
  if (m == 0 || n == 0) {
    int s = 0;
    for (int j = 0; j < n; ++j) s += costs ? costs->ins[y_label[j]] : 1;
    for (int i = 0; i < m; ++i) s += costs ? costs->del[x_label[i]] : 1;
    return s;
  }
  if (!valid_parent(x_parent, m) || !valid_parent(y_parent, n)) return -1;
  if (threads < 1) threads = 1;
  Problem P;
  P.costs = costs;
  P.m = m;
  P.n = n;
  P.xl = x_label;
  P.yl = y_label;
  P.xo = orl_of(children_of(x_parent, m));
  P.yo = orl_of(children_of(y_parent, n));
  P.xk = keyroots_of(P.xo);
  P.yk = keyroots_of(P.yo);
  std::vector<int32_t> own;
  if (!td) {
    own.resize((size_t)m * n);
    td = own.data();
  }
  P.td = td;
  std::fill(td, td + (size_t)m * n, -1);
  const int K = (int)P.xk.size(), L = (int)P.yk.size();
  auto dx = depth_of(P.xo, P.xk), dy = depth_of(P.yo, P.yk);
  // One forest scratch per thread (parallel_computing.cpp:72-76).
  std::vector<std::vector<int32_t>> F(threads, std::vector<int32_t>((size_t)(m + 1) * (n + 1)));
  // Level lists (parallel_computing.cpp:121-150), built once per level.
  int maxd = 0;
  for (int a : dx) maxd = std::max(maxd, a + *std::max_element(dy.begin(), dy.end()));
  std::vector<std::vector<int>> level(maxd + 1);
  for (int k = 0; k < K; ++k)
    for (int l = 0; l < L; ++l) level[dx[k] + dy[l]].push_back(k * L + l);
  auto t0 = std::chrono::steady_clock::now();
  // Level loop: threads stride through the level's list, join per level
  // (parallel_computing.cpp:155-182 with task() at :56-64).
  for (int d = 0; d <= maxd; ++d) {
    const auto& list = level[d];
    if (list.empty()) continue;
    if (threads == 1) {
      for (int task : list) compute_table(P, task / L, task % L, F[0].data());
      continue;
    }
    std::vector<std::thread> pool;
    for (int th = 0; th < threads; ++th) {
      pool.emplace_back([&, th]() {
        for (size_t i = th; i < list.size(); i += threads)
          compute_table(P, list[i] / L, list[i] % L, F[th].data());
      });
    }
    for (auto& t : pool) t.join();
  }
  auto t1 = std::chrono::steady_clock::now();
  if (compute_ms) *compute_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  return td[0];
}

}  // extern "C"
