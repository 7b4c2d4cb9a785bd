// Test infrastructure (oracle). C-callable wrapper around the reference's own
// X-TED_CPU sources, compiled from where they lie (see oracle/Makefile) into
// oracle/_ref/. It converts parent arrays / int label ids into the
// reference's children lists / string labels, calls standard_ted_1
// (X-TED_CPU/TED.cpp:145-174) and copies the returned D_tree out. The
// reference prints its compute-phase timer
// ("Total Time for Parallel Computing = X ms",
// X-TED_CPU/parallel_computing.cpp:184); the wrapper captures that line and
// returns X instead of letting it reach stdout.
#include <cstdint>
#include <cstdio>
#include <sstream>
#include <string>

#include "TED_C++.h"

extern "C" int ref_ted(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                       const int32_t* y_parent, const int32_t* y_label, int32_t n,
                       int32_t threads, int32_t* td, double* compute_ms) {
  if (compute_ms) *compute_ms = 0.0;
  // X-TED_CPU/TED.cpp:181-186
  if (m == 0) return n;
  if (n == 0) return m;
  vector<string> xs(m), ys(n);
  vector<vector<int>> xa(m), ya(n);
  for (int32_t i = 0; i < m; ++i) xs[i] = std::to_string(x_label[i]);
  for (int32_t i = 0; i < n; ++i) ys[i] = std::to_string(y_label[i]);
  for (int32_t i = 1; i < m; ++i) xa[x_parent[i]].push_back(i);
  for (int32_t i = 1; i < n; ++i) ya[y_parent[i]].push_back(i);
  std::ostringstream captured;
  std::streambuf* old = std::cout.rdbuf(captured.rdbuf());
  vector<vector<int>> D = standard_ted_1(xs, xa, ys, ya, threads < 1 ? 1 : threads, 1);
  std::cout.rdbuf(old);
  const std::string text = captured.str();
  const std::string key = "Total Time for Parallel Computing = ";
  size_t at = text.find(key);
  if (compute_ms && at != std::string::npos) *compute_ms = std::atof(text.c_str() + at + key.size());
  if (td)
    for (int32_t i = 0; i < m; ++i)
      for (int32_t j = 0; j < n; ++j) td[(size_t)i * n + j] = D[i][j];
  return D[0][0];
}

// outermost_right_leaves (X-TED_CPU/TED.cpp:21-42) on the children lists of a
// parent array.
extern "C" int ref_orl(const int32_t* parent, int32_t n, int32_t* orl) {
  vector<vector<int>> adj(n);
  for (int32_t i = 1; i < n; ++i) adj[parent[i]].push_back(i);
  vector<int> o = outermost_right_leaves(adj);
  for (int32_t i = 0; i < n; ++i) orl[i] = o[i];
  return 0;
}

// key_roots (X-TED_CPU/TED.cpp:46-77).
extern "C" int ref_key_roots(const int32_t* orl, int32_t n, int32_t* kr, int32_t* K) {
  vector<int> o(orl, orl + n);
  vector<int> k = key_roots(o);
  for (size_t i = 0; i < k.size(); ++i) kr[i] = k[i];
  *K = (int32_t)k.size();
  return 0;
}
