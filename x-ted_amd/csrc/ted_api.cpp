// Reference-signature C++ API (include/TED_C++.h) over the xted C ABI.
#include "../../include/TED_C++.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <unordered_map>

#include "../../include/xted.h"
#include "trees.hpp"

float total_milliseconds = 0;

namespace {

// Process-wide context on device XTED_DEVICE (default 0), created on first use.
xted_context* default_context() {
  static std::once_flag once;
  static xted_context* ctx = nullptr;
  std::call_once(once, []() {
    const char* dev = std::getenv("XTED_DEVICE");
    (void)xted_create(dev ? std::atoi(dev) : 0, &ctx);
  });
  if (!ctx)
    throw std::runtime_error(std::string("xted: no device context: ") + xted_last_error());
  return ctx;
}

bool quiet() {
  const char* q = std::getenv("XTED_QUIET");
  return q && q[0] && q[0] != '0';
}

std::vector<int32_t> parents_of(const vector<vector<int>>& adj) {
  std::vector<int32_t> parent;
  if (xted::adj_to_parent(adj, parent) != xted::OK)
    throw std::invalid_argument("xted: adjacency lists are not a preorder-numbered tree");
  return parent;
}

void check(int st) {
  if (st != XTED_OK)
    throw std::runtime_error(std::string("xted: ") + xted_status_string(st) + ": " +
                             xted_last_error());
}

// Runs one pair and prints the reference's timing lines.
vector<vector<int>> run_pair(const std::vector<int32_t>& xp, const std::vector<int32_t>& xl,
                             const std::vector<int32_t>& yp, const std::vector<int32_t>& yl) {
  const int32_t m = (int32_t)xp.size(), n = (int32_t)yp.size();
  std::vector<int32_t> td((size_t)m * n);
  int32_t dist = 0;
  xted_stats st{};
  check(xted_ted(default_context(), xp.data(), xl.data(), m, yp.data(), yl.data(), n, td.data(),
                 &dist, &st));
  total_milliseconds = (float)st.dp_ms;
  if (!quiet()) {
    // X-TED_GPU/parallel_X_TED.cu:224 and :392
    std::printf("Preprocessing Method, the time is %g ms consumed\n", st.plan_ms);
    std::printf("Measured time for total parallel execution = %.3fms\n", st.dp_ms);
  }
  vector<vector<int>> out(m, vector<int>(n));
  for (int32_t i = 0; i < m; ++i)
    for (int32_t j = 0; j < n; ++j) out[i][j] = td[(size_t)i * n + j];
  return out;
}

}  // namespace

// X-TED_GPU/TED.cu:5-19
int min3(int a, int b, int c) {
  int m = a < b ? a : b;
  return m < c ? m : c;
}

// X-TED_GPU/TED.cu:21-42 (computed on the device)
vector<int> outermost_right_leaves(vector<vector<int>>& adj) {
  const int32_t n = (int32_t)adj.size();
  if (n == 0) return {};
  std::vector<int32_t> parent = parents_of(adj), orl(n);
  check(xted_tree_structure(default_context(), parent.data(), n, orl.data(), nullptr, nullptr,
                            nullptr));
  return vector<int>(orl.begin(), orl.end());
}

// X-TED_GPU/TED.cu:44-75: the smallest node of each distinct orl value, in
// descending order. A function of orl alone, so it needs no tree upload.
vector<int> key_roots(vector<int>& orl) {
  const int m = (int)orl.size();
  std::vector<int> first(m, -1);
  for (int i = 0; i < m; ++i) {
    if (orl[i] < 0 || orl[i] >= m) throw std::invalid_argument("xted: orl value out of range");
    if (first[orl[i]] < 0) first[orl[i]] = i;
  }
  vector<int> kr;
  for (int r = m - 1; r >= 0; --r)
    if (first[r] >= 0) kr.push_back(first[r]);
  std::sort(kr.begin(), kr.end(), [](int a, int b) { return a > b; });
  return kr;
}

// X-TED_GPU/TED.cu:78-113
vector<vector<int>> standard_ted_1(vector<string>& x_node, vector<vector<int>>& x_adj,
                                   vector<string>& y_node, vector<vector<int>>& y_adj,
                                   int /*num_threads*/, int /*parallel_version*/) {
  const size_t m = x_adj.size(), n = y_adj.size();
  if (x_node.size() < m || y_node.size() < n)
    throw std::invalid_argument("xted: fewer labels than nodes");
  // Labels compare by string equality (X-TED_GPU/TED.cu:86-92): intern them.
  std::unordered_map<string, int32_t> ids;
  std::vector<int32_t> xl(m), yl(n);
  for (size_t i = 0; i < m; ++i) xl[i] = ids.emplace(x_node[i], (int32_t)ids.size()).first->second;
  for (size_t j = 0; j < n; ++j) yl[j] = ids.emplace(y_node[j], (int32_t)ids.size()).first->second;
  x_node.clear();
  y_node.clear();
  if (m == 0 || n == 0) return vector<vector<int>>(m, vector<int>(n));
  return run_pair(parents_of(x_adj), xl, parents_of(y_adj), yl);
}

// X-TED_GPU/TED.cu:116-128
int standard_ted(vector<string>& x_node, vector<vector<int>>& x_adj, vector<string>& y_node,
                 vector<vector<int>>& y_adj, int num_threads, int parallel_version) {
  const int m = (int)x_adj.size(), n = (int)y_adj.size();
  if (m == 0) return n;
  if (n == 0) return m;
  vector<vector<int>> d = standard_ted_1(x_node, x_adj, y_node, y_adj, num_threads, parallel_version);
  return d[0][0];
}

// X-TED_GPU/parallel_X_TED.cu:169-399. Labels are recovered from Delta, which
// must be a label-equality matrix; orl / keyroots are recomputed on the
// device from the adjacency lists, D is not used.
vector<vector<int>> parallel_standard_ted_test(vector<int>& /*x_orl*/, vector<int>& /*x_kr*/,
                                               vector<int>& /*y_orl*/, vector<int>& /*y_kr*/,
                                               vector<vector<int>>& Delta,
                                               vector<vector<int>>& /*D*/,
                                               vector<vector<int>>& /*D_tree*/, int m, int n,
                                               int /*num_threads*/, vector<vector<int>>& x_adj,
                                               vector<vector<int>>& y_adj) {
  if ((int)x_adj.size() != m || (int)y_adj.size() != n || (int)Delta.size() != m)
    throw std::invalid_argument("xted: size mismatch");
  // Group y nodes by the set of x nodes they match; x nodes by their row.
  std::vector<int32_t> xl(m, -1), yl(n, -1);
  int32_t next = 0;
  for (int i = 0; i < m; ++i) {
    if ((int)Delta[i].size() != n) throw std::invalid_argument("xted: Delta row size mismatch");
    if (xl[i] < 0) {
      xl[i] = next++;
      for (int k = i + 1; k < m; ++k)
        if (Delta[k] == Delta[i]) xl[k] = xl[i];
    }
  }
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) {
      if (Delta[i][j] == 0) {
        if (yl[j] >= 0 && yl[j] != xl[i])
          throw std::invalid_argument("xted: Delta is not a label-equality matrix");
        yl[j] = xl[i];
      } else if (Delta[i][j] != 1) {
        throw std::invalid_argument("xted: Delta entries must be 0 or 1");
      }
    }
    if (yl[j] < 0) yl[j] = next++;
  }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j)
      if ((Delta[i][j] == 0) != (xl[i] == yl[j]))
        throw std::invalid_argument("xted: Delta is not a label-equality matrix");
  if (m == 0 || n == 0) return vector<vector<int>>(m, vector<int>(n));
  return run_pair(parents_of(x_adj), xl, parents_of(y_adj), yl);
}

// ---- text formats (X-TED_GPU/dataset_processing.cu) ------------------------

int ToInt(const string& str) { return std::stoi(str); }

// X-TED_GPU/dataset_processing.cu:7-36: "[1, 2]" -> {1, 2}
vector<int> To_Vector(string str) {
  vector<int> out;
  if (str.size() < 2) return out;
  str = str.substr(1, str.size() - 2);
  string word;
  for (char c : str) {
    if (c == ',') {
      out.push_back(ToInt(word));
      word.clear();
    } else if (c != ' ') {
      word.push_back(c);
    }
  }
  if (!word.empty()) out.push_back(ToInt(word));
  return out;
}

// X-TED_GPU/dataset_processing.cu:38-56: the "[...]" groups of a line.
vector<string> splitting(string str) {
  vector<string> parts;
  if (str.size() < 2) return parts;
  str = str.substr(1, str.size() - 2);
  size_t i = 0;
  while ((i = str.find('[', i)) != string::npos) {
    size_t close = str.find(']', i);
    if (close == string::npos) throw std::invalid_argument("xted: unterminated '[' group");
    parts.push_back(str.substr(i, close - i + 1));
    i = close + 1;
  }
  return parts;
}

// X-TED_GPU/dataset_processing.cu:59-67
vector<vector<int>> pre_process(string str) {
  vector<vector<int>> adj;
  if (xted::parse_adjacency(str, adj) != xted::OK)
    throw std::invalid_argument("xted: malformed adjacency line");
  return adj;
}

// X-TED_GPU/dataset_processing.cu:69-82
vector<string> node_process(string str) { return xted::split_labels(str); }

// X-TED_GPU/test_TED.cu:3-56
void test(char* input_1, char* input_2) {
  auto read_lines = [](const char* path) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument(string("xted: cannot open ") + path);
    vector<string> lines;
    string line;
    while (std::getline(in, line)) lines.push_back(line);
    return lines;
  };
  vector<string> nodes = read_lines(input_1), adjs = read_lines(input_2);
  if (nodes.size() < 2 || adjs.size() < 2)
    throw std::invalid_argument("xted: the input files need at least two trees");
  vector<string> n1 = node_process(nodes[0]), n2 = node_process(nodes[1]);
  vector<vector<int>> a1 = pre_process(adjs[0]), a2 = pre_process(adjs[1]);
  const int d = standard_ted(n1, a1, n2, a2, 0, 36);
  std::printf("The distance = %d\n", d);
}
