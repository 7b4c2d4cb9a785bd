// Host-side tree model, generators and text formats.
#include "trees.hpp"

#include <algorithm>
#include <cmath>
#include <random>
#include <stdexcept>

namespace xted {

// ---------------------------------------------------------------------------
// validation / conversion
// ---------------------------------------------------------------------------

int validate_preorder(const int32_t* parent, int32_t n) {
  if (n < 0) return ERR_INVALID_ARG;
  if (n == 0) return OK;
  if (!parent || parent[0] != -1) return ERR_NOT_PREORDER;
  // path[] holds the root -> (i-1) path; node i must hang off it.
  std::vector<int32_t> path;
  path.reserve(64);
  path.push_back(0);
  for (int32_t i = 1; i < n; ++i) {
    int32_t p = parent[i];
    if (p < 0 || p >= i) return ERR_NOT_PREORDER;
    while (!path.empty() && path.back() != p) path.pop_back();
    if (path.empty()) return ERR_NOT_PREORDER;
    path.push_back(i);
  }
  return OK;
}

int adj_to_parent(const std::vector<std::vector<int>>& adj, std::vector<int32_t>& parent) {
  const int32_t n = (int32_t)adj.size();
  parent.assign(n, -2);
  if (n == 0) return OK;
  parent[0] = -1;
  for (int32_t i = 0; i < n; ++i) {
    for (int c : adj[i]) {
      if (c <= 0 || c >= n || parent[c] != -2) return ERR_NOT_PREORDER;
      parent[c] = i;
    }
  }
  for (int32_t i = 1; i < n; ++i)
    if (parent[i] == -2) return ERR_NOT_PREORDER;
  // The children lists must enumerate the tree in preorder: a DFS that follows
  // them in order visits 0, 1, 2, ... n-1.
  std::vector<int32_t> stack{0};
  int32_t expect = 0;
  while (!stack.empty()) {
    int32_t v = stack.back();
    stack.pop_back();
    if (v != expect) return ERR_NOT_PREORDER;
    ++expect;
    const auto& ch = adj[v];
    for (auto it = ch.rbegin(); it != ch.rend(); ++it) stack.push_back(*it);
  }
  return expect == n ? OK : ERR_NOT_PREORDER;
}

std::vector<std::vector<int>> parent_to_adj(const int32_t* parent, int32_t n) {
  std::vector<std::vector<int>> adj(n);
  for (int32_t i = 1; i < n; ++i) adj[parent[i]].push_back(i);
  return adj;
}

// ---------------------------------------------------------------------------
// structural arrays
// ---------------------------------------------------------------------------

void compute_orl(const int32_t* parent, int32_t n, int32_t* orl) {
  std::vector<int32_t> size(n, 1);
  for (int32_t i = n - 1; i >= 1; --i) size[parent[i]] += size[i];
  for (int32_t i = 0; i < n; ++i) orl[i] = i + size[i] - 1;
}

std::vector<int32_t> keyroots_desc(const int32_t* parent, const int32_t* orl, int32_t n) {
  std::vector<int32_t> kr;
  for (int32_t i = n - 1; i >= 0; --i)
    if (i == 0 || orl[parent[i]] != orl[i]) kr.push_back(i);
  return kr;
}

void keyroot_height(const int32_t* parent, const int32_t* orl, int32_t n, int32_t* kd) {
  // up[i] = max height over keyroots inside subtree(i) (self included), -1 if none.
  std::vector<int32_t> inner(n, -1), up(n, -1);
  for (int32_t i = n - 1; i >= 0; --i) {
    const bool is_kr = (i == 0) || orl[parent[i]] != orl[i];
    kd[i] = is_kr ? inner[i] + 1 : -1;
    up[i] = is_kr ? std::max(inner[i], kd[i]) : inner[i];
    if (i > 0) inner[parent[i]] = std::max(inner[parent[i]], up[i]);
  }
}

int32_t line_stack_depth(const int32_t* parent, const int32_t* orl, int32_t n) {
  int32_t sp = 0, deepest = 0;
  for (int32_t c = n - 1; c > 0; --c) {
    if (orl[c] < orl[parent[c]]) --sp;  // has a next sibling: its value is popped
    if (parent[c] != c - 1) deepest = std::max(deepest, ++sp);  // has a previous one
  }
  return deepest;
}

int64_t keyroot_size_sum(const int32_t* orl, const std::vector<int32_t>& kr) {
  int64_t s = 0;
  for (int32_t k : kr) s += (int64_t)orl[k] - k + 1;
  return s;
}

int32_t tree_depth(const int32_t* parent, int32_t n) {
  std::vector<int32_t> d(n, 0);
  int32_t best = 0;
  for (int32_t i = 1; i < n; ++i) {
    d[i] = d[parent[i]] + 1;
    best = std::max(best, d[i]);
  }
  return best;
}

// ---------------------------------------------------------------------------
// generators
// ---------------------------------------------------------------------------

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint64_t derive_seed(uint64_t config, uint64_t index) {
  return splitmix64(splitmix64(config) ^ (index * 0xD1B54A32D192ED03ull));
}

namespace {

struct Rng {
  std::mt19937_64 eng;
  explicit Rng(uint64_t seed) : eng(seed) {}
  // Uniform integer in [0, r) by 64x64->128 multiply-high (no library
  // distribution objects, so the stream is the same with any standard library).
  uint32_t below(uint32_t r) {
    return (uint32_t)(((unsigned __int128)eng() * (unsigned __int128)r) >> 64);
  }
  double unit() { return (double)(eng() >> 11) * (1.0 / 9007199254740992.0); }
  // Poisson(lambda) by multiplication of uniforms (Knuth); lambda is small here.
  int poisson(double lambda) {
    const double limit = std::exp(-lambda);
    int k = 0;
    double p = unit();
    while (p > limit) {
      ++k;
      p *= unit();
    }
    return k;
  }
};

// Renumbers a tree given as insertion-ordered children lists into preorder.
std::vector<int32_t> to_preorder(const std::vector<std::vector<int32_t>>& children) {
  const int32_t n = (int32_t)children.size();
  std::vector<int32_t> new_id(n, -1), order;
  order.reserve(n);
  std::vector<int32_t> stack{0};
  while (!stack.empty()) {
    int32_t v = stack.back();
    stack.pop_back();
    new_id[v] = (int32_t)order.size();
    order.push_back(v);
    const auto& ch = children[v];
    for (auto it = ch.rbegin(); it != ch.rend(); ++it) stack.push_back(*it);
  }
  std::vector<int32_t> parent(n, -1);
  for (int32_t v = 0; v < n; ++v)
    for (int32_t c : children[v]) parent[new_id[c]] = new_id[v];
  return parent;
}

// Branching-process growth: node v receives draw(v) children until n nodes
// exist. Unexpanded nodes are taken first-in-first-out (breadth first) or
// last-in-first-out (depth first). A process that dies out before n nodes is
// restarted on the same stream.
template <class Draw>
std::vector<std::vector<int32_t>> grow(int32_t n, bool depth_first, Draw draw) {
  for (;;) {
    std::vector<std::vector<int32_t>> children(1);
    size_t head = 0;
    int32_t count = 1;
    std::vector<int32_t> pending{0};
    while (count < n && head < pending.size()) {
      int32_t v;
      if (depth_first) {
        v = pending.back();
        pending.pop_back();
      } else {
        v = pending[head++];
      }
      int k = draw(v);
      k = std::min<int>(k, n - count);
      for (int c = 0; c < k; ++c) {
        children.emplace_back();
        children[v].push_back(count);
        pending.push_back(count);
        ++count;
      }
    }
    if (count == n) return children;
  }
}

}  // namespace

// Fan-out law of the deep/narrow generator: P(k) proportional to q^k on
// {0,1,2,3}; q = 0.66 puts the mean fan-out at ~1.00 (a critical branching
// process). Grown depth first, 1000-node trees have depth ~100 (about 3 in 4
// fall in [60, 120]; generate_banded() rejects the rest).
static const double kDeepQ = 0.66;

Tree generate(int kind, int32_t n, int32_t alphabet, uint64_t seed) {
  if (n <= 0) return Tree{};
  if (alphabet <= 0) alphabet = 1;
  Rng rng(seed);
  std::vector<std::vector<int32_t>> children;
  if (kind == GEN_RRT) {
    children.assign(n, {});
    for (int32_t i = 1; i < n; ++i) children[rng.below((uint32_t)i)].push_back(i);
  } else if (kind == GEN_DEEP) {
    double w[4], z = 0;
    for (int k = 0; k < 4; ++k) z += (w[k] = std::pow(kDeepQ, k));
    double cdf[4], acc = 0;
    for (int k = 0; k < 4; ++k) cdf[k] = (acc += w[k] / z);
    children = grow(n, /*depth_first=*/true, [&](int32_t) {
      double u = rng.unit();
      int k = 0;
      while (k < 3 && u >= cdf[k]) ++k;
      return k;
    });
  } else if (kind == GEN_SHALLOW) {
    const int root_fanout = 40;
    children = grow(n, /*depth_first=*/false,
                    [&](int32_t v) { return v == 0 ? root_fanout : rng.poisson(4.0); });
  } else {
    throw std::invalid_argument("unknown generator kind");
  }
  Tree t;
  t.parent = to_preorder(children);
  t.label.resize(n);
  for (int32_t i = 0; i < n; ++i) t.label[i] = (int32_t)rng.below((uint32_t)alphabet);
  return t;
}

Tree generate_banded(int kind, int32_t n, int32_t alphabet, uint64_t seed, int32_t depth_min,
                     int32_t depth_max) {
  for (uint64_t attempt = 0;; ++attempt) {
    Tree t = generate(kind, n, alphabet, attempt == 0 ? seed : derive_seed(seed, attempt));
    int32_t d = tree_depth(t.parent.data(), n);
    if (d >= depth_min && d <= depth_max) return t;
    if (attempt > 10000) throw std::runtime_error("depth band not reachable");
  }
}

// ---------------------------------------------------------------------------
// text formats
// ---------------------------------------------------------------------------

std::vector<std::string> split_labels(const std::string& line) {
  std::vector<std::string> out;
  std::string part;
  for (char c : line) {
    if (c != ' ') {
      part.push_back(c);
    } else {
      out.push_back(part);
      part.clear();
    }
  }
  return out;
}

int parse_adjacency(const std::string& line, std::vector<std::vector<int>>& adj) {
  adj.clear();
  // Drop the outer brackets the way the reference does (first and last
  // character), then read every "[...]" group.
  if (line.size() < 2) return ERR_PARSE;
  const std::string body = line.substr(1, line.size() - 2);
  size_t i = 0;
  while (i < body.size()) {
    if (body[i] != '[') {
      ++i;
      continue;
    }
    size_t close = body.find(']', i);
    if (close == std::string::npos) return ERR_PARSE;
    std::vector<int> kids;
    std::string tok;
    auto flush = [&]() -> bool {
      if (tok.empty()) return true;
      try {
        size_t used = 0;
        long v = std::stol(tok, &used);
        if (used != tok.size() || v < INT32_MIN || v > INT32_MAX) return false;
        kids.push_back((int)v);
      } catch (...) {
        return false;
      }
      tok.clear();
      return true;
    };
    for (size_t k = i + 1; k < close; ++k) {
      char c = body[k];
      if (c == ',') {
        if (tok.empty() || !flush()) return ERR_PARSE;
      } else if (c != ' ') {
        tok.push_back(c);
      }
    }
    if (!flush()) return ERR_PARSE;
    adj.push_back(std::move(kids));
    i = close + 1;
  }
  return OK;
}

std::string format_labels(const std::vector<std::string>& labels) {
  std::string s;
  for (const auto& l : labels) {
    s += l;
    s.push_back(' ');
  }
  return s;
}

std::string format_adjacency(const std::vector<std::vector<int>>& adj) {
  std::string s = "[";
  for (size_t i = 0; i < adj.size(); ++i) {
    if (i) s += ", ";
    s.push_back('[');
    for (size_t k = 0; k < adj[i].size(); ++k) {
      if (k) s += ", ";
      s += std::to_string(adj[i][k]);
    }
    s.push_back(']');
  }
  s.push_back(']');
  return s;
}

int parse_bracket(const std::string& text, std::vector<std::string>& labels,
                  std::vector<int32_t>& parent) {
  labels.clear();
  parent.clear();
  std::vector<int32_t> stack;
  size_t i = 0;
  // skip leading whitespace
  while (i < text.size() && (text[i] == ' ' || text[i] == '\n' || text[i] == '\r')) ++i;
  bool done = false;
  while (i < text.size()) {
    char c = text[i];
    if (c == '{') {
      if (done) return ERR_PARSE;
      ++i;
      size_t start = i;
      while (i < text.size() && text[i] != '{' && text[i] != '}') ++i;
      labels.push_back(text.substr(start, i - start));
      parent.push_back(stack.empty() ? -1 : stack.back());
      stack.push_back((int32_t)labels.size() - 1);
    } else if (c == '}') {
      if (stack.empty()) return ERR_PARSE;
      stack.pop_back();
      if (stack.empty()) done = true;
      ++i;
    } else if (c == ' ' || c == '\n' || c == '\r') {
      if (!done) return ERR_PARSE;
      ++i;
    } else {
      return ERR_PARSE;
    }
  }
  if (!stack.empty() || labels.empty()) return ERR_PARSE;
  return OK;
}

std::string format_bracket(const std::vector<std::string>& labels, const int32_t* parent,
                           int32_t n) {
  std::string s;
  std::vector<int32_t> path;
  for (int32_t i = 0; i < n; ++i) {
    while (!path.empty() && path.back() != parent[i]) {
      s.push_back('}');
      path.pop_back();
    }
    s.push_back('{');
    s += labels[i];
    path.push_back(i);
  }
  while (!path.empty()) {
    s.push_back('}');
    path.pop_back();
  }
  return s;
}

}  // namespace xted
