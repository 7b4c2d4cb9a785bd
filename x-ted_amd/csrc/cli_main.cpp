// X-TED_GPU <labels_file> <adjacency_file>: the reference command line
// (X-TED_GPU/main.cu:4-14): evaluates trees 0 and 1 of the files 20 times,
// one blank line after each run.
// X-TED_GPU <labels_file> <adjacency_file> --all-pairs: the distance matrix
// of every tree in the files (the dimension x dimension matrix the
// reference's test() sizes, X-TED_GPU/test_TED.cu:30-54), one row per line.
// X-TED_GPU --bracket-to-lines <trees_file> <labels_out> <adjacency_out>:
// converts a bracket-notation file (one tree per line) to the two input
// files (host only).
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <stdexcept>
#include <unordered_map>

#include "../../include/TED_C++.h"
#include "../../include/xted.h"

namespace {

vector<string> read_lines(const char* path) {
  std::ifstream in(path);
  if (!in) throw std::invalid_argument(string("cannot open ") + path);
  vector<string> lines;
  string line;
  while (std::getline(in, line)) lines.push_back(line);
  return lines;
}

int all_pairs(const char* labels_path, const char* adj_path) {
  const vector<string> nodes = read_lines(labels_path), adjs = read_lines(adj_path);
  std::unordered_map<string, int32_t> ids;
  vector<int32_t> sizes, parents, labels;
  for (size_t t = 0; t < nodes.size() && t < adjs.size(); ++t) {
    if (adjs[t].find('[') == string::npos) continue;  // blank trailing line
    const vector<vector<int>> adj = pre_process(adjs[t]);
    const vector<string> lab = node_process(nodes[t]);
    if (lab.size() < adj.size()) throw std::invalid_argument("tree with fewer labels than nodes");
    const size_t base = parents.size();
    parents.resize(base + adj.size(), -2);
    if (!adj.empty()) parents[base] = -1;
    for (size_t i = 0; i < adj.size(); ++i) {
      for (int ch : adj[i]) {
        if (ch <= 0 || (size_t)ch >= adj.size() || parents[base + ch] != -2)
          throw std::invalid_argument("adjacency line is not a tree");
        parents[base + ch] = (int32_t)i;
      }
      labels.push_back(ids.emplace(lab[i], (int32_t)ids.size()).first->second);
    }
    sizes.push_back((int32_t)adj.size());
  }
  const int32_t n = (int32_t)sizes.size();
  xted_context* ctx = nullptr;
  int st = xted_create(0, &ctx);
  vector<int32_t> dist((size_t)n * n);
  xted_stats stats;
  if (st == XTED_OK)
    st = xted_all_pairs(ctx, n, sizes.data(), parents.data(), labels.data(), 0, dist.data(),
                        &stats);
  if (st != XTED_OK) {
    std::fprintf(stderr, "%s\n", xted_last_error());
    if (ctx) xted_destroy(ctx);
    return 1;
  }
  std::printf("Distance matrix of %d trees (%.3f ms on the device)\n", n, stats.total_ms);
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t j = 0; j < n; ++j) std::printf(j ? " %d" : "%d", dist[(size_t)i * n + j]);
    std::printf("\n");
  }
  xted_destroy(ctx);
  return 0;
}

int bracket_to_lines(const char* in_path, const char* labels_path, const char* adj_path) {
  const vector<string> trees = read_lines(in_path);
  std::ofstream lo(labels_path), ao(adj_path);
  if (!lo || !ao) throw std::invalid_argument("cannot write the output files");
  for (const string& t : trees) {
    if (t.find('{') == string::npos) continue;  // blank line
    int64_t ln = 0, an = 0;
    int st = xted_bracket_to_lines(t.c_str(), nullptr, 0, nullptr, 0, &ln, &an);
    if (st != XTED_OK && st != XTED_ERR_TOO_LARGE) {
      std::fprintf(stderr, "%s\n", xted_last_error());
      return 1;
    }
    vector<char> lb((size_t)ln), ab((size_t)an);
    st = xted_bracket_to_lines(t.c_str(), lb.data(), ln, ab.data(), an, &ln, &an);
    if (st != XTED_OK) {
      std::fprintf(stderr, "%s\n", xted_last_error());
      return 1;
    }
    lo << lb.data() << '\n';
    ao << ab.data() << '\n';
  }
  return 0;
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc >= 5 && std::strcmp(argv[1], "--bracket-to-lines") == 0) {
    try {
      return bracket_to_lines(argv[2], argv[3], argv[4]);
    } catch (const std::exception& e) {
      std::fprintf(stderr, "%s\n", e.what());
      return 1;
    }
  }
  if (argc < 3) {
    std::fprintf(stderr,
                 "usage: %s <labels_file> <adjacency_file> [--all-pairs]\n"
                 "       %s --bracket-to-lines <trees_file> <labels_out> <adjacency_out>\n",
                 argv[0], argv[0]);
    return 2;
  }
  try {
    if (argc > 3 && std::strcmp(argv[3], "--all-pairs") == 0) return all_pairs(argv[1], argv[2]);
    for (int i = 0; i < 20; ++i) {
      test(argv[1], argv[2]);
      std::printf("\n");
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
