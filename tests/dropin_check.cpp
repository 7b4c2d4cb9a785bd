// Test driver for the reference-signature C++ API (include/TED_C++.h): a
// caller written against the reference header, compiled unchanged against
// this library. For trees 0 and 1 of a label file and an adjacency file it
// calls every entry point the way the reference's own callers do and prints
// one JSON object; the tables go to binary files (int32, row-major) beside
// the given output prefix.
//
//   dropin_check <labels_file> <adjacency_file> <out_prefix>
//
// Calls, with the reference call sites they mirror:
//   node_process / pre_process          X-TED_GPU/test_TED.cu:42-46
//   standard_ted(..., 0, 36)            X-TED_GPU/test_TED.cu:48
//   standard_ted_1                      X-TED_GPU/TED.cu:116-128 (via standard_ted)
//   outermost_right_leaves, key_roots   X-TED_GPU/TED.cu:97-100
//   parallel_standard_ted_test with the
//     Delta / D_forest / D_tree that
//     standard_ted_1 builds             X-TED_GPU/TED.cu:81-106
//   min3                                X-TED_GPU/TED.cu:5-19
#include <cstdio>
#include <exception>
#include <fstream>
#include <string>
#include <vector>

#include "../include/TED_C++.h"

namespace {

vector<string> lines_of(const char* path) {
  std::ifstream in(path);
  vector<string> out;
  string line;
  while (std::getline(in, line)) out.push_back(line);
  return out;
}

void write_table(const string& path, const vector<vector<int>>& t) {
  std::FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) throw std::runtime_error("cannot write " + path);
  for (const vector<int>& row : t) std::fwrite(row.data(), sizeof(int), row.size(), f);
  std::fclose(f);
}

void print_ints(const char* key, const vector<int>& v) {
  std::printf("\"%s\": [", key);
  for (size_t i = 0; i < v.size(); ++i) std::printf(i ? ", %d" : "%d", v[i]);
  std::printf("], ");
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 4) {
    std::fprintf(stderr, "usage: %s <labels_file> <adjacency_file> <out_prefix>\n", argv[0]);
    return 2;
  }
  try {
    const vector<string> nodes = lines_of(argv[1]), adjs = lines_of(argv[2]);
    if (nodes.size() < 2 || adjs.size() < 2) throw std::runtime_error("need two trees");
    const string prefix = argv[3];
    vector<string> xn = node_process(nodes[0]), yn = node_process(nodes[1]);
    vector<vector<int>> xa = pre_process(adjs[0]), ya = pre_process(adjs[1]);
    const int m = (int)xa.size(), n = (int)ya.size();

    std::printf("{\"m\": %d, \"n\": %d, ", m, n);
    std::printf("\"min3\": [%d, %d, %d], ", min3(1, 2, 3), min3(5, -4, 0), min3(7, 7, 6));

    // standard_ted on copies: it clears the label vectors it is given.
    vector<string> xn1 = xn, yn1 = yn;
    const int d = standard_ted(xn1, xa, yn1, ya, 0, 36);
    std::printf("\"standard_ted\": %d, \"standard_ted_cleared\": %s, ", d,
                xn1.empty() && yn1.empty() ? "true" : "false");

    vector<string> xn2 = xn, yn2 = yn;
    const vector<vector<int>> td1 = standard_ted_1(xn2, xa, yn2, ya, 0, 36);
    std::printf("\"standard_ted_1_cleared\": %s, \"standard_ted_1_shape\": [%zu, %zu], ",
                xn2.empty() && yn2.empty() ? "true" : "false", td1.size(),
                td1.empty() ? (size_t)0 : td1[0].size());
    write_table(prefix + "_standard_ted_1.bin", td1);

    vector<int> x_orl = outermost_right_leaves(xa), y_orl = outermost_right_leaves(ya);
    vector<int> x_kr = key_roots(x_orl), y_kr = key_roots(y_orl);
    print_ints("x_orl", x_orl);
    print_ints("x_kr", x_kr);
    print_ints("y_orl", y_orl);
    print_ints("y_kr", y_kr);

    // The scheduler entry point with the arguments standard_ted_1 builds.
    vector<vector<int>> Delta(m, vector<int>(n, 0));
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j)
        if (xn[i] != yn[j]) Delta[i][j] = 1;
    vector<vector<int>> D_forest(m + 1, vector<int>(n + 1, -1));
    vector<vector<int>> D_tree(m, vector<int>(n, -1));
    const vector<vector<int>> td2 = parallel_standard_ted_test(
        x_orl, x_kr, y_orl, y_kr, Delta, D_forest, D_tree, m, n, 0, xa, ya);
    write_table(prefix + "_parallel_standard_ted_test.bin", td2);
    std::printf("\"parallel_standard_ted_test_shape\": [%zu, %zu], ", td2.size(),
                td2.empty() ? (size_t)0 : td2[0].size());

    // Empty-tree rule of standard_ted (X-TED_GPU/TED.cu:119-124).
    vector<string> none;
    vector<vector<int>> none_adj;
    vector<string> xn3 = xn;
    std::printf("\"empty_x\": %d, \"empty_y\": %d, ",
                standard_ted(none, none_adj, xn3, xa, 0, 36),
                standard_ted(xn3, xa, none, none_adj, 0, 36));
    std::printf("\"device_ms\": %.3f}\n", total_milliseconds);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "dropin_check: %s\n", e.what());
    return 1;
  }
  return 0;
}
