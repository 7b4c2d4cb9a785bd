// Edit mapping by backtrace. The device computes the tree-distance table TD;
// the mapping is recovered on the host by re-deriving, for the pair of whole
// trees and then for each pair of subtrees the trace matches as a unit, the
// forest table of that pair (the recurrence of X-TED_CPU/parallel_computing.cpp:
// 4-45 in its preorder / outermost-right-leaf form, with every inner subtree
// pair priced by TD) and walking it from its top-left cell. The reference
// returns distances only.
#include "mapping.hpp"

#include <algorithm>

#include "trees.hpp"

namespace xted {

int64_t edit_mapping(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                     const int32_t* y_parent, const int32_t* y_label, int32_t n,
                     const int32_t* td, std::vector<std::pair<int32_t, int32_t>>& map) {
  map.clear();
  if (m == 0 || n == 0) return (int64_t)m + n;
  std::vector<int32_t> ox(m), oy(n);
  compute_orl(x_parent, m, ox.data());
  compute_orl(y_parent, n, oy.data());
  std::vector<std::pair<int32_t, int32_t>> todo{{0, 0}};  // subtree pairs matched as units
  std::vector<int32_t> F;
  while (!todo.empty()) {
    const int32_t i0 = todo.back().first, j0 = todo.back().second;
    todo.pop_back();
    const int32_t rmax = ox[i0] + 1, cmax = oy[j0] + 1;
    const int32_t C = cmax - j0 + 1;
    F.assign((size_t)(rmax - i0 + 1) * C, 0);
    // F(i, j): forest x[i, rmax) against forest y[j, cmax).
    auto at = [&](int32_t i, int32_t j) -> int32_t& { return F[(size_t)(i - i0) * C + (j - j0)]; };
    for (int32_t j = j0; j <= cmax; ++j) at(rmax, j) = cmax - j;
    for (int32_t i = i0; i <= rmax; ++i) at(i, cmax) = rmax - i;
    for (int32_t i = rmax - 1; i >= i0; --i) {
      for (int32_t j = cmax - 1; j >= j0; --j) {
        const int32_t edge = std::min(at(i + 1, j), at(i, j + 1)) + 1;
        const bool tree = ox[i] == ox[i0] && oy[j] == oy[j0];
        const int32_t match = tree ? at(i + 1, j + 1) + (x_label[i] != y_label[j])
                                   : at(ox[i] + 1, oy[j] + 1) + td[(int64_t)i * n + j];
        at(i, j) = std::min(edge, match);
      }
    }
    int32_t i = i0, j = j0;
    while (i < rmax && j < cmax) {
      const int32_t v = at(i, j);
      if (ox[i] == ox[i0] && oy[j] == oy[j0]) {
        if (at(i + 1, j + 1) + (x_label[i] != y_label[j]) == v) {
          map.emplace_back(i, j);  // the two roots of the current trees
          ++i;
          ++j;
          continue;
        }
      } else if (at(ox[i] + 1, oy[j] + 1) + td[(int64_t)i * n + j] == v) {
        todo.emplace_back(i, j);  // sub(i) against sub(j), as a unit
        i = ox[i] + 1;
        j = oy[j] + 1;
        continue;
      }
      if (at(i + 1, j) + 1 == v) ++i;  // delete x node i
      else ++j;                         // insert y node j
    }
  }
  std::sort(map.begin(), map.end());
  int64_t cost = (int64_t)m + n - 2 * (int64_t)map.size();
  for (const auto& p : map) cost += x_label[p.first] != y_label[p.second];
  return cost;
}

}  // namespace xted
