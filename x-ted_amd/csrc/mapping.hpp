// Optimal edit mapping of two preorder trees from their tree-distance table.
#pragma once
#include <cstdint>
#include <utility>
#include <vector>

namespace xted {

// Largest m*n the host backtrace takes (its first forest table is
// (m+1) x (n+1) int32).
constexpr int64_t kMaxMappingCells = int64_t(1) << 26;

// Traces an optimal edit mapping back through the forest tables, given the
// full m x n tree-distance table td (row-major, preorder, unit costs):
// `map` receives the matched node pairs (x node, y node) in increasing
// order of both; nodes of x not matched are deleted, nodes of y not matched
// are inserted, matched pairs with different labels are renamed. Returns the
// mapping's cost, which equals td[0].
int64_t edit_mapping(const int32_t* x_parent, const int32_t* x_label, int32_t m,
                     const int32_t* y_parent, const int32_t* y_label, int32_t n,
                     const int32_t* td, std::vector<std::pair<int32_t, int32_t>>& map);

}  // namespace xted
