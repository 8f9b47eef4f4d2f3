// Fill-reducing node order: nested dissection by coordinate planes or by breadth-first level
// structures, whichever fills the factor less (crates/solver/src/ordering.rs:1-281). The order
// decides the equation numbering of K_ff, so it is reproduced exactly: the same bisection
// (median of the longest bounding-box axis, ties by node index, the smaller boundary side as
// separator, parts of at most 16 nodes kept as they come), the same pseudo-peripheral search
// and middle level, and the same strict comparison of fills.
#pragma once

#include <cstdint>
#include <vector>

namespace fc_order {

using i64 = int64_t;

std::vector<i64> nested_dissection(i64 n, const double* xyz, const i64* adj_ptr, const i64* adj_idx);
std::vector<i64> graph_dissection(i64 n, const i64* adj_ptr, const i64* adj_idx);
i64 fill(i64 n, const i64* adj_ptr, const i64* adj_idx, const i64* order);
// Returns 1 when the graph dissection was kept.
int fill_reducing_order(i64 n, const double* xyz, const i64* adj_ptr, const i64* adj_idx, std::vector<i64>& order);

}  // namespace fc_order
