// Elimination tree, postorder and column counts of a symmetric pattern given by its upper
// triangle by columns (column k holds rows i <= k, sorted), as the reference's symbolic phase
// defines them (crates/solver/src/cholesky.rs:157-342):
//   - parent[k]: the elimination tree (Liu 1990), roots -1;
//   - counts[k]: entries of column k of L, diagonal included;
//   - postorder: children in increasing order, roots in increasing order (cholesky.rs:309-342).
// Counts come from the Gilbert-Ng-Peyton skeleton count (first descendants and least common
// ancestors of the row subtrees, J. R. Gilbert, E. G. Ng, B. W. Peyton, "An efficient algorithm
// to compute row and column counts for sparse Cholesky factorization", SIMAX 15(4), 1994), in
// time near-linear in the pattern rather than in the factor. The counts are a property of the
// pattern, so they equal the reference's row-subtree walk (cholesky.rs:184-212, 289-305) entry
// for entry. A matrix of whole 3x3 node blocks is analysed on its blocks' graph and expanded
// (the reference's uniform_blocks, cholesky.rs:230-287, whose expansion it proves equal to the
// columnwise result).
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

namespace fc_sym {

using i64 = int64_t;
constexpr i64 NONE = -1;

struct Upper {
  i64 n = 0;
  const i64* ptr = nullptr;
  const i64* idx = nullptr;
};

// Liu's elimination tree with path compression over the upper columns.
inline std::vector<i64> etree(const Upper& a) {
  std::vector<i64> parent(a.n, NONE), anc(a.n, NONE);
  for (i64 k = 0; k < a.n; ++k)
    for (i64 t = a.ptr[k]; t < a.ptr[k + 1]; ++t) {
      i64 i = a.idx[t];
      while (i != NONE && i < k) {
        i64 up = anc[i];
        anc[i] = k;
        if (up == NONE) parent[i] = k;
        i = up;
      }
    }
  return parent;
}

// Postorder with children visited in increasing order and roots in increasing order.
inline std::vector<i64> postorder(const std::vector<i64>& parent) {
  const i64 n = (i64)parent.size();
  std::vector<i64> head(n, NONE), next(n, NONE), post;
  post.reserve(n);
  for (i64 j = n - 1; j >= 0; --j)
    if (parent[j] != NONE) {
      next[j] = head[parent[j]];
      head[parent[j]] = j;
    }
  std::vector<i64> stack;
  for (i64 r = 0; r < n; ++r) {
    if (parent[r] != NONE) continue;
    stack.push_back(r);
    while (!stack.empty()) {
      i64 v = stack.back();
      i64 c = head[v];
      if (c == NONE) {
        post.push_back(v);
        stack.pop_back();
      } else {
        head[v] = next[c];
        stack.push_back(c);
      }
    }
  }
  return post;
}

// Column counts (diagonal included) by the skeleton-graph method.
inline std::vector<i64> column_counts(const Upper& a, const std::vector<i64>& parent, const std::vector<i64>& post) {
  const i64 n = a.n;
  // rows of the strictly lower triangle by column: column j holds i > j with A(i, j) != 0,
  // i.e. the upper columns i that hold row j
  std::vector<i64> lp(n + 1, 0), li;
  for (i64 k = 0; k < n; ++k)
    for (i64 t = a.ptr[k]; t < a.ptr[k + 1]; ++t)
      if (a.idx[t] < k) ++lp[a.idx[t] + 1];
  for (i64 j = 0; j < n; ++j) lp[j + 1] += lp[j];
  li.resize(lp[n]);
  {
    std::vector<i64> fill(lp.begin(), lp.end() - 1);
    for (i64 k = 0; k < n; ++k)
      for (i64 t = a.ptr[k]; t < a.ptr[k + 1]; ++t)
        if (a.idx[t] < k) li[fill[a.idx[t]]++] = k;
  }
  std::vector<i64> first(n, NONE), maxfirst(n, NONE), prevleaf(n, NONE), anc(n), delta(n, 0);
  for (i64 k = 0; k < n; ++k) {
    i64 j = post[k];
    delta[j] = first[j] == NONE ? 1 : 0;  // a leaf of the tree
    for (; j != NONE && first[j] == NONE; j = parent[j]) first[j] = k;
  }
  for (i64 j = 0; j < n; ++j) anc[j] = j;
  for (i64 k = 0; k < n; ++k) {
    i64 j = post[k];
    if (parent[j] != NONE) --delta[parent[j]];
    for (i64 t = lp[j]; t < lp[j + 1]; ++t) {
      i64 i = li[t];
      // is j a leaf of the row subtree of i?
      if (i <= j || first[j] <= maxfirst[i]) continue;
      maxfirst[i] = first[j];
      i64 jprev = prevleaf[i];
      prevleaf[i] = j;
      ++delta[j];
      if (jprev != NONE) {
        i64 q = jprev;
        while (q != anc[q]) q = anc[q];
        for (i64 s = jprev; s != q;) {
          i64 up = anc[s];
          anc[s] = q;
          s = up;
        }
        --delta[q];  // q is the least common ancestor of jprev and j
      }
    }
    if (parent[j] != NONE) anc[j] = parent[j];
  }
  for (i64 j = 0; j < n; ++j)
    if (parent[j] != NONE) delta[parent[j]] += delta[j];
  return delta;
}

// The blocks' graph when every node block of 3 unknowns is whole (cholesky.rs:230-266 defines
// the condition); false otherwise.
inline bool block_graph(const Upper& a, std::vector<i64>& bptr, std::vector<i64>& bidx) {
  constexpr i64 B = 3;
  if (a.n == 0 || a.n % B != 0) return false;
  const i64 nb = a.n / B;
  bptr.assign(1, 0);
  bidx.clear();
  for (i64 q = 0; q < nb; ++q) {
    const i64 c0 = B * q;
    const i64 s0 = a.ptr[c0], e0 = a.ptr[c0 + 1];
    if (e0 == s0 || a.idx[e0 - 1] != c0) return false;
    const i64 above = e0 - 1 - s0;
    if (above % B != 0) return false;
    for (i64 t = 0; t < above; t += B) {
      const i64 r = a.idx[s0 + t];
      if (r % B != 0 || a.idx[s0 + t + 1] != r + 1 || a.idx[s0 + t + 2] != r + 2) return false;
      bidx.push_back(r / B);
    }
    bidx.push_back(q);
    for (i64 c = 1; c < B; ++c) {
      const i64 s = a.ptr[c0 + c], e = a.ptr[c0 + c + 1];
      if (e - s != above + c + 1) return false;
      if (!std::equal(a.idx + s, a.idx + s + above, a.idx + s0)) return false;
      for (i64 t = 0; t <= c; ++t)
        if (a.idx[s + above + t] != c0 + t) return false;
    }
    bptr.push_back((i64)bidx.size());
  }
  return true;
}

// Tree and counts of the whole matrix; by_blocks tells whether the blocks' graph was used.
inline void tree_and_counts(const Upper& a, std::vector<i64>& parent, std::vector<i64>& counts, bool* by_blocks = nullptr) {
  std::vector<i64> bptr, bidx;
  if (block_graph(a, bptr, bidx)) {
    Upper b{a.n / 3, bptr.data(), bidx.data()};
    auto bp = etree(b);
    auto bc = column_counts(b, bp, postorder(bp));
    parent.resize(a.n);
    counts.resize(a.n);
    for (i64 q = 0; q < b.n; ++q)
      for (i64 c = 0; c < 3; ++c) {
        parent[3 * q + c] = c < 2 ? 3 * q + c + 1 : (bp[q] == NONE ? NONE : 3 * bp[q]);
        counts[3 * q + c] = 3 * bc[q] - c;
      }
    if (by_blocks) *by_blocks = true;
    return;
  }
  parent = etree(a);
  counts = column_counts(a, parent, postorder(parent));
  if (by_blocks) *by_blocks = false;
}

}  // namespace fc_sym
