// Fill-reducing order (see ordering.hpp). Each part is a range of one working buffer; a split
// rearranges its range into [first part | second part | separator] and the ranges are ordered
// depth-first, first part, then second part, then the separator (ordering.rs:93-95, 249-251).
#include "ordering.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

#include "symbolic_core.hpp"

namespace fc_order {

namespace {

constexpr i64 LEAF = 16;               // ordering.rs:25
constexpr int PERIPHERAL_SEARCHES = 8; // ordering.rs:113

// f64::total_cmp as a signed integer key.
inline int64_t total_order_key(double x) {
  int64_t b;
  std::memcpy(&b, &x, sizeof b);
  return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);
}

struct Range {
  i64 lo, hi;
  bool emit;
};

}  // namespace

std::vector<i64> nested_dissection(i64 n, const double* xyz, const i64* adj_ptr, const i64* adj_idx) {
  std::vector<i64> buf(n), out;
  out.reserve(n);
  for (i64 v = 0; v < n; ++v) buf[v] = v;
  std::vector<uint8_t> side(n, 0);  // 1 first half, 2 second half, 3 separator
  std::vector<std::pair<int64_t, i64>> keyed;
  std::vector<i64> first, second, sep;
  std::vector<Range> work{{0, n, false}};
  while (!work.empty()) {
    Range r = work.back();
    work.pop_back();
    const i64 len = r.hi - r.lo;
    if (r.emit || len <= LEAF) {
      out.insert(out.end(), buf.begin() + r.lo, buf.begin() + r.hi);
      continue;
    }
    i64* part = buf.data() + r.lo;
    double lo[3], hi[3];
    for (int d = 0; d < 3; ++d) { lo[d] = std::numeric_limits<double>::infinity(); hi[d] = -lo[d]; }
    for (i64 t = 0; t < len; ++t)
      for (int d = 0; d < 3; ++d) {
        double x = xyz[3 * part[t] + d];
        lo[d] = std::fmin(lo[d], x);
        hi[d] = std::fmax(hi[d], x);
      }
    int axis = 0;
    for (int d = 1; d < 3; ++d)
      if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
    keyed.resize(len);
    for (i64 t = 0; t < len; ++t) keyed[t] = {total_order_key(xyz[3 * part[t] + axis]), part[t]};
    std::sort(keyed.begin(), keyed.end());
    const i64 half = len / 2;
    for (i64 t = 0; t < len; ++t) {
      part[t] = keyed[t].second;
      side[part[t]] = t < half ? 1 : 2;
    }
    auto touches = [&](i64 v, uint8_t other) {
      for (i64 q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q)
        if (side[adj_idx[q]] == other) return true;
      return false;
    };
    i64 lb = 0, rb = 0;
    for (i64 t = 0; t < half; ++t) lb += touches(part[t], 2);
    for (i64 t = half; t < len; ++t) rb += touches(part[t], 1);
    const bool left_sep = lb <= rb;  // the smaller boundary side, the left one on a tie
    first.clear(); second.clear(); sep.clear();
    for (i64 t = 0; t < len; ++t) {
      const i64 v = part[t];
      const bool in_left = t < half;
      if (in_left == left_sep && touches(v, in_left ? 2 : 1)) sep.push_back(v);
      else (in_left ? first : second).push_back(v);
    }
    for (i64 t = 0; t < len; ++t) side[part[t]] = 0;
    std::copy(first.begin(), first.end(), part);
    std::copy(second.begin(), second.end(), part + first.size());
    std::copy(sep.begin(), sep.end(), part + first.size() + second.size());
    const i64 a = r.lo, b = a + (i64)first.size(), s = b + (i64)second.size();
    work.push_back({s, r.hi, true});
    work.push_back({b, s, false});
    work.push_back({a, b, false});
  }
  return out;
}

std::vector<i64> graph_dissection(i64 n, const i64* adj_ptr, const i64* adj_idx) {
  constexpr i64 UNSEEN = std::numeric_limits<i64>::max();
  std::vector<i64> buf(n), out;
  out.reserve(n);
  for (i64 v = 0; v < n; ++v) buf[v] = v;
  std::vector<i64> owner(n, 0), level(n, UNSEEN);
  i64 token = 0;
  // a level structure: nodes by level, flat, with level starts
  struct Levels {
    std::vector<i64> nodes, start;
    i64 depth() const { return (i64)start.size() - 1; }
  };
  auto grow = [&](i64 root, Levels& s) {
    s.nodes.assign(1, root);
    s.start.assign({0, 1});
    level[root] = 0;
    for (;;) {
      const i64 d = s.depth();
      const i64 b = s.start[d - 1], e = s.start[d];
      for (i64 t = b; t < e; ++t) {
        const i64 v = s.nodes[t];
        for (i64 q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q) {
          const i64 w = adj_idx[q];
          if (owner[w] == token && level[w] == UNSEEN) {
            level[w] = d;
            s.nodes.push_back(w);
          }
        }
      }
      if ((i64)s.nodes.size() == e) return;
      s.start.push_back((i64)s.nodes.size());
    }
  };
  auto forget = [&](const Levels& s) {
    for (i64 v : s.nodes) level[v] = UNSEEN;
  };
  Levels cur, alt;
  std::vector<i64> before, after, sep, rest;
  std::vector<Range> work{{0, n, false}};
  while (!work.empty()) {
    Range r = work.back();
    work.pop_back();
    const i64 len = r.hi - r.lo;
    if (r.emit || len <= LEAF) {
      out.insert(out.end(), buf.begin() + r.lo, buf.begin() + r.hi);
      continue;
    }
    i64* part = buf.data() + r.lo;
    ++token;
    for (i64 t = 0; t < len; ++t) owner[part[t]] = token;
    auto degree = [&](i64 v) {
      i64 d = 0;
      for (i64 q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q) d += owner[adj_idx[q]] == token;
      return d;
    };
    grow(part[0], cur);
    for (int it = 0; it < PERIPHERAL_SEARCHES; ++it) {
      // the least connected node of the last level, lowest index on ties
      i64 best = -1, best_deg = 0;
      for (i64 t = cur.start[cur.depth() - 1]; t < cur.start[cur.depth()]; ++t) {
        const i64 v = cur.nodes[t], d = degree(v);
        if (best < 0 || d < best_deg || (d == best_deg && v < best)) { best = v; best_deg = d; }
      }
      forget(cur);
      grow(best, alt);
      if (alt.depth() <= cur.depth()) {
        forget(alt);
        for (i64 d = 0; d < cur.depth(); ++d)
          for (i64 t = cur.start[d]; t < cur.start[d + 1]; ++t) level[cur.nodes[t]] = d;
        break;
      }
      std::swap(cur, alt);
    }
    const i64 reached = (i64)cur.nodes.size();
    if (reached < len) {
      // disconnected: the reached component (sorted), then the rest in the part's order
      rest.clear();
      for (i64 t = 0; t < len; ++t)
        if (level[part[t]] == UNSEEN) rest.push_back(part[t]);
      forget(cur);
      std::vector<i64> comp(cur.nodes);
      std::sort(comp.begin(), comp.end());
      std::copy(comp.begin(), comp.end(), part);
      std::copy(rest.begin(), rest.end(), part + reached);
      work.push_back({r.lo + reached, r.hi, false});
      work.push_back({r.lo, r.lo + reached, false});
      continue;
    }
    const i64 depth = cur.depth();
    if (depth < 3) {
      forget(cur);
      out.insert(out.end(), part, part + len);
      continue;
    }
    // the middle level: the first (from level 1 on) at which the levels from 1 hold half
    // the part, kept off the last level (ordering.rs:214-224)
    i64 middle = 1, count = 0;
    for (i64 d = 1; d < depth; ++d) {
      count += cur.start[d + 1] - cur.start[d];
      if (2 * count >= len) {
        middle = std::min(d, depth - 2);
        break;
      }
    }
    before.assign(cur.nodes.begin(), cur.nodes.begin() + cur.start[middle]);
    sep.clear();
    for (i64 t = cur.start[middle]; t < cur.start[middle + 1]; ++t) {
      const i64 v = cur.nodes[t];
      bool next = false;
      for (i64 q = adj_ptr[v]; q < adj_ptr[v + 1] && !next; ++q) {
        const i64 w = adj_idx[q];
        next = owner[w] == token && level[w] == middle + 1;
      }
      (next ? sep : before).push_back(v);
    }
    after.assign(cur.nodes.begin() + cur.start[middle + 1], cur.nodes.end());
    forget(cur);
    std::sort(before.begin(), before.end());
    std::sort(sep.begin(), sep.end());
    std::sort(after.begin(), after.end());
    std::copy(before.begin(), before.end(), part);
    std::copy(after.begin(), after.end(), part + before.size());
    std::copy(sep.begin(), sep.end(), part + before.size() + after.size());
    const i64 a = r.lo, b = a + (i64)before.size(), s = b + (i64)after.size();
    work.push_back({s, r.hi, true});
    work.push_back({b, s, false});
    work.push_back({a, b, false});
  }
  return out;
}

i64 fill(i64 n, const i64* adj_ptr, const i64* adj_idx, const i64* order) {
  // the node graph eliminated in `order`, as upper columns (ordering.rs:258-281)
  std::vector<i64> pos(n);
  for (i64 k = 0; k < n; ++k) pos[order[k]] = k;
  std::vector<i64> ptr(n + 1, 0), idx;
  idx.reserve(adj_ptr[n] / 2 + n);
  std::vector<i64> col;
  for (i64 k = 0; k < n; ++k) {
    const i64 v = order[k];
    col.clear();
    for (i64 q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q)
      if (pos[adj_idx[q]] < k) col.push_back(pos[adj_idx[q]]);
    col.push_back(k);
    std::sort(col.begin(), col.end());
    col.erase(std::unique(col.begin(), col.end()), col.end());
    idx.insert(idx.end(), col.begin(), col.end());
    ptr[k + 1] = (i64)idx.size();
  }
  fc_sym::Upper a{n, ptr.data(), idx.data()};
  std::vector<i64> parent, counts;
  fc_sym::tree_and_counts(a, parent, counts);
  i64 total = 0;
  for (i64 c : counts) total += c;
  return total;
}

int fill_reducing_order(i64 n, const double* xyz, const i64* adj_ptr, const i64* adj_idx, std::vector<i64>& order) {
  order = nested_dissection(n, xyz, adj_ptr, adj_idx);
  if (n <= LEAF) return 0;
  std::vector<i64> by_graph = graph_dissection(n, adj_ptr, adj_idx);
  if (fill(n, adj_ptr, adj_idx, by_graph.data()) < fill(n, adj_ptr, adj_idx, order.data())) {
    order.swap(by_graph);
    return 1;
  }
  return 0;
}

}  // namespace fc_order
