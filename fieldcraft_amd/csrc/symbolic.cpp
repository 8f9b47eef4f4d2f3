// Symbolic phase and execution plan (see symbolic.hpp).
#include "symbolic.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <numeric>

#include "symbolic_core.hpp"

namespace fc {

namespace {

constexpr i64 PARALLEL_WORK = i64(1) << 21;  // cholesky.rs:74
constexpr i64 SUBTREE_SHARE = 256;           // cholesky.rs:97

// sum_{c<w} (m - c)^2, the dense flops of a supernode or front (cholesky.rs:1334-1339 measure)
double trapezoid_flops(double w, double m) {
  return w * m * m - m * w * (w - 1.0) + (w - 1.0) * w * (2.0 * w - 1.0) / 6.0;
}

// First-fit allocator of the update-matrix arena (element offsets).
struct Arena {
  std::map<i64, i64> free;  // offset -> size
  i64 end = 0;
  i64 take(i64 size) {
    for (auto it = free.begin(); it != free.end(); ++it)
      if (it->second >= size) {
        i64 at = it->first, left = it->second - size;
        free.erase(it);
        if (left > 0) free[at + size] = left;
        return at;
      }
    // grow, reusing a free block that ends at the top
    if (!free.empty()) {
      auto last = std::prev(free.end());
      if (last->first + last->second == end) {
        i64 at = last->first;
        free.erase(last);
        end = at + size;
        return at;
      }
    }
    i64 at = end;
    end += size;
    return at;
  }
  void give(i64 at, i64 size) {
    auto it = free.emplace(at, size).first;
    auto nx = std::next(it);
    if (nx != free.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free.erase(nx);
    }
    if (it != free.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free.erase(it);
      }
    }
  }
};

inline i64 round_up(i64 x, i64 a) { return (x + a - 1) / a * a; }

}  // namespace

Failure analyze(i64 n, const i64* ptr, const i64* idx, i64 max_entries, const PlanOptions& opt, Symbolic& s) {
  Failure fail;
  auto internal = [&](const char* what) {
    fail.status = INTERNAL;
    fail.message = std::string("supernodal structure: ") + what;
    return fail;
  };
  if (n < 0 || (n > 0 && (!ptr || (!idx && ptr[n] > 0)))) {
    fail.status = INVALID_ARGUMENT;
    fail.message = "null pattern";
    return fail;
  }
  // the pattern must be upper columns with sorted rows (cholesky.rs:99-108)
  if (n > 0 && ptr[0] != 0) {
    fail.status = INVALID_ARGUMENT;
    fail.message = "ptr[0] != 0";
    return fail;
  }
  for (i64 k = 0; k < n; ++k) {
    if (ptr[k + 1] < ptr[k]) { fail.status = INVALID_ARGUMENT; fail.message = "ptr decreases"; return fail; }
    for (i64 t = ptr[k]; t < ptr[k + 1]; ++t)
      if (idx[t] < 0 || idx[t] > k || (t > ptr[k] && idx[t] <= idx[t - 1])) {
        fail.status = INVALID_ARGUMENT;
        fail.message = "upper columns must hold sorted rows i <= k";
        return fail;
      }
  }
  s = Symbolic();
  s.n = n;
  const i64 nnz = n > 0 ? ptr[n] : 0;
  s.upper_ptr.assign(ptr, ptr + n + 1);
  if (n == 0) s.upper_ptr.assign(1, 0);
  s.upper_idx.assign(idx, idx + nnz);
  fc_sym::Upper a{n, s.upper_ptr.data(), s.upper_idx.data()};

  // tree and counts (cholesky.rs:1215-1220), refused past the cap before anything else
  fc_sym::tree_and_counts(a, s.parent0, s.counts0, &s.by_blocks);
  i64 total = 0;
  for (i64 c : s.counts0) total += c;
  s.entries = total;
  if (total > max_entries) {
    fail.status = TOO_LARGE;
    fail.entries = total;
    fail.limit = max_entries;
    return fail;
  }
  // postorder (cholesky.rs:1221-1233)
  s.perm = fc_sym::postorder(s.parent0);
  if ((i64)s.perm.size() != n) return internal("postorder");
  std::vector<i64> inv(n), parent(n), counts(n);
  for (i64 k = 0; k < n; ++k) inv[s.perm[k]] = k;
  for (i64 k = 0; k < n; ++k) {
    i64 p = s.parent0[s.perm[k]];
    parent[k] = p == NONE ? NONE : inv[p];
    counts[k] = s.counts0[s.perm[k]];
  }
  // the lower triangle in postorder, by columns (rows unsorted; they are merged below)
  std::vector<i64> lp(n + 1, 0), li(nnz);
  for (i64 k = 0; k < n; ++k)
    for (i64 t = ptr[k]; t < ptr[k + 1]; ++t) ++lp[std::min(inv[idx[t]], inv[k]) + 1];
  for (i64 j = 0; j < n; ++j) lp[j + 1] += lp[j];
  {
    std::vector<i64> at(lp.begin(), lp.end() - 1);
    for (i64 k = 0; k < n; ++k)
      for (i64 t = ptr[k]; t < ptr[k + 1]; ++t) {
        i64 r = std::max(inv[idx[t]], inv[k]), c = std::min(inv[idx[t]], inv[k]);
        li[at[c]++] = r;
      }
  }
  // fundamental supernodes (cholesky.rs:1236-1253)
  std::vector<i64> nkids(n, 0);
  for (i64 j = 0; j < n; ++j)
    if (parent[j] != NONE) ++nkids[parent[j]];
  s.first.assign(1, 0);
  for (i64 j = 1; j < n; ++j)
    if (!(parent[j - 1] == j && nkids[j] == 1 && counts[j - 1] == counts[j] + 1)) s.first.push_back(j);
  if (n > 0) s.first.push_back(n);
  const i64 ns = s.ns();
  std::vector<i64> snode(n);
  for (i64 q = 0; q < ns; ++q)
    for (i64 j = s.first[q]; j < s.first[q + 1]; ++j) snode[j] = q;
  // supernodal tree and row structures (cholesky.rs:1254-1312)
  s.sparent.assign(ns, NONE);
  for (i64 q = 0; q < ns; ++q) {
    i64 p = parent[s.first[q + 1] - 1];
    s.sparent[q] = p == NONE ? NONE : snode[p];
  }
  std::vector<i64> kp(ns + 1, 0), kids(ns);
  for (i64 q = 0; q < ns; ++q)
    if (s.sparent[q] != NONE) ++kp[s.sparent[q] + 1];
  for (i64 q = 0; q < ns; ++q) kp[q + 1] += kp[q];
  {
    std::vector<i64> at(kp.begin(), kp.end() - 1);
    for (i64 q = 0; q < ns; ++q)
      if (s.sparent[q] != NONE) kids[at[s.sparent[q]]++] = q;
  }
  s.rptr.assign(1, 0);
  s.rows.clear();
  {
    std::vector<i64> mark(n, NONE), set;
    for (i64 q = 0; q < ns; ++q) {
      set.clear();
      auto add = [&](i64 r) {
        if (mark[r] != q) { mark[r] = q; set.push_back(r); }
      };
      for (i64 j = s.first[q]; j < s.first[q + 1]; ++j) {
        add(j);
        for (i64 t = lp[j]; t < lp[j + 1]; ++t) add(li[t]);
      }
      for (i64 t = kp[q]; t < kp[q + 1]; ++t) {
        i64 c = kids[t];
        i64 cw = s.first[c + 1] - s.first[c];
        for (i64 u = s.rptr[c] + cw; u < s.rptr[c + 1]; ++u) add(s.rows[u]);
      }
      std::sort(set.begin(), set.end());
      if ((i64)set.size() != counts[s.first[q]] || set.front() != s.first[q])
        return internal("a supernode's rows differ from its column count");
      s.rows.insert(s.rows.end(), set.begin(), set.end());
      s.rptr.push_back((i64)s.rows.size());
    }
  }
  // updates (cholesky.rs:1314-1330), by target in increasing descendant order
  {
    std::vector<i64> cnt(ns + 1, 0);
    for (int pass = 0; pass < 2; ++pass) {
      if (pass == 1) {
        s.uptr.assign(ns + 1, 0);
        for (i64 q = 0; q < ns; ++q) s.uptr[q + 1] = s.uptr[q] + cnt[q];
        s.upd.resize(s.uptr[ns]);
        std::fill(cnt.begin(), cnt.end(), 0);
      }
      for (i64 k = 0; k < ns; ++k) {
        const i64* kr = &s.rows[s.rptr[k]];
        const i64 km = s.rptr[k + 1] - s.rptr[k];
        i64 p = s.first[k + 1] - s.first[k];
        while (p < km) {
          const i64 target = snode[kr[p]];
          const i64 end = s.first[target + 1];
          i64 q = p;
          while (q < km && kr[q] < end) ++q;
          if (pass == 1) s.upd[s.uptr[target] + cnt[target]] = Update{k, p, q};
          ++cnt[target];
          p = q;
        }
      }
    }
  }
  // subtree tasks and the supernodes above them (cholesky.rs:1332-1375)
  {
    std::vector<i64> work(ns), start(ns);
    for (i64 q = 0; q < ns; ++q) {
      const i64 m = s.rptr[q + 1] - s.rptr[q], w = s.first[q + 1] - s.first[q];
      i64 sum = 0;
      for (i64 c = 0; c < w; ++c) sum += (m - c) * (m - c);
      work[q] = sum;
      start[q] = q;
    }
    for (i64 q = 0; q < ns; ++q) {
      const i64 p = s.sparent[q];
      if (p != NONE) {
        work[p] += work[q];
        start[p] = std::min(start[p], start[q]);
      }
    }
    i64 all = 0;
    for (i64 q = 0; q < ns; ++q)
      if (s.sparent[q] == NONE) all += work[q];
    const i64 share = std::max(all / SUBTREE_SHARE, PARALLEL_WORK);
    for (i64 q = 0; q < ns;) {
      i64 root = NONE;
      for (i64 r = q; r != NONE && start[r] == q && work[r] <= share; r = s.sparent[r]) root = r;
      if (root == NONE) {
        s.top.push_back(q);
        ++q;
      } else {
        s.tasks.push_back(q);
        s.tasks.push_back(root + 1);
        q = root + 1;
      }
    }
  }
  for (i64 q = 0; q < ns; ++q)
    s.flops += trapezoid_flops((double)(s.first[q + 1] - s.first[q]), (double)(s.rptr[q + 1] - s.rptr[q]));

  // ---------------- execution plan ----------------
  // fronts: chains of supernodes, each the last child of the next (postorder puts it just
  // before its parent), merged while the time model favours it
  std::vector<i64> fs0(ns), fw(ns), fm(ns);
  std::vector<char> is_top(ns, 1);
  for (i64 q = 0; q < ns; ++q) {
    const i64 w = s.first[q + 1] - s.first[q], m = s.rptr[q + 1] - s.rptr[q];
    fs0[q] = q;
    fw[q] = w;
    fm[q] = m;
    const i64 c = q - 1;
    if (c >= 0 && s.sparent[c] == q) {
      const i64 wc = fw[c], mc = fm[c];
      const i64 wn = wc + w, mn = wc + m;
      const double added = trapezoid_flops((double)wn, (double)mn) - trapezoid_flops((double)wc, (double)mc) -
                           trapezoid_flops((double)w, (double)m);
      const double moved = (double)(mc - wc) * (double)(mc - wc);
      if ((wn <= opt.merge_small || added <= opt.merge_ratio * moved) && mn <= opt.max_merge_rows) {
        fs0[q] = fs0[c];
        fw[q] = wn;
        fm[q] = mn;
        is_top[c] = 0;
      }
    }
  }
  s.sn_front.assign(ns, NONE);
  for (i64 q = 0; q < ns; ++q)
    if (is_top[q]) {
      const i64 f = s.nf++;
      s.f_sn0.push_back(fs0[q]);
      s.f_sn1.push_back(q + 1);
      s.f_col0.push_back(s.first[fs0[q]]);
      s.f_w.push_back(fw[q]);
      s.f_m.push_back(fm[q]);
      for (i64 r = fs0[q]; r <= q; ++r) s.sn_front[r] = f;
    }
  const i64 nf = s.nf;
  s.f_rptr.assign(1, 0);
  for (i64 f = 0; f < nf; ++f) {
    const i64 top = s.f_sn1[f] - 1;
    const i64 wt = s.first[top + 1] - s.first[top];
    for (i64 c = s.f_col0[f]; c < s.f_col0[f] + s.f_w[f]; ++c) s.f_rows.push_back(c);
    s.f_rows.insert(s.f_rows.end(), s.rows.begin() + s.rptr[top] + wt, s.rows.begin() + s.rptr[top + 1]);
    s.f_rptr.push_back((i64)s.f_rows.size());
    if (s.f_rptr[f + 1] - s.f_rptr[f] != s.f_m[f]) return internal("front rows");
  }
  s.f_parent.assign(nf, NONE);
  for (i64 f = 0; f < nf; ++f) {
    const i64 p = s.sparent[s.f_sn1[f] - 1];
    s.f_parent[f] = p == NONE ? NONE : s.sn_front[p];
  }
  s.f_kptr.assign(nf + 1, 0);
  for (i64 f = 0; f < nf; ++f)
    if (s.f_parent[f] != NONE) ++s.f_kptr[s.f_parent[f] + 1];
  for (i64 f = 0; f < nf; ++f) s.f_kptr[f + 1] += s.f_kptr[f];
  s.f_kids.resize(nf);
  {
    std::vector<i64> at(s.f_kptr.begin(), s.f_kptr.end() - 1);
    for (i64 f = 0; f < nf; ++f)
      if (s.f_parent[f] != NONE) s.f_kids[at[s.f_parent[f]]++] = f;
  }
  // relative maps of each front's rows below into its parent's rows
  s.f_rel_off.assign(nf + 1, 0);
  for (i64 f = 0; f < nf; ++f) s.f_rel_off[f + 1] = s.f_rel_off[f] + (s.f_m[f] - s.f_w[f]);
  s.rel.resize(s.f_rel_off[nf]);
  for (i64 f = 0; f < nf; ++f) {
    const i64 p = s.f_parent[f];
    const i64 below = s.f_m[f] - s.f_w[f];
    if (below == 0) continue;
    if (p == NONE) return internal("a root front with rows below");
    const i64* pr = &s.f_rows[s.f_rptr[p]];
    const i64 pm = s.f_m[p];
    i64 at = 0;
    for (i64 i = 0; i < below; ++i) {
      const i64 r = s.f_rows[s.f_rptr[f] + s.f_w[f] + i];
      while (at < pm && pr[at] < r) ++at;
      if (at >= pm || pr[at] != r) return internal("a front row outside its parent");
      s.rel[s.f_rel_off[f] + i] = (int32_t)at;
    }
  }
  // levels of the front tree
  s.f_level.assign(nf, 0);
  for (i64 f = 0; f < nf; ++f)
    if (s.f_parent[f] != NONE) s.f_level[s.f_parent[f]] = std::max(s.f_level[s.f_parent[f]], s.f_level[f] + 1);
  s.levels = nf ? *std::max_element(s.f_level.begin(), s.f_level.end()) + 1 : 0;
  s.lvl_ptr.assign(s.levels + 1, 0);
  for (i64 f = 0; f < nf; ++f) ++s.lvl_ptr[s.f_level[f] + 1];
  for (i64 l = 0; l < s.levels; ++l) s.lvl_ptr[l + 1] += s.lvl_ptr[l];
  s.lvl_fronts.resize(nf);
  {
    std::vector<i64> at(s.lvl_ptr.begin(), s.lvl_ptr.end() - 1);
    for (i64 f = 0; f < nf; ++f) s.lvl_fronts[at[s.f_level[f]]++] = f;
  }
  // storage: L parts back to back (aligned to 64 entries); update matrices in an arena whose
  // blocks live from their front's level to their parent's
  s.f_lpos.resize(nf);
  s.f_ld.resize(nf);
  s.l_size = 0;
  for (i64 f = 0; f < nf; ++f) {
    s.f_ld[f] = s.f_m[f];
    s.f_lpos[f] = s.l_size;
    s.l_size += round_up(s.f_m[f] * s.f_w[f], 64);
    s.stored += s.f_m[f] * s.f_w[f] - s.f_w[f] * (s.f_w[f] - 1) / 2;
    s.front_flops += trapezoid_flops((double)s.f_w[f], (double)s.f_m[f]);
  }
  s.f_upos.assign(nf, NONE);
  {
    Arena arena;
    auto usize = [&](i64 f) { return round_up((s.f_m[f] - s.f_w[f]) * (s.f_m[f] - s.f_w[f]), 64); };
    for (i64 l = 0; l < s.levels; ++l) {
      for (i64 t = s.lvl_ptr[l]; t < s.lvl_ptr[l + 1]; ++t) {
        const i64 f = s.lvl_fronts[t];
        if (s.f_m[f] > s.f_w[f]) s.f_upos[f] = arena.take(usize(f));
      }
      for (i64 t = s.lvl_ptr[l]; t < s.lvl_ptr[l + 1]; ++t) {
        const i64 f = s.lvl_fronts[t];
        for (i64 u = s.f_kptr[f]; u < s.f_kptr[f + 1]; ++u) {
          const i64 c = s.f_kids[u];
          if (s.f_upos[c] != NONE) arena.give(s.f_upos[c], usize(c));
        }
      }
    }
    s.u_size = arena.end;
  }
  // The solves' per-front vectors (M - W entries a right-hand side) in an arena of their own,
  // with the same level lifetimes: a front's contribution is written at its level and read by
  // its parent (forward); on the way back a level's blocks are written and read within the level.
  // 8-entry blocks keep every front's vector aligned to 64 bytes.
  s.f_soff.assign(nf, 0);
  {
    Arena arena;
    auto ssize = [&](i64 f) { return round_up(s.f_m[f] - s.f_w[f], 8); };
    for (i64 l = 0; l < s.levels; ++l) {
      for (i64 t = s.lvl_ptr[l]; t < s.lvl_ptr[l + 1]; ++t) {
        const i64 f = s.lvl_fronts[t];
        if (s.f_m[f] > s.f_w[f]) s.f_soff[f] = arena.take(ssize(f));
      }
      for (i64 t = s.lvl_ptr[l]; t < s.lvl_ptr[l + 1]; ++t) {
        const i64 f = s.lvl_fronts[t];
        for (i64 u = s.f_kptr[f]; u < s.f_kptr[f + 1]; ++u) {
          const i64 c = s.f_kids[u];
          if (s.f_m[c] > s.f_w[c]) arena.give(s.f_soff[c], ssize(c));
        }
      }
    }
    s.s_size = arena.end;
  }
  // both arenas checked: the blocks live at one level (a front's from its level to its
  // parent's) never overlap. Levels in order: blocks whose lifetime ended leave an ordered map of
  // the live ones, each new block is checked against its neighbours there.
  {
    std::vector<std::vector<i64>> born(s.levels + 1), dead(s.levels + 2);
    for (i64 f = 0; f < nf; ++f) {
      if (s.f_m[f] == s.f_w[f]) continue;
      born[s.f_level[f]].push_back(f);
      dead[(s.f_parent[f] == NONE ? s.levels : s.f_level[s.f_parent[f]]) + 1].push_back(f);
    }
    auto disjoint = [&](const std::vector<i64>& off, bool square) {
      std::map<i64, i64> live;  // offset -> end
      for (i64 l = 0; l < s.levels; ++l) {
        for (i64 f : dead[l]) live.erase(off[f]);
        for (i64 f : born[l]) {
          const i64 mu = s.f_m[f] - s.f_w[f], b = off[f], e = b + (square ? mu * mu : mu);
          auto nx = live.lower_bound(b);
          if (nx != live.end() && nx->first < e) return false;
          if (nx != live.begin() && std::prev(nx)->second > b) return false;
          live.emplace(b, e);
        }
      }
      return true;
    };
    if (!disjoint(s.f_upos, true) || !disjoint(s.f_soff, false))
      return internal("overlapping blocks in a work arena");
  }
  // where each entry of A lands in the L storage
  s.a_pos.resize(nnz);
  {
    std::vector<i64> col_front(n);
    for (i64 j = 0; j < n; ++j) col_front[j] = s.sn_front[snode[j]];
    for (i64 k = 0; k < n; ++k)
      for (i64 t = ptr[k]; t < ptr[k + 1]; ++t) {
        const i64 r = std::max(inv[idx[t]], inv[k]), c = std::min(inv[idx[t]], inv[k]);
        const i64 f = col_front[c];
        const i64* fr = &s.f_rows[s.f_rptr[f]];
        const i64 pos = std::lower_bound(fr, fr + s.f_m[f], r) - fr;
        if (pos >= s.f_m[f] || fr[pos] != r) return internal("an entry of A outside the structure");
        s.a_pos[t] = s.f_lpos[f] + (c - s.f_col0[f]) * s.f_ld[f] + pos;
      }
  }
  return fail;
}

}  // namespace fc
