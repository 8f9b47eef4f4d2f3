// The symbolic phase of the supernodal Cholesky (host), and the device schedule built on it.
//
// Part 1 is the reference's symbolic phase, restated so that its outputs are the same bit for
// bit (crates/solver/src/cholesky.rs:1213-1378): the postorder `perm`, the fundamental
// supernodes `first`, their row structures `rptr`/`rows`, the update lists `updates` (for each
// supernode J, the runs {k, p1, p2} of descendant K's rows that fall in J's columns) and the
// subtree schedule `tasks`/`top` of the CPU numeric phase. All are exported for parity checks.
//
// Part 2 is this library's own execution plan for the numeric phase on one GPU: a multifrontal
// schedule over *fronts*. A front is a chain of consecutive supernodes, each the last child of
// the next, factored as one dense frontal matrix; the supernode partition itself is unchanged
// and exported as is. Rows of a lower member that its own structure lacks are stored as exact
// zeros (they stay zero through the factorization, since every product that reaches them has a
// structurally zero factor). A child front is merged when the flops its zero rows add cost less
// than moving its update matrix through HBM (a time model, see amalgamate()). Fronts are grouped
// in levels of the front tree (leaves at level 0); every kernel launch works on one level.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace fc {

using i64 = int64_t;
constexpr i64 NONE = -1;

struct Update {
  i64 k, p1, p2;
};

// Tunables of the execution plan (not of the result's structure).
struct PlanOptions {
  double merge_ratio = 32.0;  // merge a child front when added flops <= ratio * (its M - W)^2
  i64 merge_small = 16;       // always merge when the merged front is at most this wide
  i64 max_merge_rows = 1 << 16;
};

struct Symbolic {
  // --- reference symbolic phase (bit-exact) ---
  i64 n = 0;
  i64 entries = 0;
  bool by_blocks = false;
  std::vector<i64> perm;          // perm[k]: original column of factor column k
  std::vector<i64> parent0, counts0;  // tree and counts in the input order
  std::vector<i64> first;         // ns + 1
  std::vector<i64> rptr, rows;    // row structures (postorder numbering)
  std::vector<i64> sparent;       // supernodal parent
  std::vector<i64> uptr;          // ns + 1
  std::vector<Update> upd;
  std::vector<i64> tasks;         // pairs [start, end)
  std::vector<i64> top;
  // pattern of the input (the numeric phase requires the same)
  std::vector<i64> upper_ptr, upper_idx;

  // --- execution plan ---
  i64 nf = 0;                       // fronts
  std::vector<i64> sn_front;        // supernode -> front
  std::vector<i64> f_sn0, f_sn1;    // front -> supernodes [sn0, sn1)
  std::vector<i64> f_col0, f_w, f_m;
  std::vector<i64> f_rptr, f_rows;  // front rows (columns first, then rows below), sorted
  std::vector<i64> f_parent, f_level;
  std::vector<i64> f_kptr, f_kids;  // children fronts in increasing order
  std::vector<i64> f_rel_off;       // per front: start of its relative map into its parent
  std::vector<int32_t> rel;         // rows below -> positions in the parent's front rows
  std::vector<i64> f_lpos, f_ld;    // storage of the front's L part (M x W, column-major)
  std::vector<i64> f_upos;          // storage of the update matrix ((M-W) x (M-W), ld M-W), -1 if none
  i64 l_size = 0, u_size = 0;       // elements
  std::vector<i64> f_soff;          // a solve's M-W entries of the front (its contribution up the
                                    // tree, then its rows' solution values) per right-hand side
  i64 s_size = 0;                   // the solve arena per right-hand side (entries)
  i64 levels = 0;
  std::vector<i64> lvl_ptr, lvl_fronts;  // fronts of each level
  std::vector<i64> a_pos;           // per upper entry: position in the L storage
  i64 stored = 0;                   // stored L entries, padding included
  double flops = 0.0;               // sum_s sum_{c<w} (m - c)^2, the reference's measure
  double front_flops = 0.0;         // the same over fronts (padding included)

  i64 ns() const { return (i64)first.size() - 1; }
};

// Status of an analysis or factorization, mirroring FactorError (cholesky.rs:346-361).
enum Status : int {
  OK = 0,
  NOT_POSITIVE_DEFINITE = 1,
  SINGULAR = 2,
  TOO_LARGE = 3,
  OUT_OF_MEMORY = 4,
  INTERNAL = 5,
  CANCELLED = 6,
  INVALID_ARGUMENT = 7,
  DEVICE = 8,
};

struct Failure {
  Status status = OK;
  i64 column = -1, entries = 0, limit = 0;
  std::string message;
};

// Symbolic::new (cholesky.rs:1213-1378) plus the execution plan. Refuses with TOO_LARGE before
// any numeric memory is planned when the factor would exceed max_entries.
Failure analyze(i64 n, const i64* ptr, const i64* idx, i64 max_entries, const PlanOptions& opt, Symbolic& s);

}  // namespace fc
