// The C ABI (include/fc_chol.h): host symbolic phase, device plan, numeric factorization by
// levels of the front tree, solves and refinement on one GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "factor_kernels.hpp"
#include "fc_chol.h"
#include "gemm_kernels.hpp"
#include "lanczos_kernels.hpp"
#include "ldlt_kernels.hpp"
#include "small_kernels.hpp"
#include "solve_kernels.hpp"
#include "symbolic.hpp"

using fc::i64;

namespace {

constexpr const char* VERSION = "fieldcraft_amd 0.1 (gfx950)";

struct DeviceError {
  hipError_t e;
  const char* what;
};

#define HIP_CHECK(x)                                 \
  do {                                               \
    hipError_t e_ = (x);                             \
    if (e_ != hipSuccess) throw DeviceError{e_, #x}; \
  } while (0)

// The largest single device allocation the library makes (FC_MAX_DEVICE_ALLOC bytes; unset or
// 0: no limit): a caller can keep the library's buffers below what other work on the GPU needs,
// and a test can make any allocation fail. An allocation above it fails as a short device would.
size_t max_device_alloc() {
  const char* e = std::getenv("FC_MAX_DEVICE_ALLOC");
  if (!e) return 0;
  return (size_t)std::strtoull(e, nullptr, 10);
}

// FC_POISON_ALLOC=1: every device allocation starts filled with 0xFF bytes (a NaN in FP64 and
// FP32), so that a kernel reading memory no step of the current call has written shows up as a
// NaN in the results instead of whatever the allocator hands back. FC_POISON_ALLOC=2 also puts
// POISON_GUARD bytes of the same filling before and after every allocation, so that a read just
// outside a buffer shows up the same way. A test facility; off unless the variable is set.
int poison_alloc() {
  const char* e = std::getenv("FC_POISON_ALLOC");
  return e && (e[0] == '1' || e[0] == '2') ? e[0] - '0' : 0;
}
constexpr size_t POISON_GUARD = (size_t)4 << 20;
std::mutex g_guard_mu;
std::vector<void*>* g_guarded = new std::vector<void*>();  // the guarded allocations' user pointers

template <typename T>
T* dalloc(size_t count) {
  if (count == 0) count = 1;
  const size_t cap = max_device_alloc();
  if (cap && count * sizeof(T) > cap) throw std::bad_alloc();
  const int poison = poison_alloc();
  const size_t guard = poison == 2 ? POISON_GUARD : 0;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T) + 2 * guard);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw std::bad_alloc();
  }
  if (poison) {
    // the null stream does not order the library's non-blocking streams: wait for the fill
    if (hipMemset(p, 0xFF, count * sizeof(T) + 2 * guard) != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
      (void)hipFree(p);
      throw DeviceError{hipErrorUnknown, "poisoning a new allocation"};
    }
  }
  if (guard) {
    p = (char*)p + guard;
    std::lock_guard<std::mutex> g(g_guard_mu);
    g_guarded->push_back(p);
  }
  return (T*)p;
}

template <typename T>
T* dupload(const std::vector<T>& v, hipStream_t st) {
  T* p = dalloc<T>(v.size());
  if (!v.empty()) HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
  return p;
}

void dfree(void* p) {
  if (!p) return;
  {
    std::lock_guard<std::mutex> g(g_guard_mu);
    auto it = std::find(g_guarded->begin(), g_guarded->end(), p);
    if (it != g_guarded->end()) {
      g_guarded->erase(it);
      p = (char*)p - POISON_GUARD;
    }
  }
  (void)hipFree(p);
}

// A device buffer that grows on demand and is kept (workspaces reused across calls).
struct Buffer {
  void* p = nullptr;
  size_t bytes = 0;
  void* get(size_t need) {
    if (need == 0) need = 1;
    if (need > bytes) {
      dfree(p);
      p = nullptr;
      bytes = 0;
      p = (void*)dalloc<char>(need);
      bytes = need;
    }
    return p;
  }
  void release() {
    dfree(p);
    p = nullptr;
    bytes = 0;
  }
};

void set_error(fc_error* err, int kind, const std::string& msg, uint64_t column = 0, uint64_t entries = 0,
               uint64_t limit = 0) {
  if (!err) return;
  err->kind = kind;
  err->pad = 0;
  err->column = column;
  err->entries = entries;
  err->limit = limit;
  std::snprintf(err->message, sizeof err->message, "%s", msg.c_str());
}

// The factorization's schedule: how many panels the side stream updates together and how far
// GEMM launches are split along k. The defaults are measured at hex8 80^3; the tuner
// (fc_tune) times candidates on the caller's matrix and keeps the fastest.
struct Schedule {
  int32_t slab = 4;                // panels of a side-stream slab update (1: one-panel lookahead)
  int32_t split_target = 1024;     // tiles the Schur-complement launches are split up to
  // The side stream's slab updates only have to keep ahead of the main stream's panel chain:
  // they are split less, so they leave CUs to the chain's kernels and need fewer partial sums.
  int32_t split_side_fp64 = 384;
  int32_t split_side_fp32 = 256;
};

// The launches of the numeric phase and of the solves, built once per analysis.
struct Program {
  std::vector<fc::Item> items;         // extend-add and diagonal-block items
  std::vector<fc::GemmItem> gitems;    // GEMM tiles (UPDATE, TRSM, SYRK launches)
  std::vector<fc::ReduceItem> ritems;  // split-K reductions
  std::vector<fc::Launch> launches;
  std::vector<fc::EaRange> earanges;   // extend-add items' child column runs (item.c, item.d)
  std::vector<i64> f_inv0;             // per front: its diagonal-block inverses in the Inv buffer
  std::vector<i64> f_bpos;             // per front: its L part in the BF16 copies (mixed form)
  i64 b_size = 0;                      // entries of each BF16 copy
  i64 inv_size = 0;                    // entries of the Inv buffer (kept with the factor)
  i64 part_size = 0;                   // entries of the split-K workspace
  i64 w_size = 0;                      // entries of the FP64 TRSM workspace
  std::vector<fc::Item> sitems;       // solve work items
  std::vector<fc::Launch> sfwd, sbwd; // forward launches (leaves first), backward (root first)
  std::vector<fc::Launch> sfwd1, sbwd1;  // the same for one right-hand side: each level's block
                                         // chain in one launch (k_fwd_front / k_bwd_front)

  std::vector<i64> uoff;              // per front: start of its solve contribution
  i64 u_solve = 0;                    // total contribution length (one right-hand side)
  double work[fc::KINDS][2] = {};     // per kind: flops, entries
  Schedule sched;                     // the schedule the launches were built with
};

// TRSM launches (Launch.pad): the in-place product run by the FP32 and mixed factors, and the
// FP64 factor's saved-copy, product, residual and correction steps
enum TrsmVariant : int32_t { TRSM_IN_PLACE = 0, TRSM_COPY = 1, TRSM_PRODUCT = 2, TRSM_RESIDUAL = 3, TRSM_FIX = 4 };

// Side-stream panel updates (Launch.pad of their UPDATE and REDUCE launches): one variant for the
// FP64 factor, one for the FP32 and mixed ones; each factor runs its own.
enum SideVariant : int32_t { SIDE_ANY = 0, SIDE_FP64 = 1, SIDE_FP32 = 2 };

// Ordering steps between the caller's stream and the factor's side stream (Launch.kind values
// outside the launch kinds): record event slot `slot` on stream `stream`, or make `stream` wait
// for that slot's last record.
constexpr int32_t L_RECORD = 100;
constexpr int32_t L_WAIT = 101;
constexpr int SYNC_SLOTS = 4;

// kinds of solve launches
enum SolveKind : int32_t {
  S_GATHER = 0,
  S_FWD_STEP = 1,
  S_BWD_GATHER = 2,
  S_BWD_BELOW = 3,
  S_BWD_FIRST = 4,
  S_BWD_STEP = 5,
  S_FWD_FRONT = 6,
  S_BWD_FRONT = 7
};
static_assert(fc::SNB == fc::NB, "solve blocks are the factorization's panels");
static_assert(fc::BCOLS == fc::SNB, "a backward column tile is one block");

// lower pairs (i >= j) of rows [r0, r1) x columns [c0, c1)
double lower_pairs(i64 r0, i64 r1, i64 c0, i64 c1) {
  double s = 0;
  for (i64 j = c0; j < c1; ++j) s += (double)std::max<i64>(0, r1 - std::max(r0, j));
  return s;
}

// An output tile of a GEMM launch with its algorithmic flops.
struct Tile {
  fc::GemmItem g;
  double flops;
};

// Output tiles a launch should have to fill the GPU (256 CUs, two workgroups each, and some
// slack); a launch with fewer is split along k into pieces of at least SPLIT_MIN_K.
constexpr i64 SPLIT_TARGET = 1024;
constexpr i64 SPLIT_MIN_K = 512;

// One GEMM launch of `tiles` (kind: UPDATE, TRSM or SYRK), split along k when the tiles are
// too few, followed by the launch that reduces the pieces.
void add_gemm(Program& p, const std::vector<Tile>& tiles, int kind, i64 level, i64 split_target = SPLIT_TARGET) {
  if (tiles.empty()) return;
  const i64 nt = (i64)tiles.size();
  const i64 g0 = (i64)p.gitems.size(), r0 = (i64)p.ritems.size();
  double flops = 0;
  i64 part = 0;
  for (const Tile& t : tiles) {
    flops += t.flops;
    i64 split = 1;
    if (nt < split_target && t.g.k >= 2 * SPLIT_MIN_K)
      split = std::min<i64>((split_target + nt - 1) / nt, t.g.k / SPLIT_MIN_K);
    if (split <= 1) {
      p.gitems.push_back(t.g);
      continue;
    }
    i64 chunk = (t.g.k + split - 1) / split;
    // whole stages of either kernel: every piece of an UPDATE or SYRK product (which all start
    // at column 0 of their front) then starts on a chunk of the BF16 copies
    chunk = (chunk + fc::BK - 1) / fc::BK * fc::BK;
    split = (t.g.k + chunk - 1) / chunk;
    fc::ReduceItem r{};
    r.c = t.g.c;
    r.part0 = part;
    r.ldc = t.g.ldc;
    r.m = t.g.m;
    r.n = t.g.n;
    r.mode = t.g.mode;
    r.diag = t.g.diag;
    r.parts = (int32_t)split;
    r.csrc = t.g.csrc;
    p.ritems.push_back(r);
    for (i64 q = 0; q < split; ++q) {
      fc::GemmItem g = t.g;
      g.a += q * chunk * (i64)t.g.lda;
      g.b += q * chunk * (i64)t.g.ldb;
      g.k = (int32_t)std::min<i64>(chunk, t.g.k - q * chunk);
      g.mode = fc::G_PART;
      g.c = part + q * (i64)fc::GB * fc::GB;
      g.csrc = fc::S_PART;
      p.gitems.push_back(g);
    }
    part += split * (i64)fc::GB * fc::GB;
  }
  p.launches.push_back(fc::Launch{kind, (int32_t)level, g0, (int32_t)((i64)p.gitems.size() - g0), 0, flops, 0.0});
  p.work[kind][0] += flops;
  if ((i64)p.ritems.size() > r0)
    p.launches.push_back(fc::Launch{fc::K_REDUCE, (int32_t)level, r0, (int32_t)((i64)p.ritems.size() - r0), 0, 0.0, 0.0});
  p.part_size = std::max(p.part_size, part);
}

// Fronts whose dense steps one workgroup of k_small_fronts takes: at most SMALL_W columns (one
// wave's diagonal block), and an update matrix of at most SMALL_SYRK multiply-adds (the part
// that sets a front's time there: a front above it is faster on the tiled launches).
bool is_small(const fc::Symbolic& s, i64 f) {
  const double w = (double)s.f_w[f], mu = (double)(s.f_m[f] - s.f_w[f]);
  return s.f_w[f] <= fc::SMALL_W && s.f_m[f] - s.f_w[f] <= fc::SMALL_MU && 0.5 * mu * mu * w <= fc::SMALL_SYRK;
}

Program build_program(const fc::Symbolic& s, const Schedule& sched) {
  using fc::Item;
  using fc::Launch;
  // The diagonal-block kernels read a panel's columns through one buffer resource, at byte
  // offsets below 2^31: a front of 2^21 rows or more is refused (the same for every schedule,
  // so only the analysis's first program can meet it).
  for (i64 f = 0; f < s.nf; ++f)
    if (s.f_ld[f] >= (i64(1) << 21))
      throw fc::Failure{fc::TOO_LARGE, -1, s.f_ld[f], (i64(1) << 21) - 1, "a front of 2^21 rows or more"};
  Program p;
  p.sched = sched;
  double flops = 0, entries = 0;
  auto add = [&](int kind, i64 level, i64 item0) {
    const i64 count = (i64)p.items.size() - item0;
    if (count > 0) {
      p.launches.push_back(Launch{kind, (int32_t)level, item0, (int32_t)count, 0, flops, entries});
      p.work[kind][0] += flops;
      p.work[kind][1] += entries;
    }
    flops = entries = 0;
  };
  // the BF16 copies of the mixed form: each front's L part by chunks of BK columns
  // (none for small fronts: no product of the tiled launches reads them)
  p.f_bpos.resize(s.nf);
  for (i64 f = 0; f < s.nf; ++f) {
    p.f_bpos[f] = p.b_size;
    if (!is_small(s, f))
      p.b_size += (s.f_ld[f] * ((s.f_w[f] + fc::BK - 1) / fc::BK) * fc::BK + 63) / 64 * 64;
  }
  // the inverses of every front's diagonal blocks (NB x NB, the last one pw x pw), kept for
  // the solves
  p.f_inv0.resize(s.nf);
  for (i64 f = 0; f < s.nf; ++f) {
    p.f_inv0[f] = p.inv_size;
    const i64 full = s.f_w[f] / fc::NB, last = s.f_w[f] % fc::NB;
    p.inv_size += full * fc::NB * fc::NB + last * last;
  }
  std::vector<double> colent;  // child entries landing in each parent column
  for (i64 l = 0; l < s.levels; ++l) {
    // the level's fronts (every one takes the extend-add launches), its small fronts
    // (k_small_fronts) and the others (the tiled launches)
    const i64* fall = &s.lvl_fronts[s.lvl_ptr[l]];
    const i64 nall = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    std::vector<i64> reg, sm;
    for (i64 t = 0; t < nall; ++t) (is_small(s, fall[t]) ? sm : reg).push_back(fall[t]);
    const i64* fl = reg.data();
    const i64 nl = (i64)reg.size();
    auto has_kids_below = [&](i64 f) {
      for (i64 q = s.f_kptr[f]; q < s.f_kptr[f + 1]; ++q)
        if (s.f_m[s.f_kids[q]] > s.f_w[s.f_kids[q]]) return true;
      return false;
    };
    auto column_entries = [&](i64 f) {
      colent.assign(s.f_m[f], 0.0);
      for (i64 q = s.f_kptr[f]; q < s.f_kptr[f + 1]; ++q) {
        const i64 c = s.f_kids[q], cmu = s.f_m[c] - s.f_w[c];
        for (i64 jc = 0; jc < cmu; ++jc) colent[s.rel[s.f_rel_off[c] + jc]] += (double)(cmu - jc);
      }
    };
    // Extend-add items: windows of EA_COLS parent columns, each with the runs of its children's
    // columns that land there (rel rises, so each child's is one run), children in increasing
    // order. Consecutive windows of one front share an item until it holds EA_ENTRIES child
    // entries, so that small fronts still give a workgroup enough to move.
    i64 ea_r0 = -1, ea_n = 0;
    int32_t ea_f = -1, ea_a = 0, ea_b = 0;
    auto ea_close = [&]() {
      if (ea_r0 >= 0 && (i64)p.earanges.size() > ea_r0)
        p.items.push_back(Item{ea_f, ea_a, ea_b, (int32_t)ea_r0, (int32_t)((i64)p.earanges.size() - ea_r0), 0, 0, 0});
      ea_r0 = -1;
      ea_n = 0;
    };
    auto ea_window = [&](i64 f, i64 a, i64 b) {
      if (ea_r0 >= 0 && (ea_f != (int32_t)f || ea_n >= fc::EA_ENTRIES)) ea_close();
      if (ea_r0 < 0) {
        ea_r0 = (i64)p.earanges.size();
        ea_f = (int32_t)f;
        ea_a = (int32_t)a;
      }
      ea_b = (int32_t)b;
      for (i64 q = s.f_kptr[f]; q < s.f_kptr[f + 1]; ++q) {
        const i64 c = s.f_kids[q], cmu = s.f_m[c] - s.f_w[c];
        const int32_t* rc = s.rel.data() + s.f_rel_off[c];
        const i64 jc0 = std::lower_bound(rc, rc + cmu, (int32_t)a) - rc;
        const i64 jc1 = std::lower_bound(rc + jc0, rc + cmu, (int32_t)b) - rc;
        if (jc1 > jc0) p.earanges.push_back(fc::EaRange{(int32_t)c, (int32_t)jc0, (int32_t)jc1, 0});
        for (i64 jc = jc0; jc < jc1; ++jc) ea_n += cmu - jc;
      }
    };
    // extend-add into the L parts
    i64 i0 = (i64)p.items.size();
    for (i64 t = 0; t < nall; ++t) {
      const i64 f = fall[t];
      if (!has_kids_below(f)) continue;
      column_entries(f);
      for (i64 a = 0; a < s.f_w[f]; a += fc::EA_COLS) {
        const i64 b = std::min(a + fc::EA_COLS, s.f_w[f]);
        ea_window(f, a, b);
        for (i64 j = a; j < b; ++j) entries += colent[j];
      }
    }
    ea_close();
    add(fc::K_EXTEND_L, l, i0);
    // the small fronts' diagonal blocks, rows below and update matrices, one item (workgroup)
    // each, costliest first (the first workgroups to start take them)
    auto small_cost = [&](i64 f) {
      const double w = (double)s.f_w[f], mu = (double)(s.f_m[f] - s.f_w[f]);
      return w * w * w / 3 + mu * w * w + mu * mu * w;
    };
    std::stable_sort(sm.begin(), sm.end(), [&](i64 x, i64 y) { return small_cost(x) > small_cost(y); });
    i0 = (i64)p.items.size();
    for (const i64 f : sm) {
      const i64 w = s.f_w[f], mu = s.f_m[f] - w;
      p.items.push_back(Item{(int32_t)f, 0, 0, 0, 0, 0, 0, 0});
      // as the tiled launches count them
      flops += (double)w * (w + 1) * (2 * w + 1) / 6.0 + (double)mu * w * w + 2.0 * w * lower_pairs(0, mu, 0, mu);
    }
    add(fc::K_SMALL, l, i0);
    i64 wmax = 0;
    for (i64 t = 0; t < nl; ++t) wmax = std::max(wmax, s.f_w[fl[t]]);
    std::vector<Tile> tiles;
    // The columns [c0, c1) (whole panels) minus their products with the front's columns
    // [k0, k1) (all before c0): one tile column per panel, from the panel's diagonal down.
    auto update = [&](i64 c0, i64 c1, i64 k0, i64 k1, i64 split_target) {
      tiles.clear();
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t];
        for (i64 b0 = c0; b0 < std::min(c1, s.f_w[f]); b0 += fc::NB) {
          const i64 b1 = std::min(b0 + fc::NB, s.f_w[f]);
          for (i64 r = b0; r < s.f_m[f]; r += fc::GB) {
            const i64 m = std::min<i64>(fc::GB, s.f_m[f] - r);
            fc::GemmItem g{};
            g.a = s.f_lpos[f] + k0 * s.f_ld[f] + r;
            g.b = s.f_lpos[f] + k0 * s.f_ld[f] + b0;
            g.c = s.f_lpos[f] + b0 * s.f_ld[f] + r;
            g.lda = g.ldb = g.ldc = (int32_t)s.f_ld[f];
            g.m = (int32_t)m;
            g.n = (int32_t)(b1 - b0);
            g.k = (int32_t)(k1 - k0);
            g.mode = fc::G_SUB;
            g.diag = (int32_t)(r - b0);
            g.asrc = g.bsrc = g.csrc = fc::S_L;
            g.front = (int32_t)f;
            tiles.push_back(Tile{g, 2.0 * (double)(k1 - k0) * lower_pairs(r, r + m, b0, b1)});
          }
        }
      }
      add_gemm(p, tiles, fc::K_UPDATE, l, split_target);
    };
    // The panel at c0 once its updates are in: diagonal block, rows below, BF16 copies.
    auto chain = [&](i64 c0) {
      // the diagonal blocks (k_potrf_blk, any panel width)
      i0 = (i64)p.items.size();
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t];
        if (s.f_w[f] <= c0) continue;
        const i64 pw = std::min<i64>(fc::NB, s.f_w[f] - c0);
        p.items.push_back(Item{(int32_t)f, (int32_t)c0, 0, 0, 0, 0, 0, 0});
        flops += (double)pw * (pw + 1) * (2 * pw + 1) / 6.0;  // sum_{t <= pw} t^2
      }
      add(fc::K_POTRF, l, i0);
      // the rows below the diagonal block: X = B Linv^T, in place
      tiles.clear();
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t];
        if (s.f_w[f] <= c0) continue;
        const i64 c1 = std::min(c0 + fc::NB, s.f_w[f]);
        const i64 pw = c1 - c0;
        for (i64 r = c1; r < s.f_m[f]; r += fc::GB) {
          const i64 m = std::min<i64>(fc::GB, s.f_m[f] - r);
          fc::GemmItem g{};
          g.a = s.f_lpos[f] + c0 * s.f_ld[f] + r;
          g.b = p.f_inv0[f] + (c0 / fc::NB) * (i64)fc::NB * fc::NB;
          g.c = g.a;
          g.lda = g.ldc = (int32_t)s.f_ld[f];
          g.ldb = (int32_t)pw;
          g.m = (int32_t)m;
          g.n = g.k = (int32_t)pw;
          g.mode = fc::G_SET;
          g.diag = fc::GB;
          g.asrc = g.csrc = fc::S_L;
          g.bsrc = fc::S_INV;
          g.front = (int32_t)f;
          tiles.push_back(Tile{g, (double)m * (double)pw * (double)pw});
        }
      }
      add_gemm(p, tiles, fc::K_TRSM, l);  // in place (FP32 and mixed factors)
      // The FP64 factor refines the product once, which makes it as accurate as substitution
      // (a product with an explicit inverse alone is not backward stable): B is saved in the
      // workspace W (one GB x GB slot per tile), X = W Linv^T, W -= X L11^T, X += W Linv^T.
      if (!tiles.empty()) {
        std::vector<Tile> copy, prod, res, fix;
        for (size_t q = 0; q < tiles.size(); ++q) {
          const fc::GemmItem& t = tiles[q].g;
          const i64 f = t.front, w0 = (i64)q * fc::GB * fc::GB;
          fc::GemmItem g = t;  // W <- B
          g.c = w0;
          g.ldc = fc::GB;
          g.csrc = fc::S_W;
          copy.push_back(Tile{g, 0.0});
          g = t;  // X = W Linv^T
          g.a = w0;
          g.lda = fc::GB;
          g.asrc = fc::S_W;
          prod.push_back(Tile{g, tiles[q].flops});
          g = t;  // W -= X L11^T
          g.a = t.c;
          g.b = s.f_lpos[f] + (i64)((t.c - s.f_lpos[f]) / s.f_ld[f]) * s.f_ld[f] + (t.c - s.f_lpos[f]) / s.f_ld[f];
          g.ldb = (int32_t)s.f_ld[f];
          g.bsrc = fc::S_L;
          g.c = w0;
          g.ldc = fc::GB;
          g.csrc = fc::S_W;
          g.mode = fc::G_SUB;
          res.push_back(Tile{g, 0.0});
          g = t;  // X += W Linv^T
          g.a = w0;
          g.lda = fc::GB;
          g.asrc = fc::S_W;
          g.mode = fc::G_ADD;
          fix.push_back(Tile{g, 0.0});
        }
        p.w_size = std::max<i64>(p.w_size, (i64)tiles.size() * fc::GB * fc::GB);
        const double before = p.work[fc::K_TRSM][0];
        int variant = TRSM_COPY;
        for (auto* list : {&copy, &prod, &res, &fix}) {
          add_gemm(p, *list, fc::K_TRSM, l);
          p.launches.back().pad = variant++;
        }
        p.work[fc::K_TRSM][0] = before;  // the algorithmic flops are the in-place product's
      }
    };
    auto sync = [&](int kind, int stream, int slot) {
      Launch q{kind, (int32_t)l, 0, 0, 0, 0.0, 0.0};
      q.stream = stream;
      q.slot = slot;
      p.launches.push_back(q);
    };
    // Panels of NB columns, left-looking inside each front (cholesky.rs:1138-1186), grouped in
    // slabs of S panels (SW = S NB columns) with one slab of lookahead: the side stream takes
    // slab j + 1's products with the columns before slab j (one product SW columns wide) while
    // the main stream factors slab j; the main stream then subtracts slab j's product from
    // slab j + 1 and factors its panels, each after the products with the slab's earlier
    // panels. The side update starts only once the main stream's own product with slab j - 1 is
    // done (that product follows slab j - 1's panels), so the two never compete for the CUs.
    // Events: slot j % 2 marks "the main stream's product of slab j done", slot 2 + j % 2 "slab
    // j's side update done"; each is waited on before it is recorded again. S = 1 is the plain
    // one-panel lookahead.
    const i64 np = (wmax + fc::NB - 1) / fc::NB;
    const i64 S = std::max<i64>(1, sched.slab), SW = S * fc::NB, nsl = (np + S - 1) / S;
    auto side = [&](i64 j) {  // slab j's products with the columns [0, (j - 1) SW)
      sync(L_WAIT, 1, (int)((j - 1) % 2));
      const double before = p.work[fc::K_UPDATE][0];
      for (const int variant : {SIDE_FP64, SIDE_FP32}) {
        const i64 from = (i64)p.launches.size();
        update(j * SW, (j + 1) * SW, 0, (j - 1) * SW, variant == SIDE_FP64 ? sched.split_side_fp64 : sched.split_side_fp32);
        for (i64 q = from; q < (i64)p.launches.size(); ++q) {
          p.launches[q].stream = 1;
          p.launches[q].pad = variant;
        }
        if (variant == SIDE_FP64) p.work[fc::K_UPDATE][0] = before;  // count the products once
      }
      sync(L_RECORD, 1, 2 + (int)(j % 2));
    };
    for (i64 j = 0; j < nsl; ++j) {
      const i64 s0 = j * SW;
      if (j >= 2) sync(L_WAIT, 0, 2 + (int)(j % 2));
      if (j >= 1) {
        update(s0, s0 + SW, s0 - SW, s0, 0);  // the product with the slab just finished
        if (j + 1 < nsl) {
          sync(L_RECORD, 0, (int)(j % 2));
          side(j + 1);
        }
      }
      for (i64 k = j * S; k < std::min((j + 1) * S, np); ++k) {
        const i64 c0 = k * fc::NB;
        if (c0 > s0) update(c0, c0 + fc::NB, s0, c0, 0);  // the slab's earlier panels
        chain(c0);
      }
    }
    // update matrices U = -L21 L21^T
    tiles.clear();
    for (i64 t = 0; t < nl; ++t) {
      const i64 f = fl[t];
      const i64 w = s.f_w[f], mu = s.f_m[f] - w;
      for (i64 a = 0; a < mu; a += fc::GB)
        for (i64 b = 0; b <= a; b += fc::GB) {
          const i64 m = std::min<i64>(fc::GB, mu - a), nn = std::min<i64>(fc::GB, mu - b);
          fc::GemmItem g{};
          g.a = s.f_lpos[f] + w + a;
          g.b = s.f_lpos[f] + w + b;
          g.c = s.f_upos[f] + b * mu + a;
          g.lda = g.ldb = (int32_t)s.f_ld[f];
          g.ldc = (int32_t)mu;
          g.m = (int32_t)m;
          g.n = (int32_t)nn;
          g.k = (int32_t)w;
          g.mode = fc::G_NEG;
          g.diag = (int32_t)(a - b);
          g.asrc = g.bsrc = fc::S_L;
          g.csrc = fc::S_U;
          g.front = (int32_t)f;
          tiles.push_back(Tile{g, 2.0 * (double)w * lower_pairs(a, a + m, b, b + nn)});
        }
    }
    add_gemm(p, tiles, fc::K_SYRK, l, sched.split_target);
    i0 = (i64)p.items.size();
    for (i64 t = 0; t < nall; ++t) {
      const i64 f = fall[t];
      if (!has_kids_below(f)) continue;
      column_entries(f);
      for (i64 a = s.f_w[f]; a < s.f_m[f]; a += fc::EA_COLS) {
        const i64 b = std::min(a + fc::EA_COLS, s.f_m[f]);
        ea_window(f, a, b);
        for (i64 j = a; j < b; ++j) entries += colent[j];
      }
    }
    ea_close();
    add(fc::K_EXTEND_U, l, i0);
  }
  // solves: per level, forward leaves first, backward root first
  auto sadd = [&](std::vector<Launch>& list, int kind, i64 level, i64 item0) {
    const i64 count = (i64)p.sitems.size() - item0;
    if (count > 0) list.push_back(Launch{kind, (int32_t)level, item0, (int32_t)count, 0, 0.0, 0.0});
  };
  for (i64 l = 0; l < s.levels; ++l) {
    const i64* fl = &s.lvl_fronts[s.lvl_ptr[l]];
    const i64 nl = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    i64 wmax = 0;
    for (i64 t = 0; t < nl; ++t) wmax = std::max(wmax, s.f_w[fl[t]]);
    i64 i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t) p.sitems.push_back(Item{(int32_t)fl[t], 0, 0, 0, 0, 0, 0, 0});
    sadd(p.sfwd, S_GATHER, l, i0);
    for (i64 c0 = 0; c0 < wmax; c0 += fc::SNB) {
      i0 = (i64)p.sitems.size();
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t];
        if (s.f_w[f] <= c0) continue;
        const i64 c1 = std::min(c0 + fc::SNB, s.f_w[f]);
        // the first row tile also forms the next block's solution
        for (i64 r = c1; r < s.f_m[f]; r += fc::SROWS_M)
          p.sitems.push_back(Item{(int32_t)f, (int32_t)c0, (int32_t)r, (r == c1 && c1 < s.f_w[f]) ? 1 : 0, 0, 0, 0, 0});
      }
      sadd(p.sfwd, S_FWD_STEP, l, i0);
    }
  }
  for (i64 l = s.levels - 1; l >= 0; --l) {
    const i64* fl = &s.lvl_fronts[s.lvl_ptr[l]];
    const i64 nl = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    i64 wmax = 0;
    i64 i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t)
      if (s.f_m[fl[t]] > s.f_w[fl[t]]) p.sitems.push_back(Item{(int32_t)fl[t], 0, 0, 0, 0, 0, 0, 0});
    sadd(p.sbwd, S_BWD_GATHER, l, i0);
    i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t) {
      const i64 f = fl[t];
      wmax = std::max(wmax, s.f_w[f]);
      if (s.f_m[f] == s.f_w[f]) continue;
      for (i64 j = 0; j < s.f_w[f]; j += fc::SCOLS) p.sitems.push_back(Item{(int32_t)f, (int32_t)j, 0, 0, 0, 0, 0, 0});
    }
    sadd(p.sbwd, S_BWD_BELOW, l, i0);
    i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t) p.sitems.push_back(Item{(int32_t)fl[t], 0, 0, 0, 0, 0, 0, 0});
    sadd(p.sbwd, S_BWD_FIRST, l, i0);
    const i64 nblocks = (wmax + fc::SNB - 1) / fc::SNB;
    for (i64 b = nblocks - 1; b >= 1; --b) {
      const i64 c0 = b * fc::SNB;
      i0 = (i64)p.sitems.size();
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t];
        if (s.f_w[f] <= c0) continue;
        // the last column tile (the previous block's columns) also forms that block's solution
        for (i64 j = 0; j < c0; j += fc::BCOLS)
          p.sitems.push_back(Item{(int32_t)f, (int32_t)c0, (int32_t)j, j == c0 - fc::BCOLS ? 1 : 0, 0, 0, 0, 0});
      }
      sadd(p.sbwd, S_BWD_STEP, l, i0);
    }
  }
  // one right-hand side: per level the gather, then the whole block chain in one launch; the
  // backward's gather, below and first steps, then the chain. A front's forward workgroups are
  // in increasing rows and its backward ones in decreasing blocks: each waits only on earlier ones.
  for (i64 l = 0; l < s.levels; ++l) {
    const i64* fl = &s.lvl_fronts[s.lvl_ptr[l]];
    const i64 nl = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    i64 i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t) p.sitems.push_back(Item{(int32_t)fl[t], 0, 0, 0, 0, 0, 0, 0});
    sadd(p.sfwd1, S_GATHER, l, i0);
    i0 = (i64)p.sitems.size();
    bool chains = false;  // a front of several blocks: the launch forms block solutions
    i64 tiles = 0;
    for (i64 t = 0; t < nl; ++t) {
      chains |= s.f_w[fl[t]] > fc::SNB;
      tiles = std::max<i64>(tiles, (s.f_m[fl[t]] + fc::FROWS - 1) / fc::FROWS);
    }
    if (!chains) {  // one block a front: the step launch of the many-right-hand-side lists
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t], c1 = s.f_w[f];
        for (i64 r = c1; r < s.f_m[f]; r += fc::SROWS)
          p.sitems.push_back(Item{(int32_t)f, 0, (int32_t)r, 0, 0, 0, 0, 0});
      }
      sadd(p.sfwd1, S_FWD_STEP, l, i0);
      continue;
    }
    // tile-major across the level's fronts, the tiles of the fronts' own rows (the chains) first:
    // every chain starts in the first workgroups, and the tiles below, which only follow the
    // chains, never hold a slot a chain tile waits for
    for (int below = 0; below < 2; ++below)
      for (i64 k = 0; k < tiles; ++k)
        for (i64 t = 0; t < nl; ++t) {
          const i64 f = fl[t], c1 = std::min<i64>(fc::SNB, s.f_w[f]), r0 = k * fc::FROWS;
          if ((r0 >= s.f_w[f]) != (below == 1)) continue;
          if (r0 < s.f_m[f] && c1 < s.f_m[f] && r0 + fc::FROWS - 1 >= c1)  // a row of the tile below the first block
            p.sitems.push_back(Item{(int32_t)f, (int32_t)r0, 0, 0, 0, 0, 0, 0});
        }
    sadd(p.sfwd1, S_FWD_FRONT, l, i0);
  }
  for (i64 l = s.levels - 1; l >= 0; --l) {
    const i64* fl = &s.lvl_fronts[s.lvl_ptr[l]];
    const i64 nl = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    i64 i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t)
      if (s.f_m[fl[t]] > s.f_w[fl[t]]) p.sitems.push_back(Item{(int32_t)fl[t], 0, 0, 0, 0, 0, 0, 0});
    sadd(p.sbwd1, S_BWD_GATHER, l, i0);
    i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t) {
      const i64 f = fl[t];
      if (s.f_m[f] == s.f_w[f]) continue;
      for (i64 j = 0; j < s.f_w[f]; j += fc::SCOLS) p.sitems.push_back(Item{(int32_t)f, (int32_t)j, 0, 0, 0, 0, 0, 0});
    }
    sadd(p.sbwd1, S_BWD_BELOW, l, i0);
    i0 = (i64)p.sitems.size();
    for (i64 t = 0; t < nl; ++t) p.sitems.push_back(Item{(int32_t)fl[t], 0, 0, 0, 0, 0, 0, 0});
    sadd(p.sbwd1, S_BWD_FIRST, l, i0);
    // by distance from the front's last block, across the level's fronts: every chain starts in
    // the first workgroups, and a block's x is formed by an earlier workgroup than its users
    i0 = (i64)p.sitems.size();
    i64 depth = 0;
    for (i64 t = 0; t < nl; ++t) depth = std::max<i64>(depth, (s.f_w[fl[t]] + fc::SNB - 1) / fc::SNB - 1);
    for (i64 k = 0; k < depth; ++k)
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t], jb = (s.f_w[f] + fc::SNB - 1) / fc::SNB - 2 - k;
        if (jb >= 0) p.sitems.push_back(Item{(int32_t)f, (int32_t)jb, 0, 0, 0, 0, 0, 0});
      }
    sadd(p.sbwd1, S_BWD_FRONT, l, i0);
  }
  p.uoff.assign(s.f_soff.begin(), s.f_soff.end());
  p.u_solve = s.s_size;
  return p;
}

// The plan on one device.
// A device's copy of the analysis's structure (fronts, maps, orderings), shared by every
// program of the analysis on that device.
struct DeviceStruct {
  int device = -1;
  fc::FrontDev* fronts = nullptr;
  int32_t* kids = nullptr;
  int32_t* rel = nullptr;
  int32_t* rows = nullptr;
  int64_t* a_pos = nullptr;
  int64_t* diag_src = nullptr;
  int64_t* perm = nullptr;
  ~DeviceStruct() {
    if (device < 0) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    for (void* q : {(void*)fronts, (void*)kids, (void*)rel, (void*)rows, (void*)a_pos, (void*)diag_src, (void*)perm})
      dfree(q);
    (void)hipSetDevice(cur);
  }
};

// One program's items on one device, with the device's structure arrays (held, and their
// pointers copied here for the launches).
struct DevicePlan {
  int device = -1;
  const Program* prog = nullptr;  // the program the items come from (held by its users)
  std::shared_ptr<DeviceStruct> st;
  fc::FrontDev* fronts = nullptr;
  int32_t* kids = nullptr;
  int32_t* rel = nullptr;
  int32_t* rows = nullptr;
  int64_t* a_pos = nullptr;
  int64_t* diag_src = nullptr;
  int64_t* perm = nullptr;
  fc::Item* items = nullptr;
  fc::EaRange* earanges = nullptr;
  fc::GemmItem* gitems = nullptr;
  fc::ReduceItem* ritems = nullptr;
  fc::Item* sitems = nullptr;
  int64_t* uoff = nullptr;
  ~DevicePlan() {
    if (device < 0) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    for (void* q : {(void*)items, (void*)earanges, (void*)gitems, (void*)ritems, (void*)sitems, (void*)uoff}) dfree(q);
    (void)hipSetDevice(cur);
  }
};

}  // namespace

namespace {
struct LdltPlan;
}

struct fc_symbolic {
  fc::Symbolic s;
  // LDL^T's host plan, built on the first factorization and shared by the later ones
  mutable std::mutex ldlt_mu;
  mutable std::shared_ptr<const LdltPlan> ldlt_cache;
  std::vector<int64_t> diag_src;  // per postorder column: the upper entry of its diagonal
  mutable std::mutex mu;
  std::shared_ptr<const Program> prog;  // the current program (a factor keeps the one it began with)
  std::vector<std::shared_ptr<DeviceStruct>> structs;
  std::vector<std::shared_ptr<DevicePlan>> plans;

  std::shared_ptr<const Program> program() const {
    std::lock_guard<std::mutex> g(mu);
    return prog;
  }
  // A new current program: the cached device plans of other programs are dropped (factors
  // that use them keep them alive).
  void install(std::shared_ptr<const Program> p) {
    std::lock_guard<std::mutex> g(mu);
    prog = std::move(p);
    plans.erase(std::remove_if(plans.begin(), plans.end(),
                               [&](const std::shared_ptr<DevicePlan>& q) { return q->prog != prog.get(); }),
                plans.end());
  }

  // The device plan of program pr (cached unless `cache` is false: the tuner's candidates).
  std::shared_ptr<DevicePlan> plan(int device, const std::shared_ptr<const Program>& pr, hipStream_t st,
                                   bool cache = true) {
    std::lock_guard<std::mutex> g(mu);
    if (cache)
      for (auto& p : plans)
        if (p->device == device && p->prog == pr.get()) return p;
    std::shared_ptr<DeviceStruct> ds;
    for (auto& d : structs)
      if (d->device == device) ds = d;
    if (!ds) {
      ds = std::make_shared<DeviceStruct>();
      ds->device = device;
      const auto& S = s;
      std::vector<fc::FrontDev> fd(S.nf);
      for (i64 f = 0; f < S.nf; ++f) {
        fc::FrontDev& d = fd[f];
        d.lpos = S.f_lpos[f];
        d.upos = S.f_upos[f] < 0 ? 0 : S.f_upos[f];
        d.rel_off = S.f_rel_off[f];
        d.rows_off = S.f_rptr[f];
        d.inv0 = pr->f_inv0[f];  // per-front layouts: the same in every program of the analysis
        d.bpos = pr->f_bpos[f];
        d.ld = (int32_t)S.f_ld[f];
        d.w = (int32_t)S.f_w[f];
        d.m = (int32_t)S.f_m[f];
        d.col0 = (int32_t)S.f_col0[f];
        d.parent = (int32_t)S.f_parent[f];
        d.kid0 = (int32_t)S.f_kptr[f];
        d.kid1 = (int32_t)S.f_kptr[f + 1];
        d.level = (int32_t)S.f_level[f];
      }
      std::vector<int32_t> kids(S.f_kids.begin(), S.f_kids.end());
      std::vector<int32_t> rows(S.f_rows.begin(), S.f_rows.end());
      ds->fronts = dupload(fd, st);
      ds->kids = dupload(kids, st);
      ds->rel = dupload(S.rel, st);
      ds->rows = dupload(rows, st);
      ds->a_pos = dupload(S.a_pos, st);
      ds->diag_src = dupload(diag_src, st);
      ds->perm = dupload(S.perm, st);
      structs.push_back(ds);
    }
    auto p = std::make_shared<DevicePlan>();
    p->device = device;
    p->prog = pr.get();
    p->st = ds;
    p->fronts = ds->fronts;
    p->kids = ds->kids;
    p->rel = ds->rel;
    p->rows = ds->rows;
    p->a_pos = ds->a_pos;
    p->diag_src = ds->diag_src;
    p->perm = ds->perm;
    p->items = dupload(pr->items, st);
    p->earanges = dupload(pr->earanges, st);
    p->gitems = dupload(pr->gitems, st);
    p->ritems = dupload(pr->ritems, st);
    p->sitems = dupload(pr->sitems, st);
    p->uoff = dupload(pr->uoff, st);
    HIP_CHECK(hipStreamSynchronize(st));
    if (cache) plans.push_back(p);
    return p;
  }
};

struct fc_factor {
  std::shared_ptr<fc_symbolic> sym_keep;  // keeps the analysis alive
  const fc_symbolic* sym = nullptr;
  std::shared_ptr<const Program> prog;  // the program this factor runs (its analysis's at creation)
  std::shared_ptr<DevicePlan> plan;
  int device = 0;
  int precision = FC_FP64;  // the current factorization's (fallbacks change it)
  int requested = FC_FP64;  // opts.precision
  int graph_precision = -1; // the precision the captured graphs were recorded for
  fc_factor_opts opts{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  hipStream_t side = nullptr;                  // the lookahead's second stream (default priority)
  hipEvent_t sync_ev[SYNC_SLOTS] = {};         // its ordering events
  void* L = nullptr;       // factor storage (double or float)
  double* diag = nullptr;  // A's diagonal (postorder)
  unsigned long long* fail = nullptr;
  Buffer U;                // update-matrix arena (kept between factorizations)
  Buffer inv, part;        // diagonal-block inverses, split-K pieces
  Buffer trsm_w;           // FP64: the rows below a panel, saved for the TRSM's refinement step
  Buffer fexp;             // mixed form: per front, the exponent of its power-of-two scale
  Buffer lb;               // mixed form: BF16 copies (hi, then lo) of the finished panels
  hipGraph_t graph = nullptr;          // opts.use_graph bit 0: the captured factorization
  hipGraphExec_t graph_exec = nullptr;
  struct SolveGraph {                  // opts.use_graph bit 1: a captured solve
    int nr;
    void *xp, *yp, *u, *err;
    hipGraph_t g;
    hipGraphExec_t exec;
  };
  mutable std::vector<SolveGraph> sgraphs;
  Buffer vals;             // the matrix values on the device (host-value entry points)
  mutable std::mutex solve_mu;
  mutable Buffer xp, yp, usolve, serr, refine_work;  // solve workspaces
  double seconds = 0.0;
  double kind_seconds[fc::KINDS] = {};
  double kind_launches[fc::KINDS] = {};
  bool valid = false;
  ~fc_factor() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    dfree(L);
    dfree(diag);
    dfree(fail);
    U.release();
    inv.release();
    part.release();
    trsm_w.release();
    fexp.release();
    lb.release();
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    for (auto& sg : sgraphs) {
      (void)hipGraphExecDestroy(sg.exec);
      (void)hipGraphDestroy(sg.g);
    }
    vals.release();
    xp.release();
    yp.release();
    usolve.release();
    serr.release();
    refine_work.release();
    for (auto e : sync_ev)
      if (e) (void)hipEventDestroy(e);
    if (side) (void)hipStreamDestroy(side);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    (void)hipSetDevice(cur);
  }
  // Graphs recorded for another precision's kernels are dropped (recaptured on next use).
  void drop_graphs() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    for (auto& sg : sgraphs) {
      (void)hipGraphExecDestroy(sg.exec);
      (void)hipGraphDestroy(sg.g);
    }
    sgraphs.clear();
  }
  // The side stream and its events, made once (before any capture) on the factor's device.
  void make_side() {
    if (side) return;
    HIP_CHECK(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    for (auto& e : sync_ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
};

struct fc_matrix {
  int device = 0;
  i64 n = 0, nnz = 0;
  int64_t* ptr = nullptr;
  int32_t* idx = nullptr;
  double* val = nullptr;
  double norm_inf = 0.0;
  std::vector<double> row_abs;   // sum |a_ij| of each row (host)
  // independent segments of equations (rows [seg[s], seg[s + 1])) and their norms; reductions
  // run over row blocks of at most RED_ROWS rows that never straddle a segment
  std::vector<i64> seg;
  std::vector<double> seg_norm;
  std::vector<int64_t> blk;      // row block starts (nblk + 1)
  std::vector<int32_t> blk_seg;  // segment of each row block
  int64_t* dblk = nullptr;
  ~fc_matrix() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    dfree(ptr);
    dfree(idx);
    dfree(val);
    dfree(dblk);
    (void)hipSetDevice(cur);
  }
};

struct fc_ldlt {
  int device = 0;
  std::shared_ptr<fc_symbolic> sym_keep;
  const fc_symbolic* sym = nullptr;
  hipStream_t stream = nullptr;
  int64_t n = 0, ns = 0, vals = 0, max_w = 0, positive = 0, negative = 0, two_by_two = 0;
  fc::LdlSupernode* S = nullptr;
  int64_t* rows = nullptr;
  double* val = nullptr;
  int8_t* piv = nullptr;
  int32_t* order = nullptr;
  int64_t* sn = nullptr;           // supernodes by level
  std::vector<int64_t> lvl_ptr;
  bool complex = false;            // values are (re, im) pairs: complex-symmetric LDL^T
  int64_t* perm = nullptr;         // postorder position -> original equation (device)
  fc::LdlUpdate* U = nullptr;      // updates and their relative maps, for the forward gather
  int32_t* rel = nullptr;
  fc::LdlTile* ftiles = nullptr;   // the forward gather's tiles, ftile_ptr[l] .. ftile_ptr[l + 1] per level
  std::vector<int64_t> ftile_ptr;
  ~fc_ldlt() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    for (void* p : {(void*)S, (void*)rows, (void*)val, (void*)piv, (void*)order, (void*)sn, (void*)U, (void*)rel,
                    (void*)ftiles, (void*)perm})
      dfree(p);
    if (stream) (void)hipStreamDestroy(stream);
    (void)hipSetDevice(cur);
  }
};

namespace {

// Registry of the live analyses, so a factor can keep its analysis alive after
// fc_symbolic_free. It is allocated once and never destroyed: an analysis the caller has not
// freed by process exit is left to the driver, since its device plan's hipFree would otherwise
// run from a static destructor, after the HIP runtime has been torn down.
struct Registry {
  std::mutex mu;
  std::vector<std::shared_ptr<fc_symbolic>> live;
};
Registry& registry() {
  static Registry* r = new Registry();
  return *r;
}

std::shared_ptr<fc_symbolic> keep(const fc_symbolic* s) {
  Registry& R = registry();
  std::lock_guard<std::mutex> g(R.mu);
  for (auto& p : R.live)
    if (p.get() == s) return p;
  return nullptr;
}

inline int blocks_for(i64 n, int bs = 256) { return (int)std::max<i64>(1, std::min<i64>((n + bs - 1) / bs, 1 << 16)); }

struct Events {
  std::vector<hipEvent_t> ev;
  ~Events() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
  hipEvent_t make() {
    hipEvent_t e;
    HIP_CHECK(hipEventCreate(&e));
    ev.push_back(e);
    return e;
  }
};

// Every launch of one numeric factorization, on stream st. With `eager` the cancel flag is
// polled between levels and the kinds selected by the profile mask are timed with events;
// otherwise (stream capture into a graph) neither.
template <typename T>
void enqueue_factor(fc_factor* F, const double* dval, i64 nnz, hipStream_t st, bool eager, Events& ev,
                    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>>& timed) {
  const fc::Symbolic& s = F->sym->s;
  DevicePlan& P = *F->plan;
  T* L = (T*)F->L;
  T* U = (T*)F->U.get((size_t)std::max<i64>(s.u_size, 1) * sizeof(T));
  const auto& pr = *F->prog;
  T* Inv = (T*)F->inv.get((size_t)std::max<i64>(pr.inv_size, 1) * sizeof(T));
  T* Part = (T*)F->part.get((size_t)std::max<i64>(pr.part_size, 1) * sizeof(T));
  T* W = F->precision == FC_FP64 ? (T*)F->trsm_w.get((size_t)std::max<i64>(pr.w_size, 1) * sizeof(T)) : nullptr;
  fc::k_fill<T><<<blocks_for(s.l_size), 256, 0, st>>>(L, s.l_size, T(0));
  fc::k_scatter_a<T><<<blocks_for(nnz), 256, 0, st>>>(dval, P.a_pos, nnz, L);
  fc::k_gather_diag<<<blocks_for(s.n), 256, 0, st>>>(dval, P.diag_src, s.n, F->diag);
  fc::k_fill<unsigned long long><<<1, 64, 0, st>>>(F->fail, 1, ~0ull);
  const bool mixed = F->precision == FC_MIXED;
  int32_t* fexp = nullptr;
  __bf16 *Lh = nullptr, *Ll = nullptr;
  if (mixed) {
    const size_t nb = (size_t)std::max<i64>(F->prog->b_size, 1);
    Lh = (__bf16*)F->lb.get(2 * nb * sizeof(__bf16));
    Ll = Lh + nb;
    fexp = (int32_t*)F->fexp.get((size_t)std::max<i64>(s.nf, 1) * sizeof(int32_t));
    fc::k_front_scale<<<(int)std::min<i64>(std::max<i64>(s.nf, 1), 1 << 16), 256, 0, st>>>(P.fronts, P.rows, F->diag,
                                                                                          s.nf, fexp);
  }
  const auto& prog = *F->prog;
  const int mask = eager ? F->opts.profile_mask : 0;
  int level = -1;
  for (const fc::Launch& l : prog.launches) {
    if (l.level != level) {
      level = l.level;
      if (eager && F->opts.cancel && *F->opts.cancel) throw fc::Failure{fc::CANCELLED, -1, 0, 0, "cancelled"};
    }
    if (l.kind == L_RECORD || l.kind == L_WAIT) {
      hipStream_t on = l.stream ? F->side : st;
      if (l.kind == L_RECORD) HIP_CHECK(hipEventRecord(F->sync_ev[l.slot], on));
      else HIP_CHECK(hipStreamWaitEvent(on, F->sync_ev[l.slot], 0));
      continue;
    }
    if (l.kind == fc::K_TRSM && (l.pad == TRSM_IN_PLACE) == (F->precision == FC_FP64)) continue;
    if ((l.kind == fc::K_UPDATE || l.kind == fc::K_REDUCE) && l.pad != SIDE_ANY &&
        (l.pad == SIDE_FP64) != (F->precision == FC_FP64))
      continue;  // the other precision's variant of a side-stream update
    const fc::Item* it = P.items + l.item0;
    hipStream_t ls = l.stream ? F->side : st;  // the launch's stream
    dim3 g(l.count), b(256);
    hipEvent_t a = nullptr, z = nullptr;
    if (mask & (1 << l.kind)) {
      a = ev.make();
      z = ev.make();
      HIP_CHECK(hipEventRecord(a, ls));
    }
    switch (l.kind) {
      case fc::K_EXTEND_L:
      case fc::K_EXTEND_U:
        fc::k_extend_add<T><<<g, b, 0, ls>>>(it, P.earanges, P.fronts, P.rel, L, U, U);
        break;
      case fc::K_UPDATE:
      case fc::K_SYRK:
        if constexpr (std::is_same<T, float>::value) {
          if (mixed) {
            if (F->opts.bf16_terms == 1)
              fc::k_gemm_b16<1><<<g, b, 0, ls>>>(P.gitems + l.item0, P.fronts, Lh, Ll, L, U, Part, fexp);
            else
              fc::k_gemm_b16<3><<<g, b, 0, ls>>>(P.gitems + l.item0, P.fronts, Lh, Ll, L, U, Part, fexp);
            break;
          }
        }
        fc::k_gemm<T><<<g, b, 0, ls>>>(P.gitems + l.item0, L, U, Inv, Part, W, P.fronts, nullptr, nullptr, nullptr);
        break;
      case fc::K_TRSM:
        if (l.pad == TRSM_COPY) {
          fc::k_copy_tiles<T><<<g, b, 0, ls>>>(P.gitems + l.item0, L, W);
        } else {
          if constexpr (std::is_same<T, float>::value) {
            if (mixed) {  // the in-place product also writes the rows' BF16 copies
              fc::k_gemm<T, true><<<g, b, 0, ls>>>(P.gitems + l.item0, L, U, Inv, Part, W, P.fronts, Lh, Ll, fexp);
              break;
            }
          }
          fc::k_gemm<T><<<g, b, 0, ls>>>(P.gitems + l.item0, L, U, Inv, Part, W, P.fronts, nullptr, nullptr, nullptr);
        }
        break;
      case fc::K_REDUCE:
        fc::k_reduce_parts<T><<<dim3(l.count * fc::RSPLIT), b, 0, ls>>>(P.ritems + l.item0, L, U, Part);
        break;
      case fc::K_POTRF:
        fc::k_potrf_blk<T><<<g, dim3(256), 0, ls>>>(it, P.fronts, L, F->diag, F->fail, Inv, Lh, Ll, fexp);
        break;
      case fc::K_SMALL:
        fc::k_small_fronts<T><<<g, b, 0, ls>>>(it, P.fronts, L, U, F->diag, F->fail, Inv);
        break;
      default:
        throw fc::Failure{fc::INTERNAL, -1, 0, 0, "unknown launch"};
    }
    if (a) {
      HIP_CHECK(hipEventRecord(z, ls));
      timed.push_back({l.kind, {a, z}});
    }
  }
  HIP_CHECK(hipGetLastError());
}

// One numeric factorization. With opts.use_graph the launches are captured into a hipGraph on
// the first call (from the factor's own copy of the values, so replays need no new arguments)
// and replayed on every call: the cancel flag is then polled once, before the replay, and no
// kind is timed.
template <typename T>
void run_factor(fc_factor* F, const double* dval, i64 nnz, hipStream_t st) {
  F->make_side();
  Events ev;
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> timed;
  hipEvent_t e0 = ev.make(), e1 = ev.make();
  if (F->opts.use_graph & 1) {
    double* own = (double*)F->vals.get((size_t)std::max<i64>(nnz, 1) * sizeof(double));
    if (dval != own && nnz) HIP_CHECK(hipMemcpyAsync(own, dval, nnz * sizeof(double), hipMemcpyDeviceToDevice, st));
    if (F->opts.cancel && *F->opts.cancel) throw fc::Failure{fc::CANCELLED, -1, 0, 0, "cancelled"};
    if (!F->graph_exec) {
      // every workspace is allocated before the capture (enqueue_factor then finds them)
      const fc::Symbolic& s = F->sym->s;
      const auto& pr = *F->prog;
      (void)F->U.get((size_t)std::max<i64>(s.u_size, 1) * sizeof(T));
      (void)F->inv.get((size_t)std::max<i64>(pr.inv_size, 1) * sizeof(T));
      (void)F->part.get((size_t)std::max<i64>(pr.part_size, 1) * sizeof(T));
      if (F->precision == FC_FP64) (void)F->trsm_w.get((size_t)std::max<i64>(pr.w_size, 1) * sizeof(T));
      if (F->precision == FC_MIXED) {
        (void)F->lb.get(2 * (size_t)std::max<i64>(pr.b_size, 1) * sizeof(__bf16));
        (void)F->fexp.get((size_t)std::max<i64>(s.nf, 1) * sizeof(int32_t));
      }
      HIP_CHECK(hipStreamSynchronize(st));
      hipGraph_t g = nullptr;
      HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      try {
        enqueue_factor<T>(F, own, nnz, st, false, ev, timed);
      } catch (...) {
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HIP_CHECK(hipStreamEndCapture(st, &g));
      F->graph = g;
      HIP_CHECK(hipGraphInstantiate(&F->graph_exec, g, nullptr, nullptr, 0));
    }
    HIP_CHECK(hipEventRecord(e0, st));
    HIP_CHECK(hipGraphLaunch(F->graph_exec, st));
  } else {
    HIP_CHECK(hipEventRecord(e0, st));
    enqueue_factor<T>(F, dval, nnz, st, true, ev, timed);
  }
  HIP_CHECK(hipEventRecord(e1, st));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  F->seconds = ms * 1e-3;
  for (int k = 0; k < fc::KINDS; ++k) F->kind_seconds[k] = F->kind_launches[k] = 0;
  for (auto& t : timed) {
    float m = 0.f;
    HIP_CHECK(hipEventElapsedTime(&m, t.second.first, t.second.second));
    F->kind_seconds[t.first] += m * 1e-3;
    F->kind_launches[t.first] += 1;
  }
}

// numeric phase from device values, failures as values
int factor_device(fc_factor* F, const double* dval, uint64_t nnz, fc_error* err) {
  const fc::Symbolic& s = F->sym->s;
  if ((i64)nnz != (i64)s.upper_idx.size()) {
    set_error(err, FC_INVALID_ARGUMENT, "the matrix's pattern differs from the analysed one (entry count)");
    return FC_INVALID_ARGUMENT;
  }
  try {
    HIP_CHECK(hipSetDevice(F->device));
    hipStream_t st = F->stream;
    F->valid = false;
    // each factorization starts from the precision asked for, unless a fallback has moved the
    // factor to FP64 storage
    if (F->precision != FC_FP64 || F->requested == FC_FP64) F->precision = F->requested;
    if (!F->L) {  // a failed fallback could not restore the storage: allocate it again
      const size_t es = F->precision == FC_FP64 ? sizeof(double) : sizeof(float);
      F->L = (void*)dalloc<char>((size_t)std::max<i64>(s.l_size, 1) * es);
    }
    double seconds = 0.0;
    for (;;) {
      if (F->graph_precision != F->precision) {
        F->drop_graphs();
        F->graph_precision = F->precision;
      }
      if (F->precision == FC_FP64) run_factor<double>(F, dval, (i64)nnz, st);
      else run_factor<float>(F, dval, (i64)nnz, st);
      seconds += F->seconds;
      unsigned long long fail = 0;
      HIP_CHECK(hipMemcpyAsync(&fail, F->fail, sizeof fail, hipMemcpyDeviceToHost, st));
      HIP_CHECK(hipStreamSynchronize(st));
      if (fail == ~0ull) break;
      if (F->precision == FC_FP64 || F->opts.fallback != 0) {
        F->seconds = seconds;
        const uint64_t col = (uint64_t)s.perm[(i64)fail];
        set_error(err, FC_NOT_POSITIVE_DEFINITE, "the matrix is not positive definite", col, s.entries);
        return FC_NOT_POSITIVE_DEFINITE;
      }
      // the reduced precision lost definiteness: the next precision, mixed -> FP32 on the same
      // FP32 storage, FP32 -> FP64 on new FP64 storage (kept for later factorizations)
      if (F->precision == FC_MIXED) {
        F->precision = FC_FP32;
      } else {
        dfree(F->L);
        F->L = nullptr;
        F->U.release();
        F->inv.release();
        F->part.release();
        try {
          F->L = (void*)dalloc<double>((size_t)std::max<i64>(s.l_size, 1));
        } catch (const std::bad_alloc&) {
          F->L = (void*)dalloc<float>((size_t)std::max<i64>(s.l_size, 1));  // the storage just freed
          throw;
        }
        F->precision = FC_FP64;
      }
    }
    F->seconds = seconds;
    F->valid = true;
    set_error(err, FC_OK, "");
    return FC_OK;
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough device memory for the factorization", 0, s.entries);
    return FC_OUT_OF_MEMORY;
  } catch (const fc::Failure& f) {
    set_error(err, f.status, f.message);
    return f.status;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
}

int factor_host(fc_factor* F, const double* upper_val, uint64_t nnz, fc_error* err) {
  try {
    HIP_CHECK(hipSetDevice(F->device));
    double* dval = (double*)F->vals.get(std::max<uint64_t>(nnz, 1) * sizeof(double));
    if (nnz) HIP_CHECK(hipMemcpyAsync(dval, upper_val, nnz * sizeof(double), hipMemcpyHostToDevice, F->stream));
    return factor_device(F, dval, nnz, err);
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough device memory for the matrix values", 0, F->sym->s.entries);
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
}

// A solve pass's device work: x and y (n x nr), the fronts' contributions, and the error word of
// the one-right-hand-side chains (set when a wait gave up).
struct SolveWork {
  double *xp, *yp, *u;
  int32_t* err;
};

// The polls a chained solve's wait may take before it gives up (FC_SOLVE_SPIN_LIMIT; negative:
// every wait gives up at once, a test's way to reach the failure path).
int64_t solve_spin_limit() {
  if (const char* e = std::getenv("FC_SOLVE_SPIN_LIMIT")) return (int64_t)std::strtoll(e, nullptr, 10);
  return (int64_t)1 << 22;
}

template <typename T, int R>
void run_solve_r(const fc_factor* F, const SolveWork& w, i64 ldx, int nrhs, hipStream_t st) {
  const Program& prog = *F->prog;
  DevicePlan& P = *F->plan;
  const T* L = (const T*)F->L;
  const T* Inv = (const T*)F->inv.p;
  double *xp = w.xp, *yp = w.yp, *u = w.u;
  int32_t* err = w.err;
  const int64_t limit = solve_spin_limit();
  const int mark = R == 1;  // the chained passes: the gathers mark the block solutions to come
  const int groups = (nrhs + fc::RSTEP - 1) / fc::RSTEP;  // block steps: right-hand-side groups
  auto run = [&](const fc::Launch& l) {
    const fc::Item* it = P.sitems + l.item0;
    dim3 g(l.count), b(256);
    switch (l.kind) {
      case S_GATHER:
        fc::k_fwd_gather<T, R><<<g, b, 0, st>>>(it, P.fronts, P.kids, P.rel, Inv, xp, yp, ldx, nrhs, u, P.uoff, mark);
        break;
      case S_FWD_STEP:  // every right-hand side in one workgroup (groups side by side would each
                        // read the step's rows of L again: measured slower)
        fc::k_fwd_step<T, R><<<g, b, 0, st>>>(it, P.fronts, L, Inv, xp, yp, ldx, nrhs, u, P.uoff, l.count);
        break;
      case S_BWD_GATHER:
        fc::k_bwd_gather<<<g, b, 0, st>>>(it, P.fronts, P.rows, xp, ldx, nrhs, u, P.uoff);
        break;
      case S_BWD_BELOW:
        fc::k_bwd_below<T, R><<<g, b, 0, st>>>(it, P.fronts, L, u, P.uoff, yp, ldx, nrhs);
        break;
      case S_BWD_FIRST:
        fc::k_bwd_first<T, R><<<g, b, 0, st>>>(it, P.fronts, Inv, yp, xp, ldx, nrhs, mark);
        break;
      case S_BWD_STEP:
        if constexpr (R == 1) fc::k_bwd_step1<T><<<g, b, 0, st>>>(it, P.fronts, L, Inv, yp, xp);
        else  // several right-hand sides: groups of RSTEP side by side (a step's tile is small)
          fc::k_bwd_step<T, fc::RSTEP><<<dim3(l.count * groups), b, 0, st>>>(it, P.fronts, L, Inv, yp, xp, ldx, nrhs,
                                                                             l.count);
        break;
      case S_FWD_FRONT:
        fc::k_fwd_front<T><<<g, b, fc::SNB * fc::SNB * sizeof(T), st>>>(it, P.fronts, L, Inv, xp, yp, u, P.uoff, err,
                                                                         limit);
        break;
      case S_BWD_FRONT:
        fc::k_bwd_front<T><<<g, b, 0, st>>>(it, P.fronts, L, Inv, yp, xp, err, limit);
        break;
    }
  };
  if constexpr (R == 1) {
    for (const fc::Launch& l : prog.sfwd1) run(l);
    for (const fc::Launch& l : prog.sbwd1) run(l);
  } else {
    for (const fc::Launch& l : prog.sfwd) run(l);
    for (const fc::Launch& l : prog.sbwd) run(l);
  }
  HIP_CHECK(hipGetLastError());
}

template <typename T>
void run_solve(const fc_factor* F, const SolveWork& w, i64 ldx, int nrhs, hipStream_t st) {
  if (nrhs == 1) run_solve_r<T, 1>(F, w, ldx, nrhs, st);
  else if (nrhs <= fc::RC / 2) run_solve_r<T, fc::RC / 2>(F, w, ldx, nrhs, st);  // a Lanczos block
  else run_solve_r<T, fc::RC>(F, w, ldx, nrhs, st);
}

// The solve's launches for nr right-hand sides in w, replayed from a hipGraph when
// opts.use_graph asks for it (bit 1): captured on first use for each (nr, workspaces).
void solve_launches(const fc_factor* F, const SolveWork& w, i64 n, int nr, hipStream_t st) {
  auto eager = [&]() {
    if (F->precision == FC_FP64) run_solve<double>(F, w, n, nr, st);
    else run_solve<float>(F, w, n, nr, st);
  };
  if (!(F->opts.use_graph & 2)) return eager();
  for (const auto& sg : F->sgraphs)
    if (sg.nr == nr && sg.xp == w.xp && sg.yp == w.yp && sg.u == w.u && sg.err == w.err) {
      HIP_CHECK(hipGraphLaunch(sg.exec, st));
      return;
    }
  fc_factor::SolveGraph sg{nr, w.xp, w.yp, w.u, w.err, nullptr, nullptr};
  HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try {
    eager();
  } catch (...) {
    (void)hipStreamEndCapture(st, &sg.g);
    if (sg.g) (void)hipGraphDestroy(sg.g);
    throw;
  }
  HIP_CHECK(hipStreamEndCapture(st, &sg.g));
  HIP_CHECK(hipGraphInstantiate(&sg.exec, sg.g, nullptr, nullptr, 0));
  F->sgraphs.push_back(sg);
  HIP_CHECK(hipGraphLaunch(sg.exec, st));
}

// The error word of a chained solve: read (waiting for the stream), cleared when set; a set word
// means a wait gave up and the solution is not valid.
void solve_check(int32_t* err, hipStream_t st) {
  int32_t e = 0;
  HIP_CHECK(hipMemcpyAsync(&e, err, sizeof e, hipMemcpyDeviceToHost, st));
  HIP_CHECK(hipStreamSynchronize(st));
  if (e) {
    fc::k_fill<int32_t><<<1, 64, 0, st>>>(err, 1, 0);
    HIP_CHECK(hipStreamSynchronize(st));
    throw DeviceError{hipErrorLaunchTimeOut, "a solve's wait for a block solution ran past its limit"};
  }
}

// b (device, n x nrhs, ldb) solved in place, MAX_RHS right-hand sides at a time; the caller
// holds solve_mu. Returns once the stream has finished (the error word is read).
void solve_device(const fc_factor* F, double* b, i64 n, int nrhs, i64 ldb, hipStream_t st) {
  if (n == 0 || nrhs == 0) return;
  if (!F->serr.p) {  // the error word starts clear (and is cleared whenever it is found set)
    F->serr.get(sizeof(int32_t));
    fc::k_fill<int32_t><<<1, 64, 0, st>>>((int32_t*)F->serr.p, 1, 0);
  }
  int32_t* err = (int32_t*)F->serr.p;
  for (int r0 = 0; r0 < nrhs; r0 += fc::MAX_RHS) {
    const int nr = std::min(fc::MAX_RHS, nrhs - r0);
    double* br = b + (i64)r0 * ldb;
    SolveWork w;
    w.xp = (double*)F->xp.get((size_t)n * nr * sizeof(double));
    w.yp = (double*)F->yp.get((size_t)n * nr * sizeof(double));
    w.u = (double*)F->usolve.get((size_t)std::max<i64>(F->prog->u_solve * nr, 1) * sizeof(double));
    w.err = err;
    fc::k_permute<<<blocks_for(n), 256, 0, st>>>(br, w.xp, F->plan->perm, n, nr, ldb, n, 0);
    solve_launches(F, w, n, nr, st);
    fc::k_permute<<<blocks_for(n), 256, 0, st>>>(w.xp, br, F->plan->perm, n, nr, n, ldb, 1);
  }
  HIP_CHECK(hipGetLastError());
  solve_check(err, st);
}

constexpr i64 RED_ROWS = 8192;  // rows of a reduction block

// Segments and row blocks of a matrix (one segment: the whole matrix).
void set_segments(fc_matrix* A, const std::vector<i64>& seg) {
  A->seg = seg;
  A->seg_norm.assign(seg.size() - 1, 0.0);
  A->blk.clear();
  A->blk_seg.clear();
  for (size_t q = 0; q + 1 < seg.size(); ++q) {
    for (i64 r = seg[q]; r < seg[q + 1]; ++r) A->seg_norm[q] = std::max(A->seg_norm[q], A->row_abs[r]);
    for (i64 r = seg[q]; r < seg[q + 1]; r += RED_ROWS) {
      A->blk.push_back(r);
      A->blk_seg.push_back((int32_t)q);
    }
  }
  A->blk.push_back(A->n);
  A->norm_inf = 0.0;
  for (double v : A->seg_norm) A->norm_inf = std::max(A->norm_inf, v);
  dfree(A->dblk);
  A->dblk = nullptr;
  A->dblk = dupload(A->blk, nullptr);
  HIP_CHECK(hipDeviceSynchronize());
}

// Column-wise reductions over the matrix's row blocks, combined on the host in order.
struct Reducer {
  const fc_matrix* A = nullptr;
  double* partial = nullptr;  // k * nblk on the device
  std::vector<double> host;
  int nblk = 0, k = 0;
  void run(hipStream_t st, const double* a, const double* b, int mode, int kk, i64 ld) {
    k = kk;
    nblk = (int)A->blk.size() - 1;
    fc::k_reduce<<<dim3(nblk, k), 256, 0, st>>>(A->dblk, nblk, a, b, ld, mode, partial);
    host.resize((size_t)nblk * k);
    HIP_CHECK(hipMemcpyAsync(host.data(), partial, host.size() * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
  }
  double sum(int c) const {
    double s = 0.0;
    for (int q = 0; q < nblk; ++q) s += host[(size_t)c * nblk + q];
    return s;
  }
  // max over the blocks of each segment, for column c
  void seg_max(int c, std::vector<double>& out) const {
    out.assign(A->seg.size() - 1, 0.0);
    for (int q = 0; q < nblk; ++q) out[A->blk_seg[q]] = std::max(out[A->blk_seg[q]], host[(size_t)c * nblk + q]);
  }
};

// Refinement of k <= MAX_RHS solutions X (N x k, ld = N) of A X = B with the factor's solves:
// plain refinement while every step at least halves the backward error, then conjugate
// gradients preconditioned by the factor (per-column scalars). The backward error of a column
// is the largest over the segments of ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf);
// iteration stops when every column meets the target. berr[c] receives each column's.
void refine_block(const fc_factor* f, const fc_matrix* a, const double* db, double* dx, i64 N, int k, fc_refine_opts o,
                  fc_refine_report& rep, double* berr_cols, Buffer& work) {
  hipStream_t st = f->stream;
  const i64 NK = N * k;
  const int nblk = (int)a->blk.size() - 1;
  double* buf = (double*)work.get(((size_t)NK * 4 + (size_t)nblk * k) * sizeof(double));
  double *dr = buf, *dz = buf + NK, *dp = buf + 2 * NK, *dq = buf + 3 * NK;
  Reducer red;
  red.A = a;
  red.partial = buf + 4 * NK;
  const int g = blocks_for(NK);
  Events ev;
  hipEvent_t e0 = ev.make(), e1 = ev.make();
  HIP_CHECK(hipEventRecord(e0, st));
  const size_t nseg = a->seg.size() - 1;
  std::vector<std::vector<double>> bn(k);
  red.run(st, db, db, 1, k, N);
  for (int c = 0; c < k; ++c) red.seg_max(c, bn[c]);
  auto spmm = [&](const double* in, double* out) {
    fc::k_spmm<<<blocks_for(N * 8), 256, 0, st>>>(N, a->ptr, a->idx, a->val, in, out, k, N);
  };
  std::vector<double> berr(k, 0.0), rn, xn;
  auto berr_of = [&](const double* r) {  // per column, into berr; returns the largest
    std::vector<std::vector<double>> rs(k);
    red.run(st, r, r, 1, k, N);
    for (int c = 0; c < k; ++c) red.seg_max(c, rs[c]);
    red.run(st, dx, dx, 1, k, N);
    double worst = 0.0;
    for (int c = 0; c < k; ++c) {
      red.seg_max(c, xn);
      berr[c] = 0.0;
      for (size_t q = 0; q < nseg; ++q) {
        const double den = a->seg_norm[q] * xn[q] + bn[c][q];
        berr[c] = std::max(berr[c], den > 0 ? rs[c][q] / den : 0.0);
      }
      worst = std::max(worst, berr[c]);
    }
    return worst;
  };
  auto residual = [&]() {  // R = B - A X
    spmm(dx, dq);
    fc::k_residual<<<g, 256, 0, st>>>(NK, db, dq, dr);
    return berr_of(dr);
  };
  HIP_CHECK(hipMemcpyAsync(dx, db, NK * sizeof(double), hipMemcpyDeviceToDevice, st));
  solve_device(f, dx, N, k, N, st);
  double worst = residual();
  rep.berr_history[0] = worst;
  int it = 0;
  bool cg = o.method == 1;
  std::vector<double> prev = berr;
  fc::Coefs one{};
  // plain refinement while it contracts
  while (!cg && worst > o.target_berr && it < o.max_iter) {
    HIP_CHECK(hipMemcpyAsync(dz, dr, NK * sizeof(double), hipMemcpyDeviceToDevice, st));
    solve_device(f, dz, N, k, N, st);
    for (int c = 0; c < k; ++c) one.v[c] = berr[c] > o.target_berr ? 1.0 : 0.0;
    fc::k_axpy_cols<<<g, 256, 0, st>>>(N, k, N, one, dz, dx);
    ++it;
    worst = residual();
    if (it < 64) rep.berr_history[it] = worst;
    for (int c = 0; c < k; ++c)
      if (o.method == 2 && berr[c] > o.target_berr && berr[c] > 0.5 * prev[c]) cg = true;  // not contracting
    prev = berr;
  }
  if (cg && worst > o.target_berr && it < o.max_iter) {
    // conjugate gradients on A D = R from the current X, preconditioned by the factor
    HIP_CHECK(hipMemcpyAsync(dz, dr, NK * sizeof(double), hipMemcpyDeviceToDevice, st));
    solve_device(f, dz, N, k, N, st);
    HIP_CHECK(hipMemcpyAsync(dp, dz, NK * sizeof(double), hipMemcpyDeviceToDevice, st));
    std::vector<double> rz(k), pq(k);
    red.run(st, dr, dz, 0, k, N);
    for (int c = 0; c < k; ++c) rz[c] = red.sum(c);
    std::vector<bool> live(k);
    for (int c = 0; c < k; ++c) live[c] = berr[c] > o.target_berr;
    while (it < o.max_iter) {
      spmm(dp, dq);
      red.run(st, dp, dq, 0, k, N);
      fc::Coefs al{}, nal{};
      bool any = false;
      for (int c = 0; c < k; ++c) {
        pq[c] = red.sum(c);
        if (live[c] && !(pq[c] > 0)) live[c] = false;  // breakdown: keep this column as it is
        al.v[c] = live[c] ? rz[c] / pq[c] : 0.0;
        nal.v[c] = -al.v[c];
        any = any || live[c];
      }
      if (!any) break;
      fc::k_axpy_cols<<<g, 256, 0, st>>>(N, k, N, al, dp, dx);
      fc::k_axpy_cols<<<g, 256, 0, st>>>(N, k, N, nal, dq, dr);
      ++it;
      // the backward error from the true residual, into dq (free until the next A p)
      spmm(dx, dq);
      fc::k_residual<<<g, 256, 0, st>>>(NK, db, dq, dq);
      worst = berr_of(dq);
      if (it < 64) rep.berr_history[it] = worst;
      if (worst <= o.target_berr) break;
      for (int c = 0; c < k; ++c) live[c] = live[c] && berr[c] > o.target_berr;
      HIP_CHECK(hipMemcpyAsync(dz, dr, NK * sizeof(double), hipMemcpyDeviceToDevice, st));
      solve_device(f, dz, N, k, N, st);
      red.run(st, dr, dz, 0, k, N);
      fc::Coefs be{};
      for (int c = 0; c < k; ++c) {
        const double rz2 = red.sum(c);
        be.v[c] = (live[c] && rz[c] != 0.0) ? rz2 / rz[c] : 0.0;
        rz[c] = rz2;
      }
      fc::k_xpby_cols<<<g, 256, 0, st>>>(N, k, N, dz, be, dp);
    }
  }
  HIP_CHECK(hipEventRecord(e1, st));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  rep.iterations = it;
  rep.berr = worst;
  rep.converged = worst <= o.target_berr;
  rep.seconds = ms * 1e-3;
  if (berr_cols)
    for (int c = 0; c < k; ++c) berr_cols[c] = berr[c];
}

// refine_block over blocks of MAX_RHS columns; the report keeps the worst block's berr, the
// iterations of the slowest block and the summed time
void refine_many(const fc_factor* f, const fc_matrix* a, const double* db, double* dx, i64 N, i64 nrhs,
                 fc_refine_opts o, fc_refine_report& rep, double* berr_cols, Buffer& work) {
  std::memset(&rep, 0, sizeof rep);
  rep.converged = 1;
  for (i64 c0 = 0; c0 < nrhs; c0 += fc::MAX_RHS) {
    const int k = (int)std::min<i64>(fc::MAX_RHS, nrhs - c0);
    fc_refine_report r;
    std::memset(&r, 0, sizeof r);
    refine_block(f, a, db + c0 * N, dx + c0 * N, N, k, o, r, berr_cols ? berr_cols + c0 : nullptr, work);
    if (r.iterations >= rep.iterations) {
      rep.iterations = r.iterations;
      std::memcpy(rep.berr_history, r.berr_history, sizeof rep.berr_history);
    }
    rep.berr = std::max(rep.berr, r.berr);
    rep.converged = rep.converged && r.converged;
    rep.seconds += r.seconds;
  }
}

fc_refine_opts refine_defaults(const fc_refine_opts* opts) {
  fc_refine_opts o{2, 50, 1e-12};
  if (opts) {
    o = *opts;
    if (o.max_iter <= 0) o.max_iter = 50;
    if (!(o.target_berr > 0)) o.target_berr = 1e-12;
  }
  return o;
}

template <typename Body>
int guarded(fc_error* err, const char* oom, Body body) {
  try {
    body();
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, oom);
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  set_error(err, FC_OK, "");
  return FC_OK;
}

// ---------------------------------------------------------------------------------------------
// Shift-invert block Lanczos (crates/solver/src/eigen.rs:281-466) on the device.

// xorshift64* (eigen.rs:159-175), uniform in [-1, 1): the start block and fresh directions.
struct XorShift {
  uint64_t s;
  explicit XorShift(uint64_t seed) : s(seed ? seed : 1) {}
  double uniform() {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    const uint64_t v = s * 0x2545F4914F6CDD1DULL;
    return (double)(v >> 11) / (double)(1ULL << 52) - 1.0;
  }
};

// Eigenvalues ascending and eigenvectors (column k of the row-major `vec`) of a dense symmetric
// n x n matrix (row-major `a`, overwritten): Householder reduction to tridiagonal form, then
// implicit QR steps with Wilkinson shifts on the tridiagonal matrix, the transformations
// accumulated (Golub & Van Loan, Matrix Computations, 4th ed., algorithms 8.3.1 and 8.3.3).
// False when the QR iteration does not converge within 30 n + 30 steps.
bool dense_symmetric_eigen(int n, std::vector<double>& a, std::vector<double>& values, std::vector<double>& vec) {
  std::vector<double> q((size_t)n * n, 0.0), v(n), p(n);
  for (int i = 0; i < n; ++i) q[(size_t)i * n + i] = 1.0;
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  for (int k = 0; k + 2 < n; ++k) {
    double nx = 0.0;
    for (int i = k + 1; i < n; ++i) nx += A(i, k) * A(i, k);
    nx = std::sqrt(nx);
    if (nx == 0.0) continue;
    const double alpha = A(k + 1, k) >= 0.0 ? -nx : nx;
    double nv = 0.0;
    for (int i = k + 1; i < n; ++i) {
      v[i] = A(i, k) - (i == k + 1 ? alpha : 0.0);
      nv += v[i] * v[i];
    }
    nv = std::sqrt(nv);
    if (nv == 0.0) continue;
    for (int i = k + 1; i < n; ++i) v[i] /= nv;
    double vp = 0.0;
    for (int i = k + 1; i < n; ++i) {
      double t = 0.0;
      for (int j = k + 1; j < n; ++j) t += A(i, j) * v[j];
      p[i] = t;
      vp += v[i] * t;
    }
    for (int i = k + 1; i < n; ++i) p[i] -= vp * v[i];
    for (int i = k + 1; i < n; ++i)
      for (int j = k + 1; j < n; ++j) A(i, j) -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
    A(k + 1, k) = A(k, k + 1) = alpha;
    for (int i = k + 2; i < n; ++i) A(i, k) = A(k, i) = 0.0;
    for (int r = 0; r < n; ++r) {
      double t = 0.0;
      for (int j = k + 1; j < n; ++j) t += q[(size_t)r * n + j] * v[j];
      for (int j = k + 1; j < n; ++j) q[(size_t)r * n + j] -= 2.0 * t * v[j];
    }
  }
  std::vector<double> d(n), e(std::max(n - 1, 0));
  for (int i = 0; i < n; ++i) d[i] = A(i, i);
  for (int i = 0; i + 1 < n; ++i) e[i] = A(i + 1, i);
  const double eps = std::numeric_limits<double>::epsilon();
  auto negligible = [&](double ev, double x, double y) { return std::fabs(ev) <= eps * (std::fabs(x) + std::fabs(y)); };
  int h = n - 1, steps = 0;
  while (h > 0) {
    if (negligible(e[h - 1], d[h - 1], d[h])) {
      e[h - 1] = 0.0;
      --h;
      continue;
    }
    int l = h - 1;
    while (l > 0 && !negligible(e[l - 1], d[l - 1], d[l])) --l;
    if (l > 0) e[l - 1] = 0.0;
    if (++steps > 30 * n + 30) return false;
    const double dd = (d[h - 1] - d[h]) / 2.0, eh = e[h - 1];
    const double mu = d[h] - eh * eh / (dd + (dd >= 0.0 ? 1.0 : -1.0) * std::hypot(dd, eh));
    double x = d[l] - mu, z = e[l];
    for (int k = l; k < h; ++k) {
      const double r = std::hypot(x, z);
      const double c = r == 0.0 ? 1.0 : x / r, sn = r == 0.0 ? 0.0 : -z / r;
      if (k > l) e[k - 1] = c * e[k - 1] - sn * z;
      const double a0 = d[k], b0 = e[k], d1 = d[k + 1];
      d[k] = c * c * a0 - 2.0 * c * sn * b0 + sn * sn * d1;
      d[k + 1] = sn * sn * a0 + 2.0 * c * sn * b0 + c * c * d1;
      e[k] = c * sn * (a0 - d1) + (c * c - sn * sn) * b0;
      if (k + 1 < h) {
        z = -sn * e[k + 1];
        e[k + 1] *= c;
        x = e[k];
      }
      for (int r2 = 0; r2 < n; ++r2) {
        const double qp = q[(size_t)r2 * n + k], qq = q[(size_t)r2 * n + k + 1];
        q[(size_t)r2 * n + k] = c * qp - sn * qq;
        q[(size_t)r2 * n + k + 1] = sn * qp + c * qq;
      }
    }
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y] || (d[x] == d[y] && x < y); });
  values.resize(n);
  vec.assign((size_t)n * n, 0.0);
  for (int k = 0; k < n; ++k) {
    values[k] = d[order[k]];
    for (int r = 0; r < n; ++r) vec[(size_t)r * n + k] = q[(size_t)r * n + order[k]];
  }
  return true;
}

// ---------------------------------------------------------------------------------------------
// LDL^T (crates/solver/src/ldlt.rs) on the device: the plan from the analysis, built on the host.

struct LdltPlan {
  std::vector<fc::LdlSupernode> sn;
  std::vector<fc::LdlUpdate> up;
  std::vector<int32_t> rel;
  std::vector<int64_t> pos;         // upper entry -> slot in the trapezoids
  std::vector<int64_t> coord_r, coord_c;  // upper entry -> postorder (row, column)
  std::vector<int64_t> lvl_ptr, lvl_sn;   // supernodes by level of the supernodal tree
  std::vector<int64_t> tile_ptr;          // tiles of each level's assemble
  std::vector<fc::LdlTile> tiles;
  // Each level lists its narrow supernodes first (lvl_narrow of them, one workgroup each for the
  // whole factorization) and then its wide ones, factored one panel per launch: panel launch i
  // of the level's wide supernodes is followed by the trailing update over the tiles
  // pan_tiles[pan_tile_ptr[i] .. pan_tile_ptr[i + 1]], panel launches pan_ptr[l] .. pan_ptr[l + 1].
  std::vector<int64_t> lvl_narrow;
  std::vector<int64_t> pan_ptr, pan_tile_ptr;
  std::vector<fc::LdlTile> pan_tiles;
  std::vector<int64_t> ftile_ptr;  // the solve's forward gather: tiles of each level
  std::vector<fc::LdlTile> ftiles;
  int64_t vals = 0, work = 0, max_w = 0;
};

constexpr int64_t LDL_TILE_COLS = fc::LDL_YC;  // target columns of one assemble or trailing workgroup
constexpr int64_t LDL_TILE_ROWS = 128;         // and its target rows
constexpr int64_t LDL_GATHER_COLS = 64;        // unknowns of one forward-gather workgroup
constexpr size_t LDL_WORK_LDS = 48 * 1024;     // dynamic LDS for a panel kernel's two work columns
constexpr unsigned LDL_PANEL_THREADS = 1024;    // threads of a panel-search workgroup (512 measured no faster)

// tiles of LDL_TILE_COLS target columns [b, e) of supernode k by bands of LDL_TILE_ROWS rows from b
void ldl_tiles(std::vector<fc::LdlTile>& out, i64 k, i64 b, i64 e, i64 m) {
  for (i64 rb = b; rb < m; rb += LDL_TILE_ROWS) out.push_back(fc::LdlTile{k, b, e, rb, std::min<i64>(rb + LDL_TILE_ROWS, m)});
}
constexpr int64_t LDL_WIDE = 192;      // wider supernodes are factored one panel per launch

LdltPlan ldlt_plan(const fc::Symbolic& s) {
  LdltPlan P;
  const i64 ns = s.ns(), n = s.n;
  std::vector<i64> snode(n);
  for (i64 k = 0; k < ns; ++k)
    for (i64 c = s.first[k]; c < s.first[k + 1]; ++c) snode[c] = k;
  P.sn.resize(ns);
  for (i64 k = 0; k < ns; ++k) {
    fc::LdlSupernode& J = P.sn[k];
    J.m = (int32_t)(s.rptr[k + 1] - s.rptr[k]);
    J.w = (int32_t)(s.first[k + 1] - s.first[k]);
    J.voff = P.vals;
    J.rows = s.rptr[k];
    J.first = s.first[k];
    J.scratch = P.work;
    P.vals += (i64)J.w * J.m - (i64)J.w * (J.w - 1) / 2;
    P.work += 2 * (i64)J.m;
    P.max_w = std::max<i64>(P.max_w, J.w);
    J.upd0 = (i64)P.up.size();
    for (i64 q = s.uptr[k]; q < s.uptr[k + 1]; ++q) {
      const fc::Update& u = s.upd[q];
      fc::LdlUpdate U;
      U.k = u.k;
      U.p1 = (int32_t)u.p1;
      U.p2 = (int32_t)u.p2;
      U.rel = (i64)P.rel.size();
      const i64 km = s.rptr[u.k + 1] - s.rptr[u.k];
      i64 at = 0;
      for (i64 i = u.p1; i < km; ++i) {
        const i64 r = s.rows[s.rptr[u.k] + i];
        while (at < J.m && s.rows[s.rptr[k] + at] < r) ++at;
        if (at >= J.m || s.rows[s.rptr[k] + at] != r) throw std::runtime_error("an update row outside the structure");
        // the update's columns must be columns of the target (subtract, cholesky.rs:668-671)
        if (i < u.p2 && at >= J.w) throw std::runtime_error("an update column outside the supernode");
        P.rel.push_back((int32_t)at);
      }
      P.up.push_back(U);
    }
    J.upd1 = (i64)P.up.size();
  }
  std::vector<i64> inv(n);
  for (i64 i = 0; i < n; ++i) inv[s.perm[i]] = i;
  const i64 nnz = (i64)s.upper_idx.size();
  P.pos.resize(nnz);
  P.coord_r.resize(nnz);
  P.coord_c.resize(nnz);
  for (i64 k = 0; k < n; ++k)
    for (i64 e = s.upper_ptr[k]; e < s.upper_ptr[k + 1]; ++e) {
      const i64 a = inv[s.upper_idx[e]], b = inv[k];
      const i64 r = std::max(a, b), c = std::min(a, b);
      const i64 sj = snode[c];
      const fc::LdlSupernode& J = P.sn[sj];
      const i64* rows = s.rows.data() + s.rptr[sj];
      const i64 lr = std::lower_bound(rows, rows + J.m, r) - rows, lc = c - J.first;
      if (lr >= J.m || rows[lr] != r || lr < lc) throw std::runtime_error("an entry of A outside the structure");
      P.pos[e] = J.voff + lc * J.m - lc * (lc - 1 > 0 ? lc - 1 : 0) / 2 + lr - lc;
      P.coord_r[e] = r;
      P.coord_c[e] = c;
    }
  std::vector<i64> level(ns, 0);
  i64 levels = 0;
  for (i64 k = 0; k < ns; ++k) {  // children come before parents in postorder
    const i64 pa = s.sparent[k];
    if (pa >= 0 && pa < ns) level[pa] = std::max(level[pa], level[k] + 1);
    levels = std::max(levels, level[k] + 1);
  }
  P.lvl_ptr.assign(levels + 1, 0);
  for (i64 k = 0; k < ns; ++k) ++P.lvl_ptr[level[k] + 1];
  for (i64 l = 0; l < levels; ++l) P.lvl_ptr[l + 1] += P.lvl_ptr[l];
  P.lvl_sn.resize(ns);
  std::vector<i64> cur(P.lvl_ptr.begin(), P.lvl_ptr.end() - 1);
  for (i64 k = 0; k < ns; ++k)
    if (P.sn[k].w <= LDL_WIDE) P.lvl_sn[cur[level[k]]++] = k;
  P.lvl_narrow.resize(levels);
  for (i64 l = 0; l < levels; ++l) P.lvl_narrow[l] = cur[l] - P.lvl_ptr[l];
  for (i64 k = 0; k < ns; ++k)
    if (P.sn[k].w > LDL_WIDE) P.lvl_sn[cur[level[k]]++] = k;
  P.pan_ptr.assign(levels + 1, 0);
  P.pan_tile_ptr.assign(1, 0);
  for (i64 l = 0; l < levels; ++l) {
    i64 wmax = 0;
    for (i64 q = P.lvl_ptr[l] + P.lvl_narrow[l]; q < P.lvl_ptr[l + 1]; ++q) wmax = std::max<i64>(wmax, P.sn[P.lvl_sn[q]].w);
    // each panel ends at or after min(start + LDL_PANEL, w), so panel p ends at or after
    // min((p + 1) LDL_PANEL, w) and a tile ending there or before has no column left to update
    const i64 panels = (wmax + fc::LDL_PANEL - 1) / fc::LDL_PANEL;
    for (i64 p = 0; p < panels; ++p) {
      const i64 done = (p + 1) * fc::LDL_PANEL;
      for (i64 q = P.lvl_ptr[l] + P.lvl_narrow[l]; q < P.lvl_ptr[l + 1]; ++q) {
        const i64 k = P.lvl_sn[q], w = P.sn[k].w;
        for (i64 b = 0; b < w; b += LDL_TILE_COLS) {
          const i64 e = std::min<i64>(b + LDL_TILE_COLS, w);
          if (e > done) ldl_tiles(P.pan_tiles, k, b, e, P.sn[k].m);
        }
      }
      P.pan_tile_ptr.push_back((i64)P.pan_tiles.size());
    }
    P.pan_ptr[l + 1] = P.pan_ptr[l] + panels;
  }
  P.tile_ptr.assign(levels + 1, 0);
  for (i64 l = 0; l < levels; ++l) {
    for (i64 q = P.lvl_ptr[l]; q < P.lvl_ptr[l + 1]; ++q) {
      const i64 k = P.lvl_sn[q];
      if (P.sn[k].upd0 == P.sn[k].upd1) continue;  // no update to apply
      for (i64 b = 0; b < P.sn[k].w; b += LDL_TILE_COLS)
        ldl_tiles(P.tiles, k, b, std::min<i64>(b + LDL_TILE_COLS, P.sn[k].w), P.sn[k].m);
    }
    P.tile_ptr[l + 1] = (i64)P.tiles.size();
  }
  P.ftile_ptr.assign(levels + 1, 0);
  for (i64 l = 0; l < levels; ++l) {
    for (i64 q = P.lvl_ptr[l]; q < P.lvl_ptr[l + 1]; ++q) {
      const i64 k = P.lvl_sn[q];
      if (P.sn[k].upd0 == P.sn[k].upd1) continue;
      for (i64 b = 0; b < P.sn[k].w; b += LDL_GATHER_COLS)
        P.ftiles.push_back(fc::LdlTile{k, b, std::min<i64>(b + LDL_GATHER_COLS, P.sn[k].w), 0, 0});
    }
    P.ftile_ptr[l + 1] = (i64)P.ftiles.size();
  }
  return P;
}

struct LanczosError {
  std::string msg;
};

// The device side of one Lanczos run: the basis Q and its M products MQ (n x cap), a block W and
// MW (n x 8), reduction buffers.
struct LanczosWork {
  const fc_matrix* M;
  hipStream_t st;
  i64 n;
  i64 cap;
  double *Q = nullptr, *MQ = nullptr, *W = nullptr, *MW = nullptr, *part = nullptr, *C = nullptr;
  int64_t* at = nullptr;
  int chunks = 1;
  std::vector<double> host;

  ~LanczosWork() {
    for (void* p : {(void*)Q, (void*)MQ, (void*)W, (void*)MW, (void*)part, (void*)C, (void*)at}) dfree(p);
  }
  double* q(i64 j) { return Q + j * n; }
  double* mq(i64 j) { return MQ + j * n; }
  // Y = M X for k columns (Csr::mul by rows, sparse.rs:42-50)
  void mass(const double* x, double* y, int k) {
    fc::k_spmm<<<blocks_for(n * 8), 256, 0, st>>>(n, M->ptr, M->idx, M->val, x, y, k, n);
  }
  // host C (m x b, row-major) = A^T B: A n x m (leading dimension n), B n x b
  const std::vector<double>& dots(const double* A, int m, const double* B, int b) {
    host.assign((size_t)m * b, 0.0);
    if (m == 0 || b == 0) return host;
    dim3 grid(chunks, (m + fc::TS_COLS - 1) / fc::TS_COLS);
    fc::k_ts_dots<<<grid, 256, 0, st>>>(n, A, n, m, B, n, b, part);
    fc::k_ts_dots_sum<<<(m * b + 255) / 256, 256, 0, st>>>(part, chunks, m, b, C);
    HIP_CHECK(hipMemcpyAsync(host.data(), C, sizeof(double) * m * b, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    return host;
  }
  // B = beta B + alpha A c for b columns; c host m x b row-major (uploaded in pieces the
  // update kernel's LDS holds)
  void update(const double* A, int m, const std::vector<double>& c, double* B, int b, double alpha, double beta) {
    const int per = std::max(1, fc::TS_MAXC / std::max(b, 1));
    for (int i0 = 0; i0 < m || (m == 0 && i0 == 0); i0 += per) {
      const int mm = std::min(per, m - i0);
      if (mm <= 0) {
        if (beta != 1.0) fc::k_ts_update<<<blocks_for(n), 256, 0, st>>>(n, A, n, 0, C, b, B, n, b, alpha, beta);
        break;
      }
      HIP_CHECK(hipMemcpyAsync(C, c.data() + (size_t)i0 * b, sizeof(double) * mm * b, hipMemcpyHostToDevice, st));
      fc::k_ts_update<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, A + (i64)i0 * n, n, mm, C, b, B, n, b, alpha,
                                                                    i0 == 0 ? beta : 1.0);
    }
  }
  void scale(double* x, double s) { fc::k_scale<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, x, s); }
  void copy(const double* x, double* y) { fc::k_copy_col<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, x, y); }
  double mnorm2(const double* x, const double* mx) { return std::max(dots(x, 1, mx, 1)[0], 0.0); }
  void random_into(double* x, XorShift& rng) {
    std::vector<double> v(n);
    for (i64 r = 0; r < n; ++r) v[r] = rng.uniform();
    HIP_CHECK(hipMemcpyAsync(x, v.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    HIP_CHECK(hipStreamSynchronize(st));
  }
};

// Orthonormalize the k columns of W against the basis Q[:, 0..mb) and one another in the M
// inner product (eigen.rs:231-279): the basis part as two passes of block classical Gram-Schmidt
// (C = MQ^T W, W -= Q C), then column by column against the block's earlier vectors, twice,
// recording r[j][i] (the coefficient of column i on new vector j; r[i][i] its norm). A column
// whose norm falls below 1e-8 of its norm before is replaced by a fresh random direction (at most
// three times) and keeps a zero coefficient. The new vectors go to Q[:, mb..), their M products
// to MQ[:, mb..). Returns how many were kept.
int orthonormalize_block(LanczosWork& L, int k, i64 mb, XorShift& rng, std::vector<double>& r) {
  r.assign((size_t)k * k, 0.0);
  L.mass(L.W, L.MW, k);
  std::vector<double> before(k);
  {
    const std::vector<double>& g = L.dots(L.W, k, L.MW, k);
    for (int i = 0; i < k; ++i) before[i] = std::sqrt(std::max(g[(size_t)i * k + i], 0.0));
  }
  auto project_basis = [&](double* w, int cols) {
    for (int pass = 0; pass < 2 && mb > 0; ++pass) {
      const std::vector<double> c = L.dots(L.MQ, (int)mb, w, cols);
      L.update(L.Q, (int)mb, c, w, cols, -1.0, 1.0);
    }
  };
  project_basis(L.W, k);
  int kept = 0;
  for (int i = 0; i < k; ++i) {
    double* v = L.W + (i64)i * L.n;
    bool fresh = false;
    for (int attempt = 0; attempt < 4; ++attempt) {
      if (fresh) {
        L.random_into(v, rng);
        L.mass(v, L.MW, 1);
        before[i] = std::sqrt(L.mnorm2(v, L.MW));
        project_basis(v, 1);
      }
      for (int pass = 0; pass < 2 && kept > 0; ++pass) {
        const std::vector<double> c = L.dots(L.mq(mb), kept, v, 1);
        L.update(L.q(mb), kept, c, v, 1, -1.0, 1.0);
        if (!fresh)
          for (int j = 0; j < kept; ++j) r[(size_t)j * k + i] += c[j];
      }
      double* mv = L.mq(mb + kept);
      L.mass(v, mv, 1);
      const double norm = std::sqrt(L.mnorm2(v, mv));
      if (norm > 1e-8 * before[i] && norm > 0.0) {
        if (!fresh) r[(size_t)i * k + i] = norm;
        L.copy(v, L.q(mb + kept));
        L.scale(L.q(mb + kept), 1.0 / norm);
        L.scale(mv, 1.0 / norm);
        ++kept;
        break;
      }
      fresh = true;
    }
  }
  return kept;
}

struct LanczosResult {
  std::vector<double> values, negative;
  i64 basis = 0;
};

// eigen.rs:281-466 with the operator (K - shift M)^{-1} M: the factor's block solves and the
// mass matrix's products on the device; the projected matrix T and its eigenproblem on the host.
// solve(W, b) overwrites the n x b block W (device, leading dimension n) with (K - shift M)^-1 W
// on stream st; *cancel (if given) is polled once per block step.
LanczosResult lanczos_run(i64 n, const std::function<void(double*, int)>& solve, const volatile int32_t* cancel,
                          const fc_matrix* M, double shift, const fc_lanczos_opts& o, double* vectors, hipStream_t st) {
  const int b = (int)std::max<i64>(1, std::min<i64>(o.block, n));
  const i64 max_basis = std::max<i64>(std::min<i64>((i64)o.max_basis, n), b);
  const i64 cap = max_basis + b;
  LanczosWork L{M, st, n, cap};
  L.chunks = (int)std::max<i64>(1, (n + fc::TS_ROWS - 1) / fc::TS_ROWS);
  L.Q = dalloc<double>((size_t)n * cap);
  L.MQ = dalloc<double>((size_t)n * cap);
  L.W = dalloc<double>((size_t)n * fc::TS_MAXB);
  L.MW = dalloc<double>((size_t)n * fc::TS_MAXB);
  L.part = dalloc<double>((size_t)L.chunks * cap * fc::TS_MAXB);
  L.C = dalloc<double>((size_t)std::max<i64>(cap * fc::TS_MAXB, fc::TS_MAXC));
  L.at = dalloc<int64_t>((size_t)L.chunks);
  XorShift rng(0x9E3779B97F4A7C15ULL);
  {
    std::vector<double> start((size_t)n * b);
    for (size_t i = 0; i < start.size(); ++i) start[i] = rng.uniform();  // column by column
    HIP_CHECK(hipMemcpyAsync(L.W, start.data(), sizeof(double) * start.size(), hipMemcpyHostToDevice, st));
  }
  std::vector<double> r;
  if (orthonormalize_block(L, b, 0, rng, r) < b) throw LanczosError{"could not build a Lanczos start block"};
  std::vector<double> prev_b;
  i64 next_check = std::max<i64>((i64)o.count + b, 2 * b);
  std::vector<double> T;  // m x m row-major
  i64 m = 0;
  for (;;) {
    if (cancel && *cancel) throw LanczosError{"cancelled"};
    const i64 j0 = m;
    // W = Op Q_j: the block's M products solved together
    HIP_CHECK(hipMemcpyAsync(L.W, L.mq(j0), sizeof(double) * n * b, hipMemcpyDeviceToDevice, st));
    solve(L.W, b);
    // A_j = MQ_j^T W, symmetrized
    std::vector<double> a = L.dots(L.mq(j0), b, L.W, b);
    for (int i = 0; i < b; ++i)
      for (int j = i + 1; j < b; ++j) {
        const double sym = 0.5 * (a[(size_t)i * b + j] + a[(size_t)j * b + i]);
        a[(size_t)i * b + j] = a[(size_t)j * b + i] = sym;
      }
    m = j0 + b;
    {
      std::vector<double> T2((size_t)m * m, 0.0);
      for (i64 i = 0; i < j0; ++i)
        for (i64 j = 0; j < j0; ++j) T2[(size_t)i * m + j] = T[(size_t)i * j0 + j];
      T.swap(T2);
    }
    for (int i = 0; i < b; ++i)
      for (int j = 0; j < b; ++j) T[(size_t)(j0 + i) * m + j0 + j] = a[(size_t)i * b + j];
    if (!prev_b.empty())
      for (int i = 0; i < b; ++i)
        for (int j = 0; j < b; ++j) {
          T[(size_t)(j0 + i) * m + j0 - b + j] = prev_b[(size_t)i * b + j];
          T[(size_t)(j0 - b + j) * m + j0 + i] = prev_b[(size_t)i * b + j];
        }
    const bool full = m >= max_basis || m + b > n;
    int kept = b;
    std::vector<double> bnext((size_t)b * b, 0.0);
    if (!full) kept = orthonormalize_block(L, b, m, rng, bnext);
    if (m >= next_check || full || kept < b) {
      next_check = m + std::max<i64>(m / 4, b);
      std::vector<double> dense(T), theta, s;
      if (!dense_symmetric_eigen((int)m, dense, theta, s)) throw LanczosError{"the tridiagonal QR iteration did not converge"};
      std::vector<int> idx(m);
      for (i64 i = 0; i < m; ++i) idx[i] = (int)i;
      std::sort(idx.begin(), idx.end(), [&](int x, int y) { return theta[x] > theta[y] || (theta[x] == theta[y] && x < y); });
      auto residual = [&](int i) {
        if (full) return 0.0;
        double res2 = 0.0;
        for (int rr = 0; rr < b; ++rr) {
          double v = 0.0;
          for (int c = 0; c < b; ++c) v += bnext[(size_t)rr * b + c] * s[(size_t)(m - b + c) * m + i];
          res2 += v * v;
        }
        return std::sqrt(res2);
      };
      bool done = false;
      std::vector<int> accepted;
      i64 found = 0;
      for (int i : idx) {
        const double th = theta[i];
        if (!(th > 0.0)) break;
        if (residual(i) > o.tolerance * th) break;
        const double lambda = shift + 1.0 / th;
        if (lambda > o.max) {
          done = true;
          break;
        }
        if (lambda >= o.min) {
          accepted.push_back(i);
          if (++found == (i64)o.count) {
            done = true;
            break;
          }
        }
      }
      if (full && !done) {
        if (m >= n) {
          done = true;
        } else {
          char msg[200];
          std::snprintf(msg, sizeof msg, "found %lld of %llu modes after %lld Lanczos vectors (the cap)", (long long)found,
                        (unsigned long long)o.count, (long long)m);
          throw LanczosError{msg};
        }
      }
      if (done) {
        LanczosResult res;
        res.basis = m;
        for (auto it = idx.rbegin(); it != idx.rend(); ++it) {
          const double th = theta[*it];
          if (!(th < 0.0)) break;
          if (residual(*it) > o.tolerance * std::fabs(th)) break;
          res.negative.push_back(shift + 1.0 / th);
        }
        for (size_t a2 = 0; a2 < accepted.size(); ++a2) {
          res.values.push_back(shift + 1.0 / theta[accepted[a2]]);
          if (!vectors) continue;
          // x = Q s, M-normalized, the largest component positive
          std::vector<double> c((size_t)m);
          for (i64 k2 = 0; k2 < m; ++k2) c[k2] = s[(size_t)k2 * m + accepted[a2]];
          double* x = vectors + (i64)a2 * n;
          L.update(L.Q, (int)m, c, x, 1, 1.0, 0.0);
          L.mass(x, L.MW, 1);
          const double norm = std::sqrt(L.mnorm2(x, L.MW));
          if (norm > 0.0) L.scale(x, 1.0 / norm);
          fc::k_argmax_abs<<<L.chunks, 256, 0, st>>>(n, x, L.at);
          std::vector<int64_t> at(L.chunks);
          HIP_CHECK(hipMemcpyAsync(at.data(), L.at, sizeof(int64_t) * L.chunks, hipMemcpyDeviceToHost, st));
          HIP_CHECK(hipStreamSynchronize(st));
          double big = 0.0;
          for (int64_t q : at) {
            double v = 0.0;
            HIP_CHECK(hipMemcpy(&v, x + q, sizeof(double), hipMemcpyDeviceToHost));
            if (std::fabs(v) > std::fabs(big)) big = v;
          }
          if (big < 0.0) L.scale(x, -1.0);
        }
        HIP_CHECK(hipStreamSynchronize(st));
        return res;
      }
    }
    if (full) throw LanczosError{"the Lanczos basis is exhausted"};
    if (kept < b) throw LanczosError{"the Lanczos block lost rank"};
    prev_b = bnext;
  }
}

}  // namespace

// ------------------------------------------------------------------------------------------

extern "C" {

const char* fc_version(void) { return VERSION; }

}  // extern "C"

namespace {
double host_mod(double x) { return std::fabs(x); }
double host_mod(fc::zd x) { return std::hypot(x.re, x.im); }  // Complex64::norm

template <typename T>
int ldlt_numeric(const fc_symbolic* h, const T* upper_val, uint64_t nnz, int32_t device, fc_ldlt** out,
                 fc_error* err) {
  constexpr bool REAL = std::is_same<T, double>::value;
  if (out) *out = nullptr;
  if (!h || !out || (nnz && !upper_val)) {
    set_error(err, FC_INVALID_ARGUMENT, "bad arguments to fc_ldlt_numeric");
    return FC_INVALID_ARGUMENT;
  }
  const fc::Symbolic& s = h->s;
  if ((i64)nnz != (i64)s.upper_idx.size()) {
    set_error(err, FC_INVALID_ARGUMENT, "the matrix's pattern differs from the analysed one (entry count)");
    return FC_INVALID_ARGUMENT;
  }
  std::unique_ptr<fc_ldlt> F(new fc_ldlt());
  F->complex = !REAL;
  F->device = device;
  F->sym_keep = keep(h);
  F->sym = h;
  F->n = s.n;
  F->ns = s.ns();
  int cur = 0;
  (void)hipGetDevice(&cur);
  int64_t failed = -1;
  std::string plan_error;
  int st = guarded(err, "not enough device memory for the LDL^T factor", [&] {
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
    hipStream_t stq = F->stream;
    std::shared_ptr<const LdltPlan> Pp;
    {
      std::lock_guard<std::mutex> g(h->ldlt_mu);
      if (!h->ldlt_cache) {
        try {
          h->ldlt_cache = std::make_shared<const LdltPlan>(ldlt_plan(s));
        } catch (const std::runtime_error& e) {
          plan_error = e.what();
          return;
        }
      }
      Pp = h->ldlt_cache;
    }
    const LdltPlan& P = *Pp;
    F->vals = P.vals;
    F->max_w = P.max_w;
    // the largest magnitude of each row (= column) of A, in postorder (cholesky.rs:1399-1414; a
    // NaN leaves it as it is, as f64::max does)
    std::vector<double> scale(std::max<i64>(s.n, 1), 0.0);
    for (i64 e = 0; e < (i64)nnz; ++e) {
      const double x = host_mod(upper_val[e]);
      if (x > scale[P.coord_r[e]]) scale[P.coord_r[e]] = x;
      if (x > scale[P.coord_c[e]]) scale[P.coord_c[e]] = x;
    }
    F->S = dupload(P.sn, stq);
    F->rows = dupload(s.rows, stq);
    F->perm = dupload(s.perm, stq);
    F->val = dalloc<double>((size_t)std::max<i64>(P.vals, 1) * (sizeof(T) / sizeof(double)));
    T* fval = reinterpret_cast<T*>(F->val);
    F->piv = dalloc<int8_t>((size_t)std::max<i64>(s.n, 1));
    F->order = dalloc<int32_t>((size_t)std::max<i64>(s.n, 1));
    fc::LdlUpdate* dU = dupload(P.up, stq);
    int32_t* drel = dupload(P.rel, stq);
    F->U = dU;
    F->rel = drel;
    F->ftiles = dupload(P.ftiles, stq);
    F->ftile_ptr = P.ftile_ptr;
    int64_t* dpos = dupload(P.pos, stq);
    int64_t* dsn = dupload(P.lvl_sn, stq);
    F->sn = dsn;
    F->lvl_ptr = P.lvl_ptr;
    double* dscale = dupload(scale, stq);
    T* daval = dalloc<T>((size_t)std::max<uint64_t>(nnz, 1));
    T* dwork = dalloc<T>((size_t)std::max<i64>(P.work, 1));
    int64_t* dfail = dalloc<int64_t>((size_t)std::max<i64>(F->ns, 1));
    int32_t* dtwo = dalloc<int32_t>((size_t)std::max<i64>(F->ns, 1));
    int64_t* dstate = dalloc<int64_t>((size_t)std::max<i64>(4 * F->ns, 1));
    struct Free {
      std::vector<void*> p;
      ~Free() {
        for (void* q : p) dfree(q);
      }
    } tmp{{dpos, dscale, daval, dwork, dfail, dtwo, dstate}};
    if (nnz) HIP_CHECK(hipMemcpyAsync(daval, upper_val, sizeof(T) * nnz, hipMemcpyHostToDevice, stq));
    if (P.vals) fc::k_ldlt_zero<<<blocks_for(P.vals), 256, 0, stq>>>(P.vals, fval);
    if (nnz) fc::k_ldlt_scatter<<<(unsigned)((nnz + 255) / 256), 256, 0, stq>>>((i64)nnz, daval, dpos, fval);
    fc::LdlTile* dtiles = dupload(P.tiles, stq);
    tmp.p.push_back(dtiles);
    fc::LdlTile* dptiles = dupload(P.pan_tiles, stq);
    tmp.p.push_back(dptiles);
    // the LDS budget of the panel kernels' work columns (FC_LDLT_WORK_LDS bytes, 0..LDL_WORK_LDS:
    // a tuning knob; a launch whose tallest supernode's two columns do not fit keeps them in
    // device memory instead)
    size_t work_lds = LDL_WORK_LDS;
    if (const char* e = std::getenv("FC_LDLT_WORK_LDS")) {
      const long v = std::strtol(e, nullptr, 10);
      if (v >= 0 && (size_t)v <= LDL_WORK_LDS) work_lds = (size_t)v;
    }
    for (size_t l = 0; l + 1 < P.lvl_ptr.size(); ++l) {
      const i64 narrow = P.lvl_narrow[l], wide = P.lvl_ptr[l + 1] - P.lvl_ptr[l] - narrow;
      const i64 nt = P.tile_ptr[l + 1] - P.tile_ptr[l];
      const int64_t* lsn = dsn + P.lvl_ptr[l];
      if (nt > 0)
        fc::k_ldlt_assemble<<<(unsigned)nt, 256, 0, stq>>>(dtiles + P.tile_ptr[l], F->S, dU, drel, fval, F->piv);
      // a level of few supernodes gives each 1024 threads; a crowded one 256
      // the work columns of a launch's supernodes in LDS when two columns of its tallest fit
      auto lds_for = [&](i64 q0, i64 q1, int* rows) {
        i64 mm = 0;
        for (i64 q = q0; q < q1; ++q) mm = std::max<i64>(mm, P.sn[P.lvl_sn[q]].m);
        const size_t bytes = 2 * (size_t)mm * sizeof(T);
        *rows = bytes <= work_lds ? (int)mm : 0;
        return bytes <= work_lds ? (unsigned)bytes : 0u;
      };
      int nrows = 0, wrows = 0;
      const unsigned nlds = lds_for(P.lvl_ptr[l], P.lvl_ptr[l] + narrow, &nrows);
      const unsigned wlds = lds_for(P.lvl_ptr[l] + narrow, P.lvl_ptr[l + 1], &wrows);
      if (narrow > 0)
        fc::k_ldlt_supernode<T, false><<<(unsigned)narrow, narrow < 256 ? LDL_PANEL_THREADS : 256, nlds, stq>>>(
            lsn, F->S, dU, drel, dscale, fval, F->piv, F->order, dwork, dfail, dtwo, dstate, nrows);
      if (wide > 0) {
        fc::k_ldlt_state_init<<<(unsigned)wide, 256, 0, stq>>>(lsn + narrow, F->S, F->piv, F->order, dfail, dtwo, dstate);
        for (i64 i = P.pan_ptr[l]; i < P.pan_ptr[l + 1]; ++i) {
          fc::k_ldlt_supernode<T, true><<<(unsigned)wide, LDL_PANEL_THREADS, wlds, stq>>>(lsn + narrow, F->S, dU, drel, dscale, fval,
                                                                         F->piv, F->order, dwork, dfail, dtwo, dstate,
                                                                         wrows);
          const i64 pt = P.pan_tile_ptr[i + 1] - P.pan_tile_ptr[i];
          if (pt > 0)
            fc::k_ldlt_trailing<<<(unsigned)pt, 256, 0, stq>>>(dptiles + P.pan_tile_ptr[i], F->S, fval, F->piv, dstate);
        }
      }
    }
    HIP_CHECK(hipGetLastError());
    std::vector<int64_t> fails(std::max<i64>(F->ns, 1));
    std::vector<int32_t> twos(std::max<i64>(F->ns, 1));
    if (F->ns) {
      HIP_CHECK(hipMemcpyAsync(fails.data(), dfail, sizeof(int64_t) * F->ns, hipMemcpyDeviceToHost, stq));
      HIP_CHECK(hipMemcpyAsync(twos.data(), dtwo, sizeof(int32_t) * F->ns, hipMemcpyDeviceToHost, stq));
    }
    HIP_CHECK(hipStreamSynchronize(stq));
    for (i64 k = 0; k < F->ns; ++k) {
      if (fails[k] >= 0 && (failed < 0 || fails[k] < failed)) failed = fails[k];
      F->two_by_two += twos[k];
    }
    if (REAL && failed < 0) {
      // inertia from D (ldlt.rs:412-445)
      std::vector<double> v((size_t)std::max<i64>(2 * s.n, 1));
      std::vector<int8_t> pv((size_t)std::max<i64>(s.n, 1));
      double* ddiag = dalloc<double>((size_t)std::max<i64>(2 * s.n, 1));
      tmp.p.push_back(ddiag);
      if (F->ns) fc::k_ldlt_diag<<<(unsigned)F->ns, 64, 0, stq>>>(F->S, F->val, F->piv, ddiag);
      HIP_CHECK(hipGetLastError());
      if (s.n) HIP_CHECK(hipMemcpyAsync(v.data(), ddiag, sizeof(double) * 2 * s.n, hipMemcpyDeviceToHost, stq));
      if (s.n) HIP_CHECK(hipMemcpyAsync(pv.data(), F->piv, s.n, hipMemcpyDeviceToHost, stq));
      HIP_CHECK(hipStreamSynchronize(stq));
      for (i64 k = 0; k < F->ns; ++k) {
        const fc::LdlSupernode& J = P.sn[k];
        // D(r, c) for r = c or c + 1 from the extracted blocks
        auto d = [&](i64 r, i64 c) { return r == c ? v[2 * (J.first + c)] : v[2 * (J.first + c) + 1]; };
        i64 c = 0;
        while (c < J.w) {
          if (pv[J.first + c] == fc::PIV_FIRST) {
            const double d11 = d(c, c), d21 = d(c + 1, c), d22 = d(c + 1, c + 1);
            const double det = d11 * d22 - d21 * d21;
            if (det < 0.0) {
              F->positive += 1;
              F->negative += 1;
            } else if (d11 + d22 > 0.0) {
              F->positive += 2;
            } else {
              F->negative += 2;
            }
            c += 2;
          } else {
            if (d(c, c) > 0.0) F->positive += 1;
            else F->negative += 1;
            c += 1;
          }
        }
      }
    }
  });
  (void)hipSetDevice(cur);
  if (st != FC_OK) return st;
  if (!plan_error.empty()) {
    set_error(err, FC_INTERNAL, plan_error);
    return FC_INTERNAL;
  }
  if (failed >= 0) {
    const uint64_t col = (uint64_t)s.perm[failed];
    set_error(err, FC_SINGULAR, "the matrix is singular (no acceptable pivot in its supernode)", col, s.entries);
    return FC_SINGULAR;
  }
  *out = F.release();
  set_error(err, FC_OK, "");
  return FC_OK;
}
}  // namespace

extern "C" {

int fc_ldlt_numeric(const fc_symbolic* h, const double* upper_val, uint64_t nnz, int32_t device, fc_ldlt** out,
                    fc_error* err) {
  return ldlt_numeric<double>(h, upper_val, nnz, device, out, err);
}

int fc_zldlt_numeric(const fc_symbolic* h, const double* upper_val, uint64_t nnz, int32_t device, fc_ldlt** out,
                     fc_error* err) {
  return ldlt_numeric<fc::zd>(h, reinterpret_cast<const fc::zd*>(upper_val), nnz, device, out, err);
}

}  // extern "C"

namespace {
// The forward and backward passes on x (postorder, device) with work vectors z and y of n entries.
template <typename T>
void ldlt_passes(const fc_ldlt* f, const T* fv, T* dx, T* dz, T* dyw, hipStream_t st) {
  // rows of one backward column staged in LDS at a time (FC_LDLT_SOLVE_ROWS, 16..3840: a tuning
  // knob; larger supernodes take the chunked branch)
  int cap = fc::LDL_BW_CAP;
  if (const char* e = std::getenv("FC_LDLT_SOLVE_ROWS")) {
    const long v = std::strtol(e, nullptr, 10);
    if (v >= 16 && v <= fc::LDL_BW_CAP) cap = (int)v;
  }
  for (size_t l = 0; l + 1 < f->lvl_ptr.size(); ++l) {
    const i64 nt = f->ftile_ptr[l + 1] - f->ftile_ptr[l], cnt = f->lvl_ptr[l + 1] - f->lvl_ptr[l];
    if (nt > 0)
      fc::k_ldlt_fwd_gather<<<(unsigned)nt, 256, 0, st>>>(f->ftiles + f->ftile_ptr[l], f->S, f->U, f->rel, fv, dz, dx);
    if (cnt > 0)
      fc::k_ldlt_fwd_own<<<(unsigned)cnt, 256, 0, st>>>(f->sn + f->lvl_ptr[l], f->S, fv, f->piv, f->order, dx, dyw, dz);
  }
  for (i64 l = (i64)f->lvl_ptr.size() - 2; l >= 0; --l) {
    const i64 cnt = f->lvl_ptr[l + 1] - f->lvl_ptr[l];
    if (cnt > 0)
      fc::k_ldlt_solve_bwd<<<(unsigned)cnt, 256, 0, st>>>(f->sn + f->lvl_ptr[l], f->S, f->rows, fv, f->piv, f->order, dx,
                                                         dyw, cap);
  }
}

// Real LDL^T solves of the n x nb block W (device, leading dimension n) in place on stream st,
// each column as a lone solve (the same bits): permuted into postorder, the passes, permuted back.
// `work` holds the three work vectors (allocated on first use, kept by the caller).
void ldlt_solve_block_device(const fc_ldlt* f, double* W, int nb, hipStream_t st, Buffer& work) {
  const i64 n = f->n;
  double* dx = (double*)work.get((size_t)3 * n * sizeof(double));
  double* dz = dx + n;
  double* dyw = dz + n;
  for (int c = 0; c < nb; ++c) {
    double* col = W + (size_t)c * n;
    fc::k_ldlt_permute<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, f->perm, col, dx, false);
    ldlt_passes(f, f->val, dx, dz, dyw, st);
    fc::k_ldlt_permute<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(n, f->perm, dx, col, true);
  }
  HIP_CHECK(hipGetLastError());
}

template <typename T>
int ldlt_solve(const fc_ldlt* f, T* b, uint64_t n, fc_error* err) {
  if (!f || (n && !b)) {
    set_error(err, FC_INVALID_ARGUMENT, "bad arguments to fc_ldlt_solve");
    return FC_INVALID_ARGUMENT;
  }
  if (f->complex != !std::is_same<T, double>::value) {
    set_error(err, FC_INVALID_ARGUMENT, "a real solve of a complex factor or the reverse");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  if (n == 0) {
    set_error(err, FC_OK, "");
    return FC_OK;
  }
  const std::vector<i64>& perm = f->sym->s.perm;
  return guarded(err, "not enough device memory for the solve", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    const T* fv = reinterpret_cast<const T*>(f->val);
    std::vector<T> x(n);
    for (uint64_t i = 0; i < n; ++i) x[i] = b[perm[i]];
    T* dx = dalloc<T>(n);
    T* dz = dalloc<T>(n);
    HIP_CHECK(hipMemcpyAsync(dx, x.data(), sizeof(T) * n, hipMemcpyHostToDevice, f->stream));
    T* dyw = dalloc<T>(n);
    // rows of one backward column staged in LDS at a time (FC_LDLT_SOLVE_ROWS, 16..3840: a tuning
    // knob; larger supernodes take the chunked branch)
    ldlt_passes(f, fv, dx, dz, dyw, f->stream);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(x.data(), dx, sizeof(T) * n, hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
    dfree(dx);
    dfree(dz);
    dfree(dyw);
    for (uint64_t i = 0; i < n; ++i) b[perm[i]] = x[i];
  });
}
}  // namespace

// The schedule between its C and C++ forms.
namespace {
bool schedule_ok(const fc_schedule& c) {
  auto in = [](int32_t v, int32_t lo, int32_t hi) { return v >= lo && v <= hi; };
  return in(c.slab, 1, 16) && in(c.split_target, 1, 65536) && in(c.split_side_fp64, 1, 65536) &&
         in(c.split_side_fp32, 1, 65536);
}
Schedule to_schedule(const fc_schedule& c) {
  Schedule r;
  r.slab = c.slab;
  r.split_target = c.split_target;
  r.split_side_fp64 = c.split_side_fp64;
  r.split_side_fp32 = c.split_side_fp32;
  return r;
}
fc_schedule from_schedule(const Schedule& r) {
  fc_schedule c;
  c.slab = r.slab;
  c.split_target = r.split_target;
  c.split_side_fp64 = r.split_side_fp64;
  c.split_side_fp32 = r.split_side_fp32;
  return c;
}
bool same(const Schedule& a, const Schedule& b) {
  return a.slab == b.slab && a.split_target == b.split_target && a.split_side_fp64 == b.split_side_fp64 &&
         a.split_side_fp32 == b.split_side_fp32;
}
}  // namespace

extern "C" {

int fc_ldlt_solve(const fc_ldlt* f, double* b, uint64_t n, fc_error* err) { return ldlt_solve<double>(f, b, n, err); }

int fc_zldlt_solve(const fc_ldlt* f, double* b, uint64_t n, fc_error* err) {
  return ldlt_solve<fc::zd>(f, reinterpret_cast<fc::zd*>(b), n, err);
}

int fc_ldlt_inertia(const fc_ldlt* f, uint64_t* positive, uint64_t* negative, uint64_t* two_by_two) {
  if (!f) return FC_INVALID_ARGUMENT;
  if (positive) *positive = (uint64_t)f->positive;
  if (negative) *negative = (uint64_t)f->negative;
  if (two_by_two) *two_by_two = (uint64_t)f->two_by_two;
  return FC_OK;
}

int fc_ldlt_export(const fc_ldlt* f, double* out, uint64_t len, fc_error* err) {
  if (!f || (len && !out) || (i64)len != f->vals * (f->complex ? 2 : 1)) {
    set_error(err, FC_INVALID_ARGUMENT, "the output needs the factor's entry count (twice for a complex factor)");
    return FC_INVALID_ARGUMENT;
  }
  return guarded(err, "device", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    if (len) HIP_CHECK(hipMemcpyAsync(out, f->val, sizeof(double) * len, hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
  });
}

void fc_ldlt_free(fc_ldlt* f) { delete f; }

int fc_lanczos_device(const fc_factor* f, const fc_matrix* m, double shift, const fc_lanczos_opts* opts,
                      double* values, double* vectors_device, fc_lanczos_report* report, fc_error* err) {
  if (!f || !f->valid || !m || !opts || (opts->count && !values) || m->n != f->sym->s.n || opts->block < 1 ||
      opts->block > 8 || !(opts->tolerance > 0.0) || !std::isfinite(shift)) {
    set_error(err, FC_INVALID_ARGUMENT, "bad arguments to fc_lanczos_device");
    return FC_INVALID_ARGUMENT;
  }
  if (f->sym->s.n == 0) {
    set_error(err, FC_INVALID_ARGUMENT, "there are no free degrees of freedom");
    return FC_INVALID_ARGUMENT;
  }
  std::lock_guard<std::mutex> g(f->solve_mu);
  LanczosResult res;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  std::string stopped;
  int st = guarded(err, "not enough device memory for the Lanczos basis", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipEventRecord(e0, f->stream));
    try {
      res = lanczos_run(
          f->sym->s.n, [&](double* W, int nb) { solve_device(f, W, f->sym->s.n, nb, f->sym->s.n, f->stream); },
          f->opts.cancel, m, shift, *opts, vectors_device, f->stream);
    } catch (const LanczosError& le) {
      stopped = le.msg;
    }
    HIP_CHECK(hipEventRecord(e1, f->stream));
    HIP_CHECK(hipEventSynchronize(e1));
  });
  if (st == FC_OK && !stopped.empty()) {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const int kind = stopped == "cancelled" ? FC_CANCELLED : FC_NOT_CONVERGED;
    set_error(err, kind, stopped);
    return kind;
  }
  float ms = 0.f;
  if (st == FC_OK && e0 && e1) (void)hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st != FC_OK) return st;
  for (size_t i = 0; i < res.values.size(); ++i) values[i] = res.values[i];
  if (report) {
    std::memset(report, 0, sizeof *report);
    report->found = res.values.size();
    report->basis = (uint64_t)res.basis;
    report->negative_count = res.negative.size();
    for (size_t i = 0; i < res.negative.size() && i < 16; ++i) report->negative[i] = res.negative[i];
    report->seconds = ms * 1e-3;
  }
  set_error(err, FC_OK, "");
  return FC_OK;
}

int fc_lanczos_device_ldlt(const fc_ldlt* f, const fc_matrix* m, double shift, const fc_lanczos_opts* opts,
                           double* values, double* vectors_device, fc_lanczos_report* report, fc_error* err) {
  if (!f || f->complex || !m || !opts || (opts->count && !values) || m->n != f->n || opts->block < 1 ||
      opts->block > 8 || !(opts->tolerance > 0.0) || !std::isfinite(shift)) {
    set_error(err, FC_INVALID_ARGUMENT, "bad arguments to fc_lanczos_device_ldlt");
    return FC_INVALID_ARGUMENT;
  }
  if (f->n == 0) {
    set_error(err, FC_INVALID_ARGUMENT, "there are no free degrees of freedom");
    return FC_INVALID_ARGUMENT;
  }
  LanczosResult res;
  std::string stopped;
  struct Work {  // the block solves' work vectors, freed however the run ends
    Buffer b;
    ~Work() { b.release(); }
  } work;
  int st = guarded(err, "not enough device memory for the Lanczos basis", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    try {
      res = lanczos_run(
          f->n, [&](double* W, int nb) { ldlt_solve_block_device(f, W, nb, f->stream, work.b); }, nullptr, m, shift,
          *opts, vectors_device, f->stream);
    } catch (const LanczosError& le) {
      stopped = le.msg;
    }
    HIP_CHECK(hipStreamSynchronize(f->stream));
  });
  if (st == FC_OK && !stopped.empty()) {
    set_error(err, FC_NOT_CONVERGED, stopped);
    return FC_NOT_CONVERGED;
  }
  if (st != FC_OK) return st;
  for (size_t i = 0; i < res.values.size(); ++i) values[i] = res.values[i];
  if (report) {
    std::memset(report, 0, sizeof *report);
    report->found = res.values.size();
    report->basis = (uint64_t)res.basis;
    report->negative_count = res.negative.size();
    for (size_t i = 0; i < res.negative.size() && i < 16; ++i) report->negative[i] = res.negative[i];
  }
  set_error(err, FC_OK, "");
  return FC_OK;
}

int fc_analyze(uint64_t n, const uint64_t* ptr, const uint64_t* idx, uint64_t max_entries, fc_symbolic** out,
               fc_error* err) {
  if (!out) {
    set_error(err, FC_INVALID_ARGUMENT, "out is NULL");
    return FC_INVALID_ARGUMENT;
  }
  *out = nullptr;
  try {
    auto h = std::make_shared<fc_symbolic>();
    fc::PlanOptions opt;
    const i64 cap = max_entries > (uint64_t)INT64_MAX ? INT64_MAX : (i64)max_entries;
    fc::Failure f = fc::analyze((i64)n, (const i64*)ptr, (const i64*)idx, cap, opt, h->s);
    if (f.status != fc::OK) {
      set_error(err, f.status, f.message, 0, (uint64_t)f.entries, (uint64_t)f.limit);
      return f.status;
    }
    const fc::Symbolic& s = h->s;
    h->diag_src.assign(s.n, -1);
    {
      std::vector<i64> inv(s.n);
      for (i64 k = 0; k < s.n; ++k) inv[s.perm[k]] = k;
      for (i64 k = 0; k < s.n; ++k)
        for (i64 t = s.upper_ptr[k]; t < s.upper_ptr[k + 1]; ++t)
          if (s.upper_idx[t] == k) h->diag_src[inv[k]] = t;
    }
    h->prog = std::make_shared<const Program>(build_program(s, Schedule{}));
    {
      Registry& R = registry();
      std::lock_guard<std::mutex> g(R.mu);
      R.live.push_back(h);
    }
    *out = h.get();
    set_error(err, FC_OK, "");
    return FC_OK;
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough host memory for the analysis");
    return FC_OUT_OF_MEMORY;
  } catch (const fc::Failure& f) {
    set_error(err, f.status, f.message);
    return f.status;
  }
}

int fc_symbolic_info_get(const fc_symbolic* h, fc_symbolic_info* info) {
  if (!h || !info) return FC_INVALID_ARGUMENT;
  const fc::Symbolic& s = h->s;
  std::memset(info, 0, sizeof *info);
  info->n = s.n;
  info->supernodes = s.ns();
  info->entries = s.entries;
  info->rows = s.rows.size();
  info->updates = s.upd.size();
  info->tasks = s.tasks.size() / 2;
  info->top = s.top.size();
  info->fronts = s.nf;
  info->levels = s.levels;
  info->stored = s.stored;
  info->flops = s.flops;
  info->front_flops = s.front_flops;
  info->l_bytes_fp64 = (uint64_t)s.l_size * 8;
  info->u_bytes_fp64 = (uint64_t)s.u_size * 8;
  info->by_blocks = s.by_blocks;
  info->launches = 0;
  const auto prog = h->program();
  for (const fc::Launch& l : prog->launches) info->launches += l.kind < fc::KINDS;  // not the ordering steps
  for (int k = 0; k < fc::KINDS && k < 8; ++k) {
    info->kind_flops[k] = prog->work[k][0];
    info->kind_entries[k] = prog->work[k][1];
  }
  info->solve_bytes = (uint64_t)(2 * s.n + prog->u_solve) * sizeof(double) + 256;
  return FC_OK;
}

int fc_symbolic_export(const fc_symbolic* h, uint64_t* perm, uint64_t* first, uint64_t* rptr, uint64_t* rows,
                       uint64_t* uptr, uint64_t* updates, uint64_t* tasks, uint64_t* top) {
  if (!h) return FC_INVALID_ARGUMENT;
  const fc::Symbolic& s = h->s;
  auto copy = [](uint64_t* dst, const std::vector<i64>& v) {
    if (dst)
      for (size_t i = 0; i < v.size(); ++i) dst[i] = (uint64_t)v[i];
  };
  copy(perm, s.perm);
  copy(first, s.first);
  copy(rptr, s.rptr);
  copy(rows, s.rows);
  copy(uptr, s.uptr);
  if (updates)
    for (size_t q = 0; q < s.upd.size(); ++q) {
      updates[3 * q] = s.upd[q].k;
      updates[3 * q + 1] = s.upd[q].p1;
      updates[3 * q + 2] = s.upd[q].p2;
    }
  copy(tasks, s.tasks);
  copy(top, s.top);
  return FC_OK;
}

void fc_symbolic_free(fc_symbolic* h) {
  // the last reference may be this one: the analysis (and its device plans) is destroyed
  // outside the registry's lock
  std::shared_ptr<fc_symbolic> last;
  {
    Registry& R = registry();
    std::lock_guard<std::mutex> g(R.mu);
    for (auto it = R.live.begin(); it != R.live.end(); ++it)
      if (it->get() == h) {
        last = std::move(*it);
        R.live.erase(it);
        break;
      }
  }
}

static int factor_create(const fc_symbolic* h, const fc_factor_opts* opts, std::unique_ptr<fc_factor>& F,
                         fc_error* err) {
  fc_factor_opts o{};
  if (opts) o = *opts;
  if (o.precision < FC_FP64 || o.precision > FC_FP32) {
    set_error(err, FC_INVALID_ARGUMENT, "unknown precision");
    return FC_INVALID_ARGUMENT;
  }
  if (o.bf16_terms == 0) o.bf16_terms = 3;
  if (o.bf16_terms != 1 && o.bf16_terms != 3) {
    set_error(err, FC_INVALID_ARGUMENT, "bf16_terms must be 1 or 3 (0: the default, 3)");
    return FC_INVALID_ARGUMENT;
  }
  if (o.fallback != 0 && o.fallback != 1) {
    set_error(err, FC_INVALID_ARGUMENT, "fallback must be 0 (refactor in the next precision) or 1 (report)");
    return FC_INVALID_ARGUMENT;
  }
  F.reset(new fc_factor());
  F->sym_keep = keep(h);
  F->sym = h;
  F->device = o.device;
  F->precision = F->requested = o.precision;
  F->opts = o;
  try {
    HIP_CHECK(hipSetDevice(o.device));
    if (o.stream) {
      F->stream = (hipStream_t)o.stream;
    } else {
      HIP_CHECK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
      F->own_stream = true;
    }
    F->prog = h->program();
    F->plan = const_cast<fc_symbolic*>(h)->plan(o.device, F->prog, F->stream);
    const fc::Symbolic& s = h->s;
    const size_t es = o.precision == FC_FP64 ? sizeof(double) : sizeof(float);
    F->L = (void*)dalloc<char>((size_t)std::max<i64>(s.l_size, 1) * es);
    F->diag = dalloc<double>((size_t)std::max<i64>(s.n, 1));
    F->fail = dalloc<unsigned long long>(1);
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough device memory for the factor", 0, h->s.entries);
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  return FC_OK;
}

int fc_factor_numeric(const fc_symbolic* h, const double* upper_val, uint64_t nnz, const fc_factor_opts* opts,
                      fc_factor** out, fc_error* err) {
  if (!h || !out || (nnz && !upper_val)) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  *out = nullptr;
  std::unique_ptr<fc_factor> F;
  int st = factor_create(h, opts, F, err);
  if (st != FC_OK) return st;
  st = factor_host(F.get(), upper_val, nnz, err);
  if (st != FC_OK) return st;
  *out = F.release();
  return FC_OK;
}

int fc_factor_numeric_device(const fc_symbolic* h, const double* upper_val_device, uint64_t nnz,
                             const fc_factor_opts* opts, fc_factor** out, fc_error* err) {
  if (!h || !out || (nnz && !upper_val_device)) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  *out = nullptr;
  std::unique_ptr<fc_factor> F;
  int st = factor_create(h, opts, F, err);
  if (st != FC_OK) return st;
  st = factor_device(F.get(), upper_val_device, nnz, err);
  if (st != FC_OK) return st;
  *out = F.release();
  return FC_OK;
}

int fc_factor_new(uint64_t n, const uint64_t* ptr, const uint64_t* idx, const double* val, uint64_t max_entries,
                  const fc_factor_opts* opts, fc_factor** out, fc_error* err) {
  fc_symbolic* s = nullptr;
  int st = fc_analyze(n, ptr, idx, max_entries, &s, err);
  if (st != FC_OK) return st;
  st = fc_factor_numeric(s, val, n ? ptr[n] : 0, opts, out, err);
  fc_symbolic_free(s);  // the factor keeps what it needs
  return st;
}

int fc_refactor(fc_factor* f, const double* upper_val, uint64_t nnz, fc_error* err) {
  if (!f || (nnz && !upper_val)) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  return factor_host(f, upper_val, nnz, err);
}

int fc_refactor_device(fc_factor* f, const double* upper_val_device, uint64_t nnz, fc_error* err) {
  if (!f || (nnz && !upper_val_device)) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  return factor_device(f, upper_val_device, nnz, err);
}

uint64_t fc_entries(const fc_factor* f) { return f ? (uint64_t)f->sym->s.entries : 0; }

double fc_factor_seconds(const fc_factor* f) { return f ? f->seconds : 0.0; }

int fc_factor_precision(const fc_factor* f) { return f ? f->precision : -1; }

int fc_factor_profile(const fc_factor* f, int32_t kind, double* launches, double* seconds, double* flops,
                      double* entries) {
  if (!f || kind < 0 || kind >= fc::KINDS) return FC_INVALID_ARGUMENT;
  if (launches) *launches = f->kind_launches[kind];
  if (seconds) *seconds = f->kind_seconds[kind];
  if (flops) *flops = f->prog->work[kind][0];
  if (entries) *entries = f->prog->work[kind][1];
  return FC_OK;
}

void fc_factor_free(fc_factor* f) { delete f; }


int fc_symbolic_schedule_get(const fc_symbolic* h, fc_schedule* out) {
  if (!h || !out) return FC_INVALID_ARGUMENT;
  *out = from_schedule(h->program()->sched);
  return FC_OK;
}

int fc_symbolic_schedule_set(fc_symbolic* h, const fc_schedule* c, fc_error* err) {
  if (!h || !c) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  if (!schedule_ok(*c)) {
    set_error(err, FC_INVALID_ARGUMENT, "schedule out of range (slab 1..16, split targets 1..65536)");
    return FC_INVALID_ARGUMENT;
  }
  try {
    h->install(std::make_shared<const Program>(build_program(h->s, to_schedule(*c))));
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough host memory for the program");
    return FC_OUT_OF_MEMORY;
  }
  set_error(err, FC_OK, "");
  return FC_OK;
}

int fc_tune(fc_symbolic* h, const double* dval, uint64_t nnz, const fc_tune_opts* opts, fc_tune_report* report,
            fc_error* err) {
  if (!h || (nnz && !dval)) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  fc_tune_opts o{};
  if (opts) o = *opts;
  const int repeats = o.repeats > 0 ? o.repeats : 2;
  fc_factor_opts fo{};
  fo.precision = o.precision;
  fo.device = o.device;
  fo.stream = o.stream;
  std::unique_ptr<fc_factor> F;
  int st = factor_create(h, &fo, F, err);
  if (st != FC_OK) return st;
  const std::shared_ptr<const Program> start = F->prog;
  struct Tried {
    Schedule sc;
    double seconds;
    std::shared_ptr<const Program> prog;
  };
  std::vector<Tried> tried;
  // the fastest of `repeats` factorizations (after an untimed one) with schedule sc
  auto time = [&](const Schedule& sc, double& secs) -> int {
    for (const Tried& t : tried)
      if (same(t.sc, sc)) {
        secs = t.seconds;
        return FC_OK;
      }
    std::shared_ptr<const Program> pr;
    try {
      pr = same(sc, start->sched) ? start : std::make_shared<const Program>(build_program(h->s, sc));
      F->prog = pr;
      F->plan = h->plan(F->device, pr, F->stream, pr == start);
    } catch (const std::bad_alloc&) {
      set_error(err, FC_OUT_OF_MEMORY, "not enough memory for a candidate's program");
      return FC_OUT_OF_MEMORY;
    } catch (const DeviceError& d) {
      set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
      return FC_DEVICE;
    }
    secs = 1e300;
    for (int r = 0; r <= repeats; ++r) {
      const int s2 = factor_device(F.get(), dval, nnz, err);
      if (s2 != FC_OK) return s2;
      if (r > 0) secs = std::min(secs, F->seconds);
    }
    tried.push_back(Tried{sc, secs, pr});
    return FC_OK;
  };
  Schedule best = start->sched;
  double best_s = 0;
  if ((st = time(best, best_s)) != FC_OK) return st;
  const double start_s = best_s;
  auto search = [&](int32_t Schedule::*field, std::initializer_list<int32_t> values) -> int {
    const Schedule base = best;
    for (const int32_t v : values) {
      if (tried.size() >= FC_TUNE_MAX) break;
      Schedule c = base;
      c.*field = v;
      double t = 0;
      const int s2 = time(c, t);
      if (s2 != FC_OK) return s2;
      if (t < best_s) {
        best_s = t;
        best = c;
      }
      for (Tried& q : tried)  // only the fastest candidate's program is kept
        if (!same(q.sc, best)) q.prog.reset();
    }
    return FC_OK;
  };
  if ((st = search(&Schedule::slab, {1, 2, 4, 8})) != FC_OK) return st;
  if ((st = search(o.precision == FC_FP64 ? &Schedule::split_side_fp64 : &Schedule::split_side_fp32,
                   {128, 256, 384, 512})) != FC_OK)
    return st;
  if ((st = search(&Schedule::split_target, {512, 1024, 2048})) != FC_OK) return st;
  for (const Tried& t : tried)
    if (same(t.sc, best)) {
      if (t.prog != start) h->install(t.prog);
      break;
    }
  F.reset();  // the factor and the candidates' device plans go before the report is written
  if (report) {
    std::memset(report, 0, sizeof *report);
    report->chosen = from_schedule(best);
    report->seconds_start = start_s;
    report->seconds_chosen = best_s;
    report->candidates = (int32_t)tried.size();
    for (size_t q = 0; q < tried.size() && q < FC_TUNE_MAX; ++q) {
      report->tried[q] = from_schedule(tried[q].sc);
      report->tried_seconds[q] = tried[q].seconds;
    }
  }
  set_error(err, FC_OK, "");
  return FC_OK;
}

int fc_solve_device(const fc_factor* f, double* b, uint64_t n, uint64_t nrhs, uint64_t ldb, fc_error* err) {
  if (!f || !f->valid) {
    set_error(err, FC_INVALID_ARGUMENT, "no valid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n || ldb < n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  std::lock_guard<std::mutex> g(f->solve_mu);
  return guarded(err, "not enough device memory for the solve", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    solve_device(f, b, (i64)n, (int)nrhs, (i64)ldb, f->stream);
    HIP_CHECK(hipStreamSynchronize(f->stream));
  });
}

// ---------------------------------------------------------------------------------------------
// The solve that allocates nothing (fc_solve_device_noalloc): its work vectors (the permuted
// right-hand sides, the forward result and the fronts' contributions) live in `scratch`, one
// copy per device, held by every device the library runs this entry point on until the process
// ends. Host code uses the device address hipGetSymbolAddress gives for it (kept per device),
// never the array's name. The solve kernels write every work entry before reading it, so nothing
// has to be cleared between calls.

}  // extern "C"

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_DEVICES = 64;
struct NoallocDevice {
  std::mutex mu;        // held through a call: calls on one device never overlap
  void* ptr = nullptr;  // this device's copy of scratch
};
// never destroyed: no HIP call or lock may run from a static destructor
NoallocDevice* const g_noalloc = new NoallocDevice[NOALLOC_DEVICES];

// The scratch bytes a call may use: NOALLOC_SCRATCH_BYTES, or less when FC_NOALLOC_LIMIT sets a
// lower figure (a test of the size check needs no large matrix).
size_t noalloc_limit() {
  size_t cap = NOALLOC_SCRATCH_BYTES;
  if (const char* e = std::getenv("FC_NOALLOC_LIMIT")) {
    const size_t v = (size_t)std::strtoull(e, nullptr, 10);
    if (v > 0 && v < cap) cap = v;
  }
  return cap;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

// Bytes of scratch a pass over nr right-hand sides takes: x and y (n x nr each), the
// contributions, and the chained solve's error word (the last 256 bytes), each block 256-byte
// aligned.
size_t noalloc_need(i64 n, i64 u_solve, int nr) {
  return 2 * align256((size_t)n * nr * sizeof(double)) + align256((size_t)std::max<i64>(u_solve * nr, 1) * sizeof(double)) +
         256;
}

}  // namespace

extern "C" {

int fc_solve_device_noalloc(const fc_factor* f, double* b, uint64_t n, uint64_t nrhs, uint64_t ldb, fc_error* err) {
  if (!f || !f->valid) {
    set_error(err, FC_INVALID_ARGUMENT, "no valid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n || ldb < n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  if (f->device < 0 || f->device >= NOALLOC_DEVICES) {
    set_error(err, FC_INVALID_ARGUMENT, "device ordinal beyond the allocation-free solve's table");
    return FC_INVALID_ARGUMENT;
  }
  const i64 us = f->prog->u_solve;
  const size_t cap = noalloc_limit(), one = noalloc_need((i64)n, us, 1);
  if (one > cap) {
    set_error(err, FC_TOO_LARGE,
              "the solve needs " + std::to_string(one) + " bytes of device work per right-hand side; the "
              "allocation-free solve holds " + std::to_string(cap),
              0, one, cap);
    return FC_TOO_LARGE;
  }
  if (n == 0 || nrhs == 0) {
    set_error(err, FC_OK, "");
    return FC_OK;
  }
  // as many right-hand sides a pass as fit (at most MAX_RHS, the solve's own pass width)
  int per = (int)std::min<uint64_t>(nrhs, (uint64_t)fc::MAX_RHS);
  while (per > 1 && noalloc_need((i64)n, us, per) > cap) --per;
  NoallocDevice& D = g_noalloc[f->device];
  std::lock_guard<std::mutex> gd(D.mu);
  std::lock_guard<std::mutex> g(f->solve_mu);
  return guarded(err, "the allocation-free solve allocated", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    if (!D.ptr) {  // once per device, with the device current
      void* p = nullptr;
      HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      D.ptr = p;
    }
    hipStream_t st = f->stream;
    // the chained solve's error word at the array's end, cleared first
    int32_t* serr = (int32_t*)((char*)D.ptr + cap - 256);
    fc::k_fill<int32_t><<<1, 64, 0, st>>>(serr, 1, 0);
    for (uint64_t r0 = 0; r0 < nrhs; r0 += (uint64_t)per) {
      const int nr = (int)std::min<uint64_t>((uint64_t)per, nrhs - r0);
      double* br = b + r0 * ldb;
      char* base = (char*)D.ptr;
      SolveWork w;
      w.xp = (double*)base;
      w.yp = (double*)(base + align256((size_t)n * nr * sizeof(double)));
      w.u = (double*)(base + 2 * align256((size_t)n * nr * sizeof(double)));
      w.err = serr;
      fc::k_permute<<<blocks_for((i64)n), 256, 0, st>>>(br, w.xp, f->plan->perm, (i64)n, nr, (i64)ldb, (i64)n, 0);
      if (f->precision == FC_FP64) run_solve<double>(f, w, (i64)n, nr, st);  // eager: no graph
      else run_solve<float>(f, w, (i64)n, nr, st);
      fc::k_permute<<<blocks_for((i64)n), 256, 0, st>>>(w.xp, br, f->plan->perm, (i64)n, nr, (i64)n, (i64)ldb, 1);
    }
    HIP_CHECK(hipGetLastError());
    solve_check(serr, st);
  });
}

int fc_solve_many(const fc_factor* f, double* const* bs, uint64_t nrhs, uint64_t n, fc_error* err) {
  if (!f || !f->valid) {
    set_error(err, FC_INVALID_ARGUMENT, "no valid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n) {
    set_error(err, FC_INTERNAL, "right-hand side length");  // the reference's Internal (cholesky.rs:866-868)
    return FC_INTERNAL;
  }
  if (n == 0 || nrhs == 0) {
    set_error(err, FC_OK, "");
    return FC_OK;
  }
  std::lock_guard<std::mutex> g(f->solve_mu);
  Buffer tmp;
  int st = guarded(err, "not enough device memory for the solve", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    double* d = (double*)tmp.get((size_t)n * nrhs * sizeof(double));
    for (uint64_t r = 0; r < nrhs; ++r)
      HIP_CHECK(hipMemcpyAsync(d + r * n, bs[r], n * sizeof(double), hipMemcpyHostToDevice, f->stream));
    solve_device(f, d, (i64)n, (int)nrhs, (i64)n, f->stream);
    for (uint64_t r = 0; r < nrhs; ++r)
      HIP_CHECK(hipMemcpyAsync(bs[r], d + r * n, n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
  });
  tmp.release();
  return st;
}

int fc_solve(const fc_factor* f, double* b, uint64_t n, fc_error* err) {
  double* bs[1] = {b};
  return fc_solve_many(f, bs, 1, n, err);
}

int fc_factor_export(const fc_factor* f, double* out, uint64_t len, fc_error* err) {
  if (!f || !f->valid || !out) {
    set_error(err, FC_INVALID_ARGUMENT, "no valid factor");
    return FC_INVALID_ARGUMENT;
  }
  const fc::Symbolic& s = f->sym->s;
  if ((i64)len != s.entries) {
    set_error(err, FC_INVALID_ARGUMENT, "export length differs from the factor's entries");
    return FC_INVALID_ARGUMENT;
  }
  return guarded(err, "not enough host memory for the export", [&] {
    HIP_CHECK(hipSetDevice(f->device));
    std::vector<double> Lh((size_t)s.l_size);
    if (f->precision == FC_FP64) {
      HIP_CHECK(hipMemcpy(Lh.data(), f->L, s.l_size * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      std::vector<float> tmp((size_t)s.l_size);
      HIP_CHECK(hipMemcpy(tmp.data(), f->L, s.l_size * sizeof(float), hipMemcpyDeviceToHost));
      for (i64 e = 0; e < s.l_size; ++e) Lh[e] = tmp[e];
    }
    i64 at = 0;
    for (i64 q = 0; q < s.ns(); ++q) {
      const i64 fr = s.sn_front[q];
      const i64* frows = &s.f_rows[s.f_rptr[fr]];
      const i64 fm = s.f_m[fr];
      const i64 m = s.rptr[q + 1] - s.rptr[q];
      for (i64 c = s.first[q]; c < s.first[q + 1]; ++c) {
        const i64 col = (c - s.f_col0[fr]) * s.f_ld[fr] + s.f_lpos[fr];
        for (i64 t = s.rptr[q] + (c - s.first[q]); t < s.rptr[q] + m; ++t) {
          const i64 pos = std::lower_bound(frows, frows + fm, s.rows[t]) - frows;
          out[at++] = Lh[col + pos];
        }
      }
    }
  });
}

int fc_matrix_new(uint64_t n, const uint64_t* ptr, const uint64_t* idx, const double* val, int32_t device,
                  fc_matrix** out, fc_error* err) {
  if (!out || (n && (!ptr || !idx || !val))) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  if (n >= (uint64_t)INT32_MAX) {
    set_error(err, FC_INVALID_ARGUMENT, "the residual matrix keeps 32-bit column indices");
    return FC_INVALID_ARGUMENT;
  }
  std::unique_ptr<fc_matrix> A(new fc_matrix());
  A->device = device;
  A->n = (i64)n;
  A->nnz = n ? (i64)ptr[n] : 0;
  int st = guarded(err, "not enough device memory for the matrix", [&] {
    HIP_CHECK(hipSetDevice(device));
    std::vector<int64_t> p(ptr, ptr + n + 1);
    std::vector<int32_t> ix(A->nnz);
    for (i64 t = 0; t < A->nnz; ++t) ix[t] = (int32_t)idx[t];
    std::vector<double> v(val, val + A->nnz);
    A->row_abs.assign(n, 0.0);
    for (i64 r = 0; r < (i64)n; ++r) {
      double s = 0.0;
      for (i64 t = p[r]; t < p[r + 1]; ++t) s += std::fabs(v[t]);
      A->row_abs[r] = s;
    }
    A->ptr = dupload(p, nullptr);
    A->idx = dupload(ix, nullptr);
    A->val = dupload(v, nullptr);
    set_segments(A.get(), std::vector<i64>{0, (i64)n});
  });
  if (st != FC_OK) return st;
  *out = A.release();
  return FC_OK;
}

void fc_matrix_free(fc_matrix* a) { delete a; }

int fc_matrix_set_segments(fc_matrix* a, uint64_t count, const uint64_t* offsets, fc_error* err) {
  if (!a || !offsets || count == 0) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument or no segment");
    return FC_INVALID_ARGUMENT;
  }
  std::vector<i64> seg(offsets, offsets + count + 1);
  bool ok = seg.front() == 0 && seg.back() == a->n;
  for (size_t q = 0; ok && q + 1 < seg.size(); ++q) ok = seg[q] <= seg[q + 1];
  if (!ok) {
    set_error(err, FC_INVALID_ARGUMENT, "segment offsets must run from 0 to n without decreasing");
    return FC_INVALID_ARGUMENT;
  }
  return guarded(err, "not enough device memory for the segments", [&] {
    HIP_CHECK(hipSetDevice(a->device));
    set_segments(a, seg);
  });
}

int fc_solve_refined(const fc_factor* f, const fc_matrix* a, const double* b, double* x, uint64_t n,
                     const fc_refine_opts* opts, fc_refine_report* report, fc_error* err) {
  if (!f || !f->valid || !a || !b || !x) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument or invalid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n || a->n != (i64)n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  const fc_refine_opts o = refine_defaults(opts);
  fc_refine_report rep;
  std::memset(&rep, 0, sizeof rep);
  std::lock_guard<std::mutex> g(f->solve_mu);
  Buffer vec, work;
  int st = guarded(err, "not enough device memory for refinement", [&] {
    if (n == 0) {
      rep.converged = 1;
      return;
    }
    HIP_CHECK(hipSetDevice(f->device));
    const i64 N = (i64)n;
    double* db = (double*)vec.get((size_t)2 * N * sizeof(double));
    double* dx = db + N;
    HIP_CHECK(hipMemcpyAsync(db, b, N * sizeof(double), hipMemcpyHostToDevice, f->stream));
    refine_many(f, a, db, dx, N, 1, o, rep, nullptr, work);
    HIP_CHECK(hipMemcpyAsync(x, dx, N * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
  });
  vec.release();
  work.release();
  if (report) *report = rep;
  return st;
}

int fc_solve_refined_device(const fc_factor* f, const fc_matrix* a, const double* b, double* x, uint64_t n,
                            const fc_refine_opts* opts, fc_refine_report* report, fc_error* err) {
  if (!f || !f->valid || !a || !b || !x) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument or invalid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n || a->n != (i64)n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  const fc_refine_opts o = refine_defaults(opts);
  fc_refine_report rep;
  std::memset(&rep, 0, sizeof rep);
  std::lock_guard<std::mutex> g(f->solve_mu);
  int st = guarded(err, "not enough device memory for refinement", [&] {
    if (n == 0) {
      rep.converged = 1;
      return;
    }
    HIP_CHECK(hipSetDevice(f->device));
    refine_many(f, a, b, x, (i64)n, 1, o, rep, nullptr, f->refine_work);
  });
  if (report) *report = rep;
  return st;
}

int fc_solve_refined_many_device(const fc_factor* f, const fc_matrix* a, const double* B, double* X, uint64_t n,
                                 uint64_t nrhs, const fc_refine_opts* opts, fc_refine_report* report, double* berr,
                                 fc_error* err) {
  if (!f || !f->valid || !a || (nrhs && (!B || !X))) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument or invalid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n || a->n != (i64)n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  const fc_refine_opts o = refine_defaults(opts);
  fc_refine_report rep;
  std::memset(&rep, 0, sizeof rep);
  rep.converged = 1;
  std::lock_guard<std::mutex> g(f->solve_mu);
  int st = guarded(err, "not enough device memory for refinement", [&] {
    if (n == 0 || nrhs == 0) {
      if (berr)
        for (uint64_t c = 0; c < nrhs; ++c) berr[c] = 0.0;
      return;
    }
    HIP_CHECK(hipSetDevice(f->device));
    refine_many(f, a, B, X, (i64)n, (i64)nrhs, o, rep, berr, f->refine_work);
  });
  if (report) *report = rep;
  return st;
}

}  // extern "C"
