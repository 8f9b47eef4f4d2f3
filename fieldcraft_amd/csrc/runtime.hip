// The C ABI (include/fc_chol.h): host symbolic phase, device plan, numeric factorization by
// levels of the front tree, solves and refinement on one GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "fc_chol.h"
#include "factor_kernels.hpp"
#include "solve_kernels.hpp"
#include "symbolic.hpp"

using fc::i64;

namespace {

constexpr const char* VERSION = "fieldcraft_amd 0.1 (gfx950)";

struct DeviceError {
  hipError_t e;
  const char* what;
};

#define HIP_CHECK(x)                                   \
  do {                                                 \
    hipError_t e_ = (x);                               \
    if (e_ != hipSuccess) throw DeviceError{e_, #x};   \
  } while (0)

template <typename T>
T* dalloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw std::bad_alloc();
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
  if (p) (void)hipFree(p);
}

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

// The launches of the numeric phase and of the solves, built once per analysis.
struct Program {
  std::vector<fc::Item> items;
  std::vector<fc::Launch> launches;
  std::vector<int32_t> solve_fronts;   // fronts by level
  std::vector<i64> solve_lvl;          // level starts in solve_fronts
  std::vector<i64> uoff;               // per front: start of its solve contribution
  i64 u_solve = 0;                     // total contribution length (one right-hand side)
};

Program build_program(const fc::Symbolic& s) {
  using fc::Item;
  using fc::Launch;
  Program p;
  auto add = [&](int kind, i64 level, i64 item0) {
    const i64 count = (i64)p.items.size() - item0;
    if (count > 0) p.launches.push_back(Launch{kind, (int32_t)level, item0, (int32_t)count, 0});
  };
  for (i64 l = 0; l < s.levels; ++l) {
    const i64* fl = &s.lvl_fronts[s.lvl_ptr[l]];
    const i64 nl = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
    auto has_kids_below = [&](i64 f) {
      for (i64 q = s.f_kptr[f]; q < s.f_kptr[f + 1]; ++q)
        if (s.f_m[s.f_kids[q]] > s.f_w[s.f_kids[q]]) return true;
      return false;
    };
    // extend-add into the L parts
    i64 i0 = (i64)p.items.size();
    for (i64 t = 0; t < nl; ++t) {
      const i64 f = fl[t];
      if (!has_kids_below(f)) continue;
      for (i64 a = 0; a < s.f_w[f]; a += fc::EA_COLS)
        p.items.push_back(Item{(int32_t)f, (int32_t)a, (int32_t)std::min(a + fc::EA_COLS, s.f_w[f]), 0, 0, 0, 0, 0});
    }
    add(fc::K_EXTEND_L, l, i0);
    // panels
    i64 wmax = 0;
    for (i64 t = 0; t < nl; ++t) wmax = std::max(wmax, s.f_w[fl[t]]);
    for (i64 c0 = 0; c0 < wmax; c0 += fc::PANEL) {
      if (c0 > 0) {
        i0 = (i64)p.items.size();
        for (i64 t = 0; t < nl; ++t) {
          const i64 f = fl[t];
          if (s.f_w[f] <= c0) continue;
          const i64 c1 = std::min(c0 + fc::PANEL, s.f_w[f]);
          for (i64 r = c0; r < s.f_m[f]; r += fc::TILE_M)
            p.items.push_back(Item{(int32_t)f, (int32_t)r, (int32_t)c0, (int32_t)c1, 0, 0, 0, 0});
        }
        add(fc::K_UPDATE, l, i0);
      }
      i0 = (i64)p.items.size();
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t];
        if (s.f_w[f] > c0) p.items.push_back(Item{(int32_t)f, (int32_t)c0, 0, 0, 0, 0, 0, 0});
      }
      add(fc::K_POTRF, l, i0);
      i0 = (i64)p.items.size();
      for (i64 t = 0; t < nl; ++t) {
        const i64 f = fl[t];
        if (s.f_w[f] <= c0) continue;
        const i64 c1 = std::min(c0 + fc::PANEL, s.f_w[f]);
        for (i64 r = c1; r < s.f_m[f]; r += fc::TRSM_ROWS)
          p.items.push_back(Item{(int32_t)f, (int32_t)c0, (int32_t)r, 0, 0, 0, 0, 0});
      }
      add(fc::K_TRSM, l, i0);
    }
    // update matrices
    i0 = (i64)p.items.size();
    for (i64 t = 0; t < nl; ++t) {
      const i64 f = fl[t];
      const i64 mu = s.f_m[f] - s.f_w[f];
      for (i64 a = 0; a < mu; a += fc::TILE_M)
        for (i64 b = 0; b <= a; b += fc::TILE_N) p.items.push_back(Item{(int32_t)f, (int32_t)a, (int32_t)b, 0, 0, 0, 0, 0});
    }
    add(fc::K_SYRK, l, i0);
    i0 = (i64)p.items.size();
    for (i64 t = 0; t < nl; ++t) {
      const i64 f = fl[t];
      if (!has_kids_below(f)) continue;
      for (i64 a = s.f_w[f]; a < s.f_m[f]; a += fc::EA_COLS)
        p.items.push_back(Item{(int32_t)f, (int32_t)a, (int32_t)std::min(a + fc::EA_COLS, s.f_m[f]), 0, 0, 0, 0, 0});
    }
    add(fc::K_EXTEND_U, l, i0);
  }
  // solves
  p.solve_lvl.assign(1, 0);
  for (i64 l = 0; l < s.levels; ++l) {
    for (i64 t = s.lvl_ptr[l]; t < s.lvl_ptr[l + 1]; ++t) p.solve_fronts.push_back((int32_t)s.lvl_fronts[t]);
    p.solve_lvl.push_back((i64)p.solve_fronts.size());
  }
  p.uoff.resize(s.nf);
  for (i64 f = 0; f < s.nf; ++f) {
    p.uoff[f] = p.u_solve;
    p.u_solve += s.f_m[f] - s.f_w[f];
  }
  return p;
}

// The plan on one device.
struct DevicePlan {
  int device = -1;
  fc::FrontDev* fronts = nullptr;
  int32_t* kids = nullptr;
  int32_t* rel = nullptr;
  int32_t* rows = nullptr;
  int64_t* a_pos = nullptr;
  int64_t* diag_src = nullptr;
  int64_t* perm = nullptr;
  fc::Item* items = nullptr;
  int32_t* solve_fronts = nullptr;
  int64_t* uoff = nullptr;
  ~DevicePlan() {
    if (device < 0) return;
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    for (void* q : {(void*)fronts, (void*)kids, (void*)rel, (void*)rows, (void*)a_pos, (void*)diag_src, (void*)perm,
                    (void*)items, (void*)solve_fronts, (void*)uoff})
      dfree(q);
    (void)hipSetDevice(cur);
  }
};

}  // namespace

struct fc_symbolic {
  fc::Symbolic s;
  Program prog;
  std::vector<int64_t> diag_src;  // per postorder column: the upper entry of its diagonal
  std::mutex mu;
  std::vector<std::shared_ptr<DevicePlan>> plans;

  std::shared_ptr<DevicePlan> plan(int device, hipStream_t st) {
    std::lock_guard<std::mutex> g(mu);
    for (auto& p : plans)
      if (p->device == device) return p;
    auto p = std::make_shared<DevicePlan>();
    p->device = device;
    const auto& S = s;
    std::vector<fc::FrontDev> fd(S.nf);
    for (i64 f = 0; f < S.nf; ++f) {
      fc::FrontDev& d = fd[f];
      d.lpos = S.f_lpos[f];
      d.upos = S.f_upos[f] < 0 ? 0 : S.f_upos[f];
      d.rel_off = S.f_rel_off[f];
      d.rows_off = S.f_rptr[f];
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
    p->fronts = dupload(fd, st);
    p->kids = dupload(kids, st);
    p->rel = dupload(S.rel, st);
    p->rows = dupload(rows, st);
    p->a_pos = dupload(S.a_pos, st);
    p->diag_src = dupload(diag_src, st);
    p->perm = dupload(S.perm, st);
    p->items = dupload(prog.items, st);
    p->solve_fronts = dupload(prog.solve_fronts, st);
    p->uoff = dupload(prog.uoff, st);
    HIP_CHECK(hipStreamSynchronize(st));
    plans.push_back(p);
    return p;
  }
};

struct fc_factor {
  std::shared_ptr<fc_symbolic> sym_keep;  // keeps the analysis alive
  const fc_symbolic* sym = nullptr;
  std::shared_ptr<DevicePlan> plan;
  int device = 0;
  int precision = FC_FP64;
  fc_factor_opts opts{};
  hipStream_t stream = nullptr;
  bool own_stream = false;
  void* L = nullptr;      // factor storage (double or float)
  double* diag = nullptr; // A's diagonal (postorder)
  unsigned long long* fail = nullptr;
  double seconds = 0.0;
  bool valid = false;
  ~fc_factor() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    dfree(L);
    dfree(diag);
    dfree(fail);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    (void)hipSetDevice(cur);
  }
};

struct fc_matrix {
  int device = 0;
  i64 n = 0, nnz = 0;
  int64_t* ptr = nullptr;
  int32_t* idx = nullptr;
  double* val = nullptr;
  double norm_inf = 0.0;
  ~fc_matrix() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    dfree(ptr);
    dfree(idx);
    dfree(val);
    (void)hipSetDevice(cur);
  }
};

namespace {

// Registry so a factor can keep its analysis alive after fc_symbolic_free.
std::mutex g_reg_mu;
std::vector<std::shared_ptr<fc_symbolic>> g_symbolics;

std::shared_ptr<fc_symbolic> keep(const fc_symbolic* s) {
  std::lock_guard<std::mutex> g(g_reg_mu);
  for (auto& p : g_symbolics)
    if (p.get() == s) return p;
  return nullptr;
}

inline int blocks_for(i64 n, int bs = 256) { return (int)std::max<i64>(1, std::min<i64>((n + bs - 1) / bs, 1 << 16)); }

template <typename T>
void run_factor(fc_factor* F, const double* dval, i64 nnz, hipStream_t st) {
  const fc::Symbolic& s = F->sym->s;
  DevicePlan& P = *F->plan;
  T* L = (T*)F->L;
  T* U = dalloc<T>((size_t)std::max<i64>(s.u_size, 1));
  struct Free {
    T* p;
    ~Free() { dfree(p); }
  } ufree{U};
  fc::k_fill<T><<<blocks_for(s.l_size), 256, 0, st>>>(L, s.l_size, T(0));
  fc::k_scatter_a<T><<<blocks_for(nnz), 256, 0, st>>>(dval, P.a_pos, nnz, L);
  fc::k_gather_diag<<<blocks_for(s.n), 256, 0, st>>>(dval, P.diag_src, s.n, F->diag);
  HIP_CHECK(hipMemsetAsync(F->fail, 0xff, sizeof(unsigned long long), st));
  hipEvent_t e0, e1;
  HIP_CHECK(hipEventCreate(&e0));
  HIP_CHECK(hipEventCreate(&e1));
  HIP_CHECK(hipEventRecord(e0, st));
  const auto& prog = F->sym->prog;
  int level = -1;
  for (const fc::Launch& l : prog.launches) {
    if (l.level != level) {
      level = l.level;
      if (F->opts.cancel && *F->opts.cancel) throw fc::Failure{fc::CANCELLED, -1, 0, 0, "cancelled"};
    }
    const fc::Item* it = P.items + l.item0;
    dim3 g(l.count), b(256);
    switch (l.kind) {
      case fc::K_EXTEND_L:
      case fc::K_EXTEND_U:
        fc::k_extend_add<T><<<g, b, 0, st>>>(it, P.fronts, P.kids, P.rel, L, U, U);
        break;
      case fc::K_UPDATE:
        fc::k_update<T><<<g, b, 0, st>>>(it, P.fronts, L);
        break;
      case fc::K_POTRF:
        fc::k_potrf<T><<<g, b, 0, st>>>(it, P.fronts, L, F->diag, F->fail);
        break;
      case fc::K_TRSM:
        fc::k_trsm<T><<<g, b, 0, st>>>(it, P.fronts, L);
        break;
      case fc::K_SYRK:
        fc::k_syrk<T><<<g, b, 0, st>>>(it, P.fronts, L, U);
        break;
      default:
        throw fc::Failure{fc::INTERNAL, -1, 0, 0, "unknown launch"};
    }
  }
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipEventRecord(e1, st));
  HIP_CHECK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  F->seconds = ms * 1e-3;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

// numeric phase, with failures as values
int factor_numeric(fc_factor* F, const double* upper_val, uint64_t nnz, fc_error* err) {
  const fc::Symbolic& s = F->sym->s;
  if ((i64)nnz != (i64)s.upper_idx.size()) {
    set_error(err, FC_INVALID_ARGUMENT, "the matrix's pattern differs from the analysed one (entry count)");
    return FC_INVALID_ARGUMENT;
  }
  try {
    HIP_CHECK(hipSetDevice(F->device));
    hipStream_t st = F->stream;
    double* dval = dalloc<double>(std::max<uint64_t>(nnz, 1));
    struct Free {
      double* p;
      ~Free() { dfree(p); }
    } vfree{dval};
    if (nnz) HIP_CHECK(hipMemcpyAsync(dval, upper_val, nnz * sizeof(double), hipMemcpyHostToDevice, st));
    F->valid = false;
    if (F->precision == FC_FP64) run_factor<double>(F, dval, (i64)nnz, st);
    else run_factor<float>(F, dval, (i64)nnz, st);
    unsigned long long fail = 0;
    HIP_CHECK(hipMemcpyAsync(&fail, F->fail, sizeof fail, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (fail != ~0ull) {
      const uint64_t col = (uint64_t)s.perm[(i64)fail];
      set_error(err, FC_NOT_POSITIVE_DEFINITE, "the matrix is not positive definite", col, s.entries);
      return FC_NOT_POSITIVE_DEFINITE;
    }
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

template <typename T>
void run_solve(const fc_factor* F, double* xp, i64 ldx, int nrhs, hipStream_t st) {
  const fc::Symbolic& s = F->sym->s;
  const Program& prog = F->sym->prog;
  DevicePlan& P = *F->plan;
  const T* L = (const T*)F->L;
  double* u = dalloc<double>((size_t)std::max<i64>(prog.u_solve * nrhs, 1));
  struct Free {
    double* p;
    ~Free() { dfree(p); }
  } ufree{u};
  const i64 nl = (i64)prog.solve_lvl.size() - 1;
  for (i64 l = 0; l < nl; ++l) {
    const i64 a = prog.solve_lvl[l], b = prog.solve_lvl[l + 1];
    if (b > a)
      fc::k_forward<T><<<(unsigned)(b - a), 256, 0, st>>>(P.solve_fronts + a, P.fronts, P.kids, P.rel, L, xp, ldx, nrhs, u,
                                                           P.uoff);
  }
  for (i64 l = nl - 1; l >= 0; --l) {
    const i64 a = prog.solve_lvl[l], b = prog.solve_lvl[l + 1];
    if (b > a)
      fc::k_backward<T><<<(unsigned)(b - a), 256, 0, st>>>(P.solve_fronts + a, P.fronts, P.rows, L, xp, ldx, nrhs);
  }
  HIP_CHECK(hipGetLastError());
  (void)s;
}

// b (device, n x nrhs, ldb) solved in place
void solve_device(const fc_factor* F, double* b, i64 n, int nrhs, i64 ldb, hipStream_t st) {
  if (n == 0 || nrhs == 0) return;
  double* xp = dalloc<double>((size_t)n * nrhs);
  struct Free {
    double* p;
    ~Free() { dfree(p); }
  } xfree{xp};
  fc::k_permute<<<blocks_for(n), 256, 0, st>>>(b, xp, F->plan->perm, n, nrhs, ldb, n, 0);
  if (F->precision == FC_FP64) run_solve<double>(F, xp, n, nrhs, st);
  else run_solve<float>(F, xp, n, nrhs, st);
  fc::k_permute<<<blocks_for(n), 256, 0, st>>>(xp, b, F->plan->perm, n, nrhs, n, ldb, 1);
  HIP_CHECK(hipGetLastError());
}

}  // namespace

// ------------------------------------------------------------------------------------------

extern "C" {

const char* fc_version(void) { return VERSION; }

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
    h->prog = build_program(s);
    {
      std::lock_guard<std::mutex> g(g_reg_mu);
      g_symbolics.push_back(h);
    }
    *out = h.get();
    set_error(err, FC_OK, "");
    return FC_OK;
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough host memory for the analysis");
    return FC_OUT_OF_MEMORY;
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
  std::lock_guard<std::mutex> g(g_reg_mu);
  g_symbolics.erase(std::remove_if(g_symbolics.begin(), g_symbolics.end(),
                                   [h](const std::shared_ptr<fc_symbolic>& p) { return p.get() == h; }),
                    g_symbolics.end());
}

int fc_factor_numeric(const fc_symbolic* h, const double* upper_val, uint64_t nnz, const fc_factor_opts* opts,
                      fc_factor** out, fc_error* err) {
  if (!h || !out || (nnz && !upper_val)) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  *out = nullptr;
  fc_factor_opts o{};
  if (opts) o = *opts;
  if (o.precision < FC_FP64 || o.precision > FC_FP32) {
    set_error(err, FC_INVALID_ARGUMENT, "unknown precision");
    return FC_INVALID_ARGUMENT;
  }
  std::unique_ptr<fc_factor> F(new fc_factor());
  F->sym_keep = keep(h);
  F->sym = h;
  F->device = o.device;
  F->precision = o.precision;
  F->opts = o;
  try {
    HIP_CHECK(hipSetDevice(o.device));
    if (o.stream) {
      F->stream = (hipStream_t)o.stream;
    } else {
      HIP_CHECK(hipStreamCreateWithFlags(&F->stream, hipStreamNonBlocking));
      F->own_stream = true;
    }
    F->plan = const_cast<fc_symbolic*>(h)->plan(o.device, F->stream);
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
  int st = factor_numeric(F.get(), upper_val, nnz, err);
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
  if (!f) {
    set_error(err, FC_INVALID_ARGUMENT, "null factor");
    return FC_INVALID_ARGUMENT;
  }
  return factor_numeric(f, upper_val, nnz, err);
}

uint64_t fc_entries(const fc_factor* f) { return f ? (uint64_t)f->sym->s.entries : 0; }

double fc_factor_seconds(const fc_factor* f) { return f ? f->seconds : 0.0; }

void fc_factor_free(fc_factor* f) { delete f; }

int fc_solve_device(const fc_factor* f, double* b, uint64_t n, uint64_t nrhs, uint64_t ldb, fc_error* err) {
  if (!f || !f->valid) {
    set_error(err, FC_INVALID_ARGUMENT, "no valid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n || ldb < n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  try {
    HIP_CHECK(hipSetDevice(f->device));
    solve_device(f, b, (i64)n, (int)nrhs, (i64)ldb, f->stream);
    HIP_CHECK(hipStreamSynchronize(f->stream));
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough device memory for the solve");
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  set_error(err, FC_OK, "");
  return FC_OK;
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
  try {
    HIP_CHECK(hipSetDevice(f->device));
    double* d = dalloc<double>((size_t)n * nrhs);
    struct Free {
      double* p;
      ~Free() { dfree(p); }
    } dfree_{d};
    for (uint64_t r = 0; r < nrhs; ++r)
      HIP_CHECK(hipMemcpyAsync(d + r * n, bs[r], n * sizeof(double), hipMemcpyHostToDevice, f->stream));
    solve_device(f, d, (i64)n, (int)nrhs, (i64)n, f->stream);
    for (uint64_t r = 0; r < nrhs; ++r)
      HIP_CHECK(hipMemcpyAsync(bs[r], d + r * n, n * sizeof(double), hipMemcpyDeviceToHost, f->stream));
    HIP_CHECK(hipStreamSynchronize(f->stream));
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough device memory for the solve");
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  set_error(err, FC_OK, "");
  return FC_OK;
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
  try {
    HIP_CHECK(hipSetDevice(f->device));
    const bool dbl = f->precision == FC_FP64;
    std::vector<double> Lh((size_t)s.l_size);
    if (dbl) {
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
          const i64 r = s.rows[t];
          const i64 pos = std::lower_bound(frows, frows + fm, r) - frows;
          out[at++] = Lh[col + pos];
        }
      }
    }
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  set_error(err, FC_OK, "");
  return FC_OK;
}

int fc_matrix_new(uint64_t n, const uint64_t* ptr, const uint64_t* idx, const double* val, int32_t device,
                  fc_matrix** out, fc_error* err) {
  if (!out || (n && (!ptr || !idx || !val))) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument");
    return FC_INVALID_ARGUMENT;
  }
  std::unique_ptr<fc_matrix> A(new fc_matrix());
  A->device = device;
  A->n = (i64)n;
  A->nnz = n ? (i64)ptr[n] : 0;
  try {
    HIP_CHECK(hipSetDevice(device));
    std::vector<int64_t> p(ptr, ptr + n + 1);
    std::vector<int32_t> ix(A->nnz);
    for (i64 t = 0; t < A->nnz; ++t) ix[t] = (int32_t)idx[t];
    std::vector<double> v(val, val + A->nnz);
    double norm = 0.0;
    for (i64 r = 0; r < (i64)n; ++r) {
      double s = 0.0;
      for (i64 t = p[r]; t < p[r + 1]; ++t) s += std::fabs(v[t]);
      norm = std::max(norm, s);
    }
    A->norm_inf = norm;
    A->ptr = dupload(p, nullptr);
    A->idx = dupload(ix, nullptr);
    A->val = dupload(v, nullptr);
    HIP_CHECK(hipDeviceSynchronize());
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough device memory for the matrix");
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  *out = A.release();
  set_error(err, FC_OK, "");
  return FC_OK;
}

void fc_matrix_free(fc_matrix* a) { delete a; }

}  // extern "C"

namespace {

struct Reducer {
  double* partial = nullptr;
  std::vector<double> host = std::vector<double>(fc::RED_BLOCKS);
  double run(hipStream_t st, i64 n, const double* a, const double* b, int mode) {
    fc::k_reduce<<<fc::RED_BLOCKS, 256, 0, st>>>(n, a, b, mode, partial);
    HIP_CHECK(hipMemcpyAsync(host.data(), partial, fc::RED_BLOCKS * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double s = 0.0;
    for (double v : host) s = mode == 0 ? s + v : std::max(s, v);
    return s;
  }
};

}  // namespace

extern "C" int fc_solve_refined(const fc_factor* f, const fc_matrix* a, const double* b, double* x, uint64_t n,
                                const fc_refine_opts* opts, fc_refine_report* report, fc_error* err) {
  if (!f || !f->valid || !a || !b || !x) {
    set_error(err, FC_INVALID_ARGUMENT, "null argument or invalid factor");
    return FC_INVALID_ARGUMENT;
  }
  if ((i64)n != f->sym->s.n || a->n != (i64)n) {
    set_error(err, FC_INTERNAL, "right-hand side length");
    return FC_INTERNAL;
  }
  fc_refine_opts o{2, 50, 1e-12};
  if (opts) {
    o = *opts;
    if (o.max_iter <= 0) o.max_iter = 50;
    if (!(o.target_berr > 0)) o.target_berr = 1e-12;
  }
  fc_refine_report rep;
  std::memset(&rep, 0, sizeof rep);
  if (n == 0) {
    rep.converged = 1;
    if (report) *report = rep;
    set_error(err, FC_OK, "");
    return FC_OK;
  }
  try {
    HIP_CHECK(hipSetDevice(f->device));
    hipStream_t st = f->stream;
    const i64 N = (i64)n;
    double* buf = dalloc<double>((size_t)N * 6 + fc::RED_BLOCKS);
    struct Free {
      double* p;
      ~Free() { dfree(p); }
    } bfree{buf};
    double *db = buf, *dx = buf + N, *dr = buf + 2 * N, *dz = buf + 3 * N, *dp = buf + 4 * N, *dq = buf + 5 * N;
    Reducer red;
    red.partial = buf + 6 * N;
    const int g = blocks_for(N);
    hipEvent_t e0, e1;
    HIP_CHECK(hipEventCreate(&e0));
    HIP_CHECK(hipEventCreate(&e1));
    HIP_CHECK(hipMemcpyAsync(db, b, N * sizeof(double), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipEventRecord(e0, st));
    const double bnorm = red.run(st, N, db, db, 1);
    auto spmv = [&](const double* in, double* out) {
      fc::k_spmv<<<blocks_for(N), 256, 0, st>>>(N, a->ptr, a->idx, a->val, in, out);
    };
    auto residual = [&]() {  // dr = b - A x, returns the backward error
      spmv(dx, dq);
      fc::k_residual<<<g, 256, 0, st>>>(N, db, dq, dr);
      const double rn = red.run(st, N, dr, dr, 1), xn = red.run(st, N, dx, dx, 1);
      const double den = a->norm_inf * xn + bnorm;
      return den > 0 ? rn / den : 0.0;
    };
    HIP_CHECK(hipMemcpyAsync(dx, db, N * sizeof(double), hipMemcpyDeviceToDevice, st));
    solve_device(f, dx, N, 1, N, st);
    double berr = residual();
    rep.berr_history[0] = berr;
    int it = 0;
    bool cg = o.method == 1;
    double prev = berr;
    // plain refinement while it contracts
    while (!cg && berr > o.target_berr && it < o.max_iter) {
      HIP_CHECK(hipMemcpyAsync(dz, dr, N * sizeof(double), hipMemcpyDeviceToDevice, st));
      solve_device(f, dz, N, 1, N, st);
      fc::k_axpy<<<g, 256, 0, st>>>(N, 1.0, dz, dx);
      ++it;
      berr = residual();
      if (it < 64) rep.berr_history[it] = berr;
      if (o.method == 2 && berr > 0.5 * prev) cg = true;  // not contracting: conjugate gradients
      prev = berr;
    }
    if (cg && berr > o.target_berr && it < o.max_iter) {
      // conjugate gradients on A d = r from the current x, preconditioned by the factor
      HIP_CHECK(hipMemcpyAsync(dz, dr, N * sizeof(double), hipMemcpyDeviceToDevice, st));
      solve_device(f, dz, N, 1, N, st);
      HIP_CHECK(hipMemcpyAsync(dp, dz, N * sizeof(double), hipMemcpyDeviceToDevice, st));
      double rz = red.run(st, N, dr, dz, 0);
      // keep a copy of the recursive residual in dq's partner: use dz for z, dr for r
      double* rr = dr;
      while (it < o.max_iter) {
        spmv(dp, dq);
        const double pq = red.run(st, N, dp, dq, 0);
        if (!(pq > 0)) break;
        const double alpha = rz / pq;
        fc::k_axpy<<<g, 256, 0, st>>>(N, alpha, dp, dx);
        fc::k_axpy<<<g, 256, 0, st>>>(N, -alpha, dq, rr);
        ++it;
        // the backward error from the true residual, into dq (dq is free until the next A p)
        spmv(dx, dq);
        fc::k_residual<<<g, 256, 0, st>>>(N, db, dq, dq);
        const double rn = red.run(st, N, dq, dq, 1), xn = red.run(st, N, dx, dx, 1);
        berr = rn / (a->norm_inf * xn + bnorm);
        if (it < 64) rep.berr_history[it] = berr;
        if (berr <= o.target_berr) break;
        HIP_CHECK(hipMemcpyAsync(dz, rr, N * sizeof(double), hipMemcpyDeviceToDevice, st));
        solve_device(f, dz, N, 1, N, st);
        const double rz2 = red.run(st, N, rr, dz, 0);
        fc::k_xpby<<<g, 256, 0, st>>>(N, dz, rz2 / rz, dp);
        rz = rz2;
      }
    }
    HIP_CHECK(hipEventRecord(e1, st));
    HIP_CHECK(hipMemcpyAsync(x, dx, N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    rep.iterations = it;
    rep.berr = berr;
    rep.converged = berr <= o.target_berr;
    rep.seconds = ms * 1e-3;
  } catch (const std::bad_alloc&) {
    set_error(err, FC_OUT_OF_MEMORY, "not enough device memory for refinement");
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    set_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  if (report) *report = rep;
  set_error(err, FC_OK, "");
  return FC_OK;
}
