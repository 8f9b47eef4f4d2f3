/*
 * ORACLE — test infrastructure only. Not part of the shipped library.
 *
 * A plain-C restatement of FieldCraft's supernodal Cholesky path (crates/solver/src/cholesky.rs,
 * ordering.rs, sparse.rs), used by the tests, by smoke() and by bench.py's CPU-baseline leg as
 * the checker of the HIP library. Nothing in fieldcraft_amd/ links, loads or calls it.
 *
 * The restatement keeps the reference's operation order entry by entry:
 *   - every update product S(i,c) = sum_k L(p+i,k) L(p+c,k) is a k-ordered sum started from 0.0
 *     and subtracted once (cholesky.rs:531-630, 636-686);
 *   - the diagonal block subtracts one product at a time (cholesky.rs:1146-1158);
 *   - the rows below a panel subtract one product at a time, then divide (cholesky.rs:691-720);
 *   - updates reach a supernode in increasing descendant order (cholesky.rs:1314-1330).
 * Build with -ffp-contract=off and without fast-math so no multiply-add is fused, as the
 * reference states for itself (cholesky.rs:35).
 *
 * Pinning status (each check is a test under tests/):
 *   postorder                      known answers cholesky.rs:2201-2206
 *   factor + solve                 1x1 ([4] b=[2] -> [0.5]), empty, 7x5 grid error < 1e-13,
 *                                  TooLarge at cap 10, NotPositiveDefinite{column 34}
 *                                  (cholesky.rs:1696-1737); dense scipy Cholesky of P A P^T on
 *                                  random SPD matrices (factor values and solutions)
 *   etree / counts / supernodes    brute-force dense boolean elimination on random graphs;
 *                                  block symbolic == columnwise symbolic (cholesky.rs:2021-2062)
 *   rows / updates                 invariants: |rows(s)| = count(first(s)); the updates of each
 *                                  supernode partition its rows below its columns exactly once
 *   ordering                       12x12 grid: permutation, top separator one grid line
 *                                  (ordering.rs:288-312); cylinder wall graph fill < planes fill
 *                                  and the chosen order's fill equals the graph's
 *                                  (ordering.rs:381-387); 16^3 chosen <= planes (390-395);
 *                                  fill() equals brute-force fill on small graphs
 *   tasks / top                    invariants (tasks are whole subtrees, top are the rest, in
 *                                  order); parity unpinned beyond that (no reference vectors)
 *   Csr::mul                       closed form on small matrices
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;
#define NONE ((i64)-1)

/* cholesky.rs:48-97 */
#define PIVOT_TOLERANCE 1e-10
#define TILE 64
#define PANEL 64
#define PARALLEL_WORK ((i64)1 << 21)
#define SUBTREE_SHARE 256
#define BLOCK 3
/* ordering.rs:25, 113 */
#define LEAF 16
#define PERIPHERAL_SEARCHES 8

enum { ORC_OK = 0, ORC_NOT_POSITIVE_DEFINITE = 1, ORC_SINGULAR = 2, ORC_TOO_LARGE = 3,
       ORC_OUT_OF_MEMORY = 4, ORC_INTERNAL = 5, ORC_CANCELLED = 6 };

/* ------------------------------------------------------------------------------------------ */
/* growable i64 vector                                                                          */
typedef struct { i64* v; i64 n, cap; } vec;
static void vpush(vec* a, i64 x) {
  if (a->n == a->cap) { a->cap = a->cap ? 2 * a->cap : 16; a->v = (i64*)realloc(a->v, sizeof(i64) * a->cap); }
  a->v[a->n++] = x;
}
static void vfree(vec* a) { free(a->v); a->v = NULL; a->n = a->cap = 0; }
static int cmp_i64(const void* a, const void* b) {
  i64 x = *(const i64*)a, y = *(const i64*)b;
  return (x > y) - (x < y);
}

/* Upper triangle by columns (cholesky.rs:99-108). */
typedef struct { i64 n; i64* ptr; i64* idx; double* val; } upper_t;

/* ------------------------------------------------------------------------------------------ */
/* Symbolic                                                                                     */

/* etree (cholesky.rs:157-179) */
static void etree(const upper_t* a, i64* parent) {
  i64 n = a->n;
  i64* ancestor = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  for (i64 k = 0; k < n; ++k) { parent[k] = NONE; ancestor[k] = NONE; }
  for (i64 k = 0; k < n; ++k) {
    for (i64 t = a->ptr[k]; t < a->ptr[k + 1]; ++t) {
      i64 i = a->idx[t];
      while (i != NONE && i < k) {
        i64 next = ancestor[i];
        ancestor[i] = k;
        if (next == NONE) parent[i] = k;
        i = next;
      }
    }
  }
  free(ancestor);
}

/* ereach (cholesky.rs:184-212) */
static i64 ereach(const upper_t* a, i64 k, const i64* parent, i64* stack, i64* mark, i64* path) {
  i64 top = a->n;
  mark[k] = k;
  for (i64 t = a->ptr[k]; t < a->ptr[k + 1]; ++t) {
    i64 i = a->idx[t];
    if (i > k) continue;
    i64 len = 0;
    while (i != NONE && mark[i] != k) {
      path[len++] = i;
      mark[i] = k;
      i = parent[i];
    }
    while (len > 0) stack[--top] = path[--len];
  }
  return top;
}

/* symbolic_by_columns (cholesky.rs:289-305) */
static void symbolic_by_columns(const upper_t* a, i64* parent, i64* counts) {
  i64 n = a->n;
  etree(a, parent);
  i64* stack = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  i64* mark = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  i64* path = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  for (i64 k = 0; k < n; ++k) { counts[k] = 1; mark[k] = NONE; }
  for (i64 k = 0; k < n; ++k) {
    i64 top = ereach(a, k, parent, stack, mark, path);
    for (i64 t = top; t < n; ++t) counts[stack[t]] += 1;
  }
  free(stack); free(mark); free(path);
}

/* uniform_blocks (cholesky.rs:230-266). Returns 1 and fills *b when the matrix is made of whole
 * blocks of three unknowns. */
static int uniform_blocks(const upper_t* a, upper_t* b) {
  i64 n = a->n;
  if (n == 0 || n % BLOCK != 0) return 0;
  i64 nb = n / BLOCK;
  i64* ptr = (i64*)malloc(sizeof(i64) * (nb + 1));
  vec idx = {0};
  ptr[0] = 0;
  for (i64 q = 0; q < nb; ++q) {
    i64 s0 = a->ptr[BLOCK * q], e0 = a->ptr[BLOCK * q + 1];
    if (e0 == s0) goto fail;
    i64 own = a->idx[e0 - 1];
    i64 above = e0 - 1 - s0;
    if (own != BLOCK * q || above % BLOCK != 0) goto fail;
    for (i64 t = 0; t < above; t += BLOCK) {
      i64 first = a->idx[s0 + t];
      if (first % BLOCK != 0) goto fail;
      for (i64 i = 0; i < BLOCK; ++i)
        if (a->idx[s0 + t + i] != first + i) goto fail;
      vpush(&idx, first / BLOCK);
    }
    vpush(&idx, q);
    for (i64 c = 1; c < BLOCK; ++c) {
      i64 s = a->ptr[BLOCK * q + c], e = a->ptr[BLOCK * q + c + 1];
      if (e - s < above) goto fail;
      for (i64 t = 0; t < above; ++t)
        if (a->idx[s + t] != a->idx[s0 + t]) goto fail;
      if (e - s - above != c + 1) goto fail;
      for (i64 i = 0; i < c + 1; ++i)
        if (a->idx[s + above + i] != BLOCK * q + i) goto fail;
    }
    ptr[q + 1] = idx.n;
  }
  b->n = nb; b->ptr = ptr; b->idx = idx.v; b->val = NULL;
  if (!b->idx) b->idx = (i64*)malloc(sizeof(i64));
  return 1;
fail:
  free(ptr); vfree(&idx);
  return 0;
}

/* symbolic (cholesky.rs:274-287): the elimination tree and column counts, from the blocks'
 * matrix when the matrix is made of uniform blocks. */
static void symbolic(const upper_t* a, i64* parent, i64* counts, int* by_blocks) {
  upper_t b;
  if (!uniform_blocks(a, &b)) {
    if (by_blocks) *by_blocks = 0;
    symbolic_by_columns(a, parent, counts);
    return;
  }
  if (by_blocks) *by_blocks = 1;
  i64* bp = (i64*)malloc(sizeof(i64) * (b.n ? b.n : 1));
  i64* bc = (i64*)malloc(sizeof(i64) * (b.n ? b.n : 1));
  symbolic_by_columns(&b, bp, bc);
  for (i64 q = 0; q < b.n; ++q)
    for (i64 c = 0; c < BLOCK; ++c) {
      i64 j = BLOCK * q + c;
      parent[j] = (c + 1 < BLOCK) ? j + 1 : (bp[q] == NONE ? NONE : BLOCK * bp[q]);
      counts[j] = BLOCK * bc[q] - c;
    }
  free(bp); free(bc); free(b.ptr); free(b.idx);
}

/* postorder (cholesky.rs:309-342) */
static void postorder(i64 n, const i64* parent, i64* post) {
  i64* head = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  i64* next = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  i64* stack = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  for (i64 j = 0; j < n; ++j) { head[j] = NONE; next[j] = NONE; }
  for (i64 j = n - 1; j >= 0; --j) {
    i64 p = parent[j];
    if (p == NONE) continue;
    next[j] = head[p];
    head[p] = j;
  }
  i64 k = 0;
  for (i64 root = 0; root < n; ++root) {
    if (parent[root] != NONE) continue;
    i64 sp = 0;
    stack[sp++] = root;
    while (sp > 0) {
      i64 top = stack[sp - 1];
      i64 child = head[top];
      if (child == NONE) {
        post[k++] = top;
        --sp;
      } else {
        head[top] = next[child];
        stack[sp++] = child;
      }
    }
  }
  free(head); free(next); free(stack);
}

/* Pattern::new (cholesky.rs:738-803): A's lower triangle in postorder, with the position of
 * every upper entry. */
typedef struct { i64 r, e; } re_t;
static int cmp_re(const void* a, const void* b) {
  i64 x = ((const re_t*)a)->r, y = ((const re_t*)b)->r;
  return (x > y) - (x < y);
}
typedef struct { i64* ptr; i64* idx; i64* slot; i64* crow; i64* ccol; i64 nnz; } pattern_t;

static void pattern_new(const upper_t* a, const i64* inverse, pattern_t* p) {
  i64 n = a->n, nnz = a->ptr[n];
  p->nnz = nnz;
  p->ptr = (i64*)calloc(n + 1, sizeof(i64));
  for (i64 k = 0; k < n; ++k)
    for (i64 t = a->ptr[k]; t < a->ptr[k + 1]; ++t) {
      i64 i = a->idx[t], ni = inverse[i], nk = inverse[k];
      p->ptr[(ni < nk ? ni : nk) + 1] += 1;
    }
  for (i64 j = 0; j < n; ++j) p->ptr[j + 1] += p->ptr[j];
  re_t* entries = (re_t*)malloc(sizeof(re_t) * (nnz ? nnz : 1));
  i64* next = (i64*)malloc(sizeof(i64) * (n + 1));
  memcpy(next, p->ptr, sizeof(i64) * (n + 1));
  p->crow = (i64*)malloc(sizeof(i64) * (nnz ? nnz : 1));
  p->ccol = (i64*)malloc(sizeof(i64) * (nnz ? nnz : 1));
  for (i64 k = 0; k < n; ++k)
    for (i64 t = a->ptr[k]; t < a->ptr[k + 1]; ++t) {
      i64 ni = inverse[a->idx[t]], nk = inverse[k];
      i64 r = ni > nk ? ni : nk, c = ni < nk ? ni : nk;
      p->crow[t] = r; p->ccol[t] = c;
      entries[next[c]].r = r; entries[next[c]].e = t;
      next[c] += 1;
    }
  /* rows within a column are distinct, so the stable sort of the reference is a plain sort */
  for (i64 j = 0; j < n; ++j)
    qsort(entries + p->ptr[j], (size_t)(p->ptr[j + 1] - p->ptr[j]), sizeof(re_t), cmp_re);
  p->slot = (i64*)malloc(sizeof(i64) * (nnz ? nnz : 1));
  p->idx = (i64*)malloc(sizeof(i64) * (nnz ? nnz : 1));
  for (i64 pos = 0; pos < nnz; ++pos) {
    p->slot[entries[pos].e] = pos;
    p->idx[pos] = entries[pos].r;
  }
  free(entries); free(next);
}

/* One update (cholesky.rs:382-389). */
typedef struct { i64 k, p1, p2; } update_t;

typedef struct orc_factor {
  i64 n, ns, entries;
  int by_blocks;
  i64* perm;
  i64* parent0; /* elimination tree in the input order */
  i64* counts0;
  i64* first;   /* ns+1 */
  i64* rptr;    /* ns+1 */
  i64* rows;
  i64* sparent;
  i64* uptr;    /* ns+1: updates of supernode j are upd[uptr[j]..uptr[j+1]] */
  update_t* upd;
  i64 ntasks;
  i64* tasks;   /* 2 * ntasks: [start, end) */
  i64 ntop;
  i64* top;
  pattern_t pat;
  i64* upper_ptr;
  i64* upper_idx;
  /* numeric */
  double* lval; /* pattern-ordered values */
  double* diag;
  i64* voff;    /* ns+1 offsets of each supernode's trapezoid in val */
  double* val;
  int have_values;
  /* LDL^T (ldlt.rs): pivot kinds per postorder column (0 one, 1 first, 2 second of a 2 x 2),
     each supernode's local pivot order (order[first[s] + c] = the symbolic local column at
     pivot position c), the largest magnitude of each row of A, the count of 2 x 2 pivots */
  signed char* piv;
  i64* order;
  double* scale;
  i64 two_by_two;
} orc_factor;

static i64 s_first(const orc_factor* f, i64 s) { return f->first[s]; }
static i64 s_w(const orc_factor* f, i64 s) { return f->first[s + 1] - f->first[s]; }
static i64 s_m(const orc_factor* f, i64 s) { return f->rptr[s + 1] - f->rptr[s]; }
static const i64* s_rows(const orc_factor* f, i64 s) { return f->rows + f->rptr[s]; }

/* trapezoid / offset (cholesky.rs:412-420) */
static i64 trapezoid(i64 m, i64 w) {
  i64 t = 0;
  for (i64 c = 0; c < w; ++c) t += (m > c ? m - c : 0);
  return t;
}
static i64 offset(i64 m, i64 c) { return c * m - c * (c > 0 ? c - 1 : 0) / 2; }

void orc_free(orc_factor* f) {
  if (!f) return;
  free(f->perm); free(f->parent0); free(f->counts0); free(f->first); free(f->rptr); free(f->rows);
  free(f->sparent); free(f->uptr); free(f->upd); free(f->tasks); free(f->top);
  free(f->pat.ptr); free(f->pat.idx); free(f->pat.slot); free(f->pat.crow); free(f->pat.ccol);
  free(f->upper_ptr); free(f->upper_idx);
  free(f->lval); free(f->diag); free(f->voff); free(f->val);
  free(f->piv); free(f->order); free(f->scale);
  free(f);
}

/* Symbolic::new (cholesky.rs:1213-1378). Returns an ORC_* code; TooLarge fills *entries. */
int orc_analyze(i64 n, const i64* aptr, const i64* aidx, i64 max_entries, orc_factor** out, i64* entries_out) {
  upper_t a = {n, (i64*)aptr, (i64*)aidx, NULL};
  orc_factor* f = (orc_factor*)calloc(1, sizeof(orc_factor));
  f->n = n;
  i64 nn = n ? n : 1;
  f->parent0 = (i64*)malloc(sizeof(i64) * nn);
  f->counts0 = (i64*)malloc(sizeof(i64) * nn);
  symbolic(&a, f->parent0, f->counts0, &f->by_blocks);
  i64 total = 0;
  for (i64 j = 0; j < n; ++j) total += f->counts0[j];
  f->entries = total;
  if (entries_out) *entries_out = total;
  if (total > max_entries) { orc_free(f); return ORC_TOO_LARGE; }
  f->perm = (i64*)malloc(sizeof(i64) * nn);
  postorder(n, f->parent0, f->perm);
  i64* inverse = (i64*)malloc(sizeof(i64) * nn);
  for (i64 k = 0; k < n; ++k) inverse[f->perm[k]] = k;
  i64* parent = (i64*)malloc(sizeof(i64) * nn);
  i64* counts = (i64*)malloc(sizeof(i64) * nn);
  for (i64 k = 0; k < n; ++k) {
    i64 p = f->parent0[f->perm[k]];
    parent[k] = p == NONE ? NONE : inverse[p];
    counts[k] = f->counts0[f->perm[k]];
  }
  pattern_new(&a, inverse, &f->pat);
  f->upper_ptr = (i64*)malloc(sizeof(i64) * (n + 1));
  memcpy(f->upper_ptr, aptr, sizeof(i64) * (n + 1));
  f->upper_idx = (i64*)malloc(sizeof(i64) * (aptr[n] ? aptr[n] : 1));
  memcpy(f->upper_idx, aidx, sizeof(i64) * aptr[n]);

  /* fundamental supernodes (cholesky.rs:1236-1253) */
  i64* children = (i64*)calloc(nn, sizeof(i64));
  for (i64 j = 0; j < n; ++j)
    if (parent[j] != NONE) children[parent[j]] += 1;
  vec first = {0};
  vpush(&first, 0);
  for (i64 j = 1; j < n; ++j) {
    int joins = parent[j - 1] == j && children[j] == 1 && counts[j - 1] == counts[j] + 1;
    if (!joins) vpush(&first, j);
  }
  if (n > 0) vpush(&first, n);
  f->first = first.v;
  i64 ns = first.n - 1;
  f->ns = ns;
  i64* snode = (i64*)malloc(sizeof(i64) * nn);
  for (i64 s = 0; s < ns; ++s)
    for (i64 j = f->first[s]; j < f->first[s + 1]; ++j) snode[j] = s;
  /* supernodal tree and row structures (cholesky.rs:1254-1312) */
  f->sparent = (i64*)malloc(sizeof(i64) * (ns ? ns : 1));
  for (i64 s = 0; s < ns; ++s) {
    i64 last = f->first[s + 1] - 1;
    f->sparent[s] = parent[last] == NONE ? NONE : snode[parent[last]];
  }
  i64* kptr = (i64*)calloc(ns + 1, sizeof(i64));
  for (i64 s = 0; s < ns; ++s)
    if (f->sparent[s] != NONE) kptr[f->sparent[s] + 1] += 1;
  for (i64 s = 0; s < ns; ++s) kptr[s + 1] += kptr[s];
  i64* kids = (i64*)malloc(sizeof(i64) * (ns ? ns : 1));
  i64* kfill = (i64*)malloc(sizeof(i64) * (ns + 1));
  memcpy(kfill, kptr, sizeof(i64) * (ns + 1));
  for (i64 s = 0; s < ns; ++s)
    if (f->sparent[s] != NONE) kids[kfill[f->sparent[s]]++] = s;
  vec rptr = {0}, rows = {0}, set = {0};
  vpush(&rptr, 0);
  i64* mark = (i64*)malloc(sizeof(i64) * nn);
  for (i64 j = 0; j < n; ++j) mark[j] = NONE;
  int bad = 0;
  for (i64 s = 0; s < ns && !bad; ++s) {
    i64 fs = f->first[s], ls = f->first[s + 1];
    set.n = 0;
#define ADD(r) do { i64 r_ = (r); if (mark[r_] != s) { mark[r_] = s; vpush(&set, r_); } } while (0)
    for (i64 j = fs; j < ls; ++j) {
      ADD(j);
      for (i64 t = f->pat.ptr[j]; t < f->pat.ptr[j + 1]; ++t) ADD(f->pat.idx[t]);
    }
    for (i64 q = kptr[s]; q < kptr[s + 1]; ++q) {
      i64 c = kids[q];
      i64 cs = rptr.v[c], ce = rptr.v[c + 1];
      i64 cw = f->first[c + 1] - f->first[c];
      for (i64 t = cs + cw; t < ce; ++t) ADD(rows.v[t]);
    }
#undef ADD
    qsort(set.v, (size_t)set.n, sizeof(i64), cmp_i64);
    if (set.n != counts[fs] || set.v[0] != fs) bad = 1;
    for (i64 t = 0; t < set.n; ++t) vpush(&rows, set.v[t]);
    vpush(&rptr, rows.n);
  }
  f->rptr = rptr.v;
  f->rows = rows.v ? rows.v : (i64*)malloc(sizeof(i64));
  vfree(&set); free(mark); free(kptr); free(kids); free(kfill); free(children);
  if (bad) { free(inverse); free(parent); free(counts); free(snode); orc_free(f); return ORC_INTERNAL; }

  /* updates (cholesky.rs:1314-1330) */
  i64* ucount = (i64*)calloc(ns + 1, sizeof(i64));
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      f->uptr = (i64*)calloc(ns + 1, sizeof(i64));
      for (i64 s = 0; s < ns; ++s) f->uptr[s + 1] = f->uptr[s] + ucount[s];
      f->upd = (update_t*)malloc(sizeof(update_t) * (f->uptr[ns] ? f->uptr[ns] : 1));
      memset(ucount, 0, sizeof(i64) * (ns + 1));
    }
    for (i64 k = 0; k < ns; ++k) {
      const i64* kr = s_rows(f, k);
      i64 km = s_m(f, k);
      i64 p = s_w(f, k);
      while (p < km) {
        i64 target = snode[kr[p]];
        i64 end = f->first[target + 1];
        i64 q = p;
        while (q < km && kr[q] < end) ++q;
        if (pass == 1) {
          update_t u = {k, p, q};
          f->upd[f->uptr[target] + ucount[target]] = u;
        }
        ucount[target] += 1;
        p = q;
      }
    }
  }
  free(ucount);

  /* subtree tasks and top (cholesky.rs:1332-1375) */
  i64* subtree = (i64*)malloc(sizeof(i64) * (ns ? ns : 1));
  i64* start = (i64*)malloc(sizeof(i64) * (ns ? ns : 1));
  for (i64 s = 0; s < ns; ++s) {
    i64 m = s_m(f, s), w = s_w(f, s), sum = 0;
    for (i64 c = 0; c < w; ++c) sum += (m - c) * (m - c);
    subtree[s] = sum;
    start[s] = s;
  }
  for (i64 s = 0; s < ns; ++s) {
    i64 p = f->sparent[s];
    if (p != NONE) {
      subtree[p] += subtree[s];
      if (start[s] < start[p]) start[p] = start[s];
    }
  }
  i64 all = 0;
  for (i64 s = 0; s < ns; ++s)
    if (f->sparent[s] == NONE) all += subtree[s];
  i64 share = all / SUBTREE_SHARE;
  if (share < PARALLEL_WORK) share = PARALLEL_WORK;
  vec tasks = {0}, top = {0};
  for (i64 s = 0; s < ns;) {
    i64 root = NONE, r = s;
    while (r != NONE && start[r] == s && subtree[r] <= share) {
      root = r;
      r = f->sparent[r];
    }
    if (root == NONE) {
      vpush(&top, s);
      s += 1;
    } else {
      vpush(&tasks, s);
      vpush(&tasks, root + 1);
      s = root + 1;
    }
  }
  f->ntasks = tasks.n / 2;
  f->tasks = tasks.v ? tasks.v : (i64*)malloc(sizeof(i64));
  f->ntop = top.n;
  f->top = top.v ? top.v : (i64*)malloc(sizeof(i64));
  free(subtree); free(start); free(inverse); free(parent); free(counts); free(snode);
  *out = f;
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Numeric                                                                                      */

/* Symbolic::numeric (cholesky.rs:1392-1416): values in postorder and the diagonal. */
static int numeric(orc_factor* f, const double* aval) {
  i64 nnz = f->pat.nnz;
  free(f->lval); free(f->diag);
  f->lval = (double*)malloc(sizeof(double) * (nnz ? nnz : 1));
  f->diag = (double*)calloc(f->n ? f->n : 1, sizeof(double));
  for (i64 e = 0; e < nnz; ++e) {
    double v = aval[e];
    f->lval[f->pat.slot[e]] = v;
    if (f->pat.crow[e] == f->pat.ccol[e]) f->diag[f->pat.ccol[e]] = v;
  }
  return ORC_OK;
}

/* L(r, c) of a trapezoid (cholesky.rs:428-437) */
#define AT(data, m, r, c) ((data)[offset((m), (c)) + (r) - (c)])

/* products + subtract for one column c of an update (cholesky.rs:531-686): S(i,c) summed over
 * k in order from 0.0, then own(rel(i), rel(c)) -= S(i,c) for i >= c. */
static void update_column(double* own, i64 m, const double* lk, i64 mk, i64 k0, i64 kw, i64 p,
                          i64 mu, i64 c, const i64* rel) {
  i64 cc = rel[c];
  double* target = own + offset(m, cc);
  for (i64 i = c; i < mu; ++i) {
    double s = 0.0;
    for (i64 k = 0; k < kw; ++k) s += AT(lk, mk, p + i, k0 + k) * AT(lk, mk, p + c, k0 + k);
    target[rel[i] - cc] -= s;
  }
}

static int g_parallel_inner = 1; /* parallelise inside top supernodes (values do not change) */

static void apply_update(double* own, i64 m, const double* lk, i64 mk, i64 k0, i64 kw, i64 p,
                         i64 mu, i64 wu, const i64* rel, int inner) {
  i64 cols = wu < mu ? wu : mu;
  if (inner && mu * wu * kw >= PARALLEL_WORK) {
#pragma omp parallel for schedule(dynamic, 1)
    for (i64 c = 0; c < cols; ++c) update_column(own, m, lk, mk, k0, kw, p, mu, c, rel);
  } else {
    for (i64 c = 0; c < cols; ++c) update_column(own, m, lk, mk, k0, kw, p, mu, c, rel);
  }
}

/* Context::assemble + Context::supernode (cholesky.rs:1047-1189) for supernode j. */
static int supernode(orc_factor* f, i64 j, int inner, i64* failed_column) {
  const i64* rows = s_rows(f, j);
  i64 m = s_m(f, j), w = s_w(f, j), fc = s_first(f, j);
  double* own = f->val + f->voff[j];
  memset(own, 0, sizeof(double) * trapezoid(m, w));
  /* the columns of A (cholesky.rs:1054-1068) */
  for (i64 c = 0; c < w; ++c) {
    for (i64 t = f->pat.ptr[fc + c]; t < f->pat.ptr[fc + c + 1]; ++t) {
      i64 r = f->pat.idx[t];
      i64 lo = 0, hi = m;
      while (lo < hi) { i64 mid = (lo + hi) / 2; if (rows[mid] < r) lo = mid + 1; else hi = mid; }
      if (lo >= m || rows[lo] != r || lo < c) return ORC_INTERNAL;
      own[offset(m, c) + lo - c] = f->lval[t];
    }
  }
  /* the descendants' updates in supernode order (cholesky.rs:1069-1124) */
  i64* rel = (i64*)malloc(sizeof(i64) * (m ? m : 1));
  for (i64 q = f->uptr[j]; q < f->uptr[j + 1]; ++q) {
    update_t u = f->upd[q];
    const i64* kr = s_rows(f, u.k);
    i64 km = s_m(f, u.k), kw = s_w(f, u.k);
    i64 mu = km - u.p1, wu = u.p2 - u.p1;
    i64 at = 0;
    for (i64 i = 0; i < mu; ++i) {
      i64 r = kr[u.p1 + i];
      while (at < m && rows[at] < r) ++at;
      if (at >= m || rows[at] != r) { free(rel); return ORC_INTERNAL; }
      rel[i] = at;
    }
    /* every column of the update lies in J after the one before it (cholesky.rs:661-672) */
    for (i64 c = 0; c < (wu < mu ? wu : mu); ++c)
      if (rel[c] >= m) { free(rel); return ORC_INTERNAL; }
    apply_update(own, m, f->val + f->voff[u.k], km, 0, kw, u.p1, mu, wu, rel, inner);
  }
  /* dense factorization by panels (cholesky.rs:1138-1186) */
  for (i64 i = 0; i < m; ++i) rel[i] = i;
  for (i64 c0 = 0; c0 < w; c0 += PANEL) {
    i64 c1 = c0 + PANEL < w ? c0 + PANEL : w;
    i64 pw = c1 - c0;
    if (c0 > 0) apply_update(own, m, own, m, 0, c0, c0, m - c0, pw, rel + c0, inner);
    /* the diagonal block, column by column (cholesky.rs:1146-1171) */
    for (i64 c = c0; c < c1; ++c) {
      i64 dst = offset(m, c);
      for (i64 k = c0; k < c; ++k) {
        i64 src = offset(m, k) + c - k;
        double lck = own[src];
        for (i64 r = 0; r < c1 - c; ++r) own[dst + r] -= own[src + r] * lck;
      }
      double d = own[dst];
      double diag = f->diag[fc + c];
      if (!(d > PIVOT_TOLERANCE * fabs(diag)) || !isfinite(d)) {
        free(rel);
        *failed_column = fc + c;
        return ORC_NOT_POSITIVE_DEFINITE;
      }
      d = sqrt(d);
      own[dst] = d;
      for (i64 r = 1; r < c1 - c; ++r) own[dst + r] /= d;
    }
    /* the rows below the panel (solve_below, cholesky.rs:691-720), one row at a time: each
     * entry sees the same operations as in the reference's row tiles */
    if (c1 < m) {
      i64 rows_below = m - c1;
#pragma omp parallel for schedule(static) if (inner && rows_below * pw * pw >= PARALLEL_WORK)
      for (i64 i = c1; i < m; ++i) {
        double x[PANEL];
        for (i64 c = 0; c < pw; ++c) x[c] = AT(own, m, i, c0 + c);
        for (i64 c = 0; c < pw; ++c) {
          for (i64 k = 0; k < c; ++k) x[c] -= x[k] * AT(own, m, c0 + c, c0 + k);
          x[c] /= AT(own, m, c0 + c, c0 + c);
        }
        for (i64 c = 0; c < pw; ++c) AT(own, m, i, c0 + c) = x[c];
      }
    }
  }
  free(rel);
  return ORC_OK;
}

/* Symbolic::factor (cholesky.rs:1443-1529): subtree tasks (in parallel), then the supernodes
 * above them in order. A failure is the first failing column in postorder, named by its
 * original column. */
int orc_factor_numeric(orc_factor* f, const double* aval, const volatile int* cancel, i64* column_out) {
  numeric(f, aval);
  i64 ns = f->ns;
  free(f->voff); free(f->val);
  f->voff = (i64*)malloc(sizeof(i64) * (ns + 1));
  f->voff[0] = 0;
  for (i64 s = 0; s < ns; ++s) f->voff[s + 1] = f->voff[s] + trapezoid(s_m(f, s), s_w(f, s));
  f->val = (double*)calloc(f->voff[ns] ? f->voff[ns] : 1, sizeof(double));
  if (!f->val) return ORC_OUT_OF_MEMORY;
  f->have_values = 0;
  i64 nt = f->ntasks;
  i64* fail_col = (i64*)malloc(sizeof(i64) * (nt ? nt : 1));
  int* fail_kind = (int*)calloc(nt ? nt : 1, sizeof(int));
  int broken = ORC_OK;
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 t = 0; t < nt; ++t) {
    for (i64 j = f->tasks[2 * t]; j < f->tasks[2 * t + 1]; ++j) {
      if (cancel && *cancel) {
#pragma omp critical
        { if (broken == ORC_OK) broken = ORC_CANCELLED; }
        break;
      }
      i64 col = NONE;
      int e = supernode(f, j, 0, &col);
      if (e == ORC_NOT_POSITIVE_DEFINITE) { fail_kind[t] = e; fail_col[t] = col; break; }
      if (e != ORC_OK) {
#pragma omp critical
        { if (broken == ORC_OK) broken = e; }
        break;
      }
    }
  }
  if (broken != ORC_OK) { free(fail_col); free(fail_kind); return broken; }
  i64 earliest = NONE;
  for (i64 t = 0; t < nt; ++t)
    if (fail_kind[t] && (earliest == NONE || fail_col[t] < earliest)) earliest = fail_col[t];
  free(fail_col); free(fail_kind);
  for (i64 q = 0; q < f->ntop; ++q) {
    i64 j = f->top[q];
    if (cancel && *cancel) return ORC_CANCELLED;
    if (earliest != NONE && f->first[j] > earliest) break;
    i64 col = NONE;
    int e = supernode(f, j, g_parallel_inner, &col);
    if (e == ORC_NOT_POSITIVE_DEFINITE) { *column_out = f->perm[col]; return e; }
    if (e != ORC_OK) return e;
  }
  if (earliest != NONE) { *column_out = f->perm[earliest]; return ORC_NOT_POSITIVE_DEFINITE; }
  f->have_values = 1;
  return ORC_OK;
}

/* Factor::new (cholesky.rs:1545-1553) */
int orc_factor_new(i64 n, const i64* ptr, const i64* idx, const double* val, i64 max_entries,
                   const volatile int* cancel, orc_factor** out, i64* column_out, i64* entries_out) {
  orc_factor* f = NULL;
  int e = orc_analyze(n, ptr, idx, max_entries, &f, entries_out);
  if (e != ORC_OK) return e;
  e = orc_factor_numeric(f, val, cancel, column_out);
  if (e != ORC_OK) { orc_free(f); return e; }
  *out = f;
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Solve (solve_group, cholesky.rs:904-1037), one right-hand side at a time: each right-hand
 * side sees the same operations whatever the group size. */
static void solve_one(const orc_factor* f, double* b) {
  i64 n = f->n;
  double* x = (double*)malloc(sizeof(double) * (n ? n : 1));
  for (i64 i = 0; i < n; ++i) x[i] = b[f->perm[i]];
  for (i64 k = 0; k < f->ns; ++k) {
    const i64* rows = s_rows(f, k);
    i64 m = s_m(f, k), fc = s_first(f, k), w = s_w(f, k);
    const double* t = f->val + f->voff[k];
    double* y = x + fc;
    for (i64 q = 0; q < w; ++q) {
      double yq = y[q];
      yq /= t[offset(m, q)];
      y[q] = yq;
      for (i64 r = q + 1; r < w; ++r) y[r] -= AT(t, m, r, q) * yq;
      for (i64 r = w; r < m; ++r) x[rows[r]] -= AT(t, m, r, q) * yq;
    }
  }
  for (i64 k = f->ns - 1; k >= 0; --k) {
    const i64* rows = s_rows(f, k);
    i64 m = s_m(f, k), fc = s_first(f, k), w = s_w(f, k);
    const double* t = f->val + f->voff[k];
    double* y = x + fc;
    for (i64 c = w - 1; c >= 0; --c) {
      double sum = y[c];
      for (i64 r = c + 1; r < w; ++r) sum -= AT(t, m, r, c) * y[r];
      for (i64 r = w; r < m; ++r) sum -= AT(t, m, r, c) * x[rows[r]];
      sum /= t[offset(m, c)];
      y[c] = sum;
    }
  }
  for (i64 i = 0; i < n; ++i) b[f->perm[i]] = x[i];
  free(x);
}

/* Factor::solve / solve_many (cholesky.rs:1562-1570); bs is n x nrhs, column-major. */
int orc_solve(const orc_factor* f, double* bs, i64 nrhs) {
  if (!f->have_values) return ORC_INTERNAL;
#pragma omp parallel for schedule(dynamic, 1) if (nrhs > 1)
  for (i64 r = 0; r < nrhs; ++r) solve_one(f, bs + r * f->n);
  return ORC_OK;
}

/* ------------------------------------------------------------------------------------------ */
/* Exports                                                                                      */
i64 orc_n(const orc_factor* f) { return f->n; }
i64 orc_supernodes(const orc_factor* f) { return f->ns; }
i64 orc_entries(const orc_factor* f) { return f->entries; }
i64 orc_row_count(const orc_factor* f) { return f->rptr[f->ns]; }
i64 orc_update_count(const orc_factor* f) { return f->uptr[f->ns]; }
i64 orc_task_count(const orc_factor* f) { return f->ntasks; }
i64 orc_top_count(const orc_factor* f) { return f->ntop; }
int orc_by_blocks(const orc_factor* f) { return f->by_blocks; }
void orc_export_symbolic(const orc_factor* f, i64* perm, i64* parent0, i64* counts0, i64* first,
                         i64* rptr, i64* rows, i64* uptr, i64* upd, i64* tasks, i64* top) {
  i64 n = f->n, ns = f->ns;
  if (perm) memcpy(perm, f->perm, sizeof(i64) * n);
  if (parent0) memcpy(parent0, f->parent0, sizeof(i64) * n);
  if (counts0) memcpy(counts0, f->counts0, sizeof(i64) * n);
  if (first) memcpy(first, f->first, sizeof(i64) * (ns + 1));
  if (rptr) memcpy(rptr, f->rptr, sizeof(i64) * (ns + 1));
  if (rows) memcpy(rows, f->rows, sizeof(i64) * f->rptr[ns]);
  if (uptr) memcpy(uptr, f->uptr, sizeof(i64) * (ns + 1));
  if (upd)
    for (i64 q = 0; q < f->uptr[ns]; ++q) {
      upd[3 * q] = f->upd[q].k; upd[3 * q + 1] = f->upd[q].p1; upd[3 * q + 2] = f->upd[q].p2;
    }
  if (tasks) memcpy(tasks, f->tasks, sizeof(i64) * 2 * f->ntasks);
  if (top) memcpy(top, f->top, sizeof(i64) * f->ntop);
}
/* the trapezoids, concatenated in supernode order (the reference layout, cholesky.rs:393-420) */
int orc_export_values(const orc_factor* f, double* out) {
  if (!f->have_values) return ORC_INTERNAL;
  memcpy(out, f->val, sizeof(double) * f->voff[f->ns]);
  return ORC_OK;
}
void orc_set_parallel_inner(int on) { g_parallel_inner = on; }
int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

/* The elimination tree and column counts alone (cholesky.rs:274-305), column by column or by
 * blocks. */
void orc_symbolic(i64 n, const i64* ptr, const i64* idx, int by_columns, i64* parent, i64* counts) {
  upper_t a = {n, (i64*)ptr, (i64*)idx, NULL};
  if (by_columns) symbolic_by_columns(&a, parent, counts);
  else symbolic(&a, parent, counts, NULL);
}
void orc_postorder(i64 n, const i64* parent, i64* post) { postorder(n, parent, post); }

/* UpperColumns::from_symmetric (cholesky.rs:112-144): CSR of both triangles to upper columns.
 * Call with idx == NULL first to get the entry count. */
i64 orc_from_symmetric(i64 n, const i64* rptr, const i64* ridx, const double* rval, i64* ptr, i64* idx, double* val) {
  i64 total = 0;
  if (ptr) ptr[0] = 0;
  for (i64 k = 0; k < n; ++k) {
    for (i64 t = rptr[k]; t < rptr[k + 1]; ++t)
      if (ridx[t] <= k) {
        if (idx) { idx[total] = ridx[t]; val[total] = rval[t]; }
        total += 1;
      }
    if (ptr) ptr[k + 1] = total;
  }
  return total;
}

/* Csr::mul (sparse.rs:42-50): each row summed in column order; Rust's f64 sum starts from -0.0 */
void orc_csr_mul(i64 n, const i64* ptr, const i64* idx, const double* val, const double* x, double* y) {
#pragma omp parallel for schedule(static)
  for (i64 r = 0; r < n; ++r) {
    double s = -0.0;
    for (i64 t = ptr[r]; t < ptr[r + 1]; ++t) s += val[t] * x[idx[t]];
    y[r] = s;
  }
}

/* ------------------------------------------------------------------------------------------ */
/* Ordering (ordering.rs)                                                                        */

/* f64::total_cmp as an integer key */
static i64 total_key(double x) {
  i64 b;
  memcpy(&b, &x, sizeof b);
  b ^= (i64)(((uint64_t)(b >> 63)) >> 1);
  return b;
}
typedef struct { i64 key, v; } kv_t;
static int cmp_kv(const void* a, const void* b) {
  const kv_t* x = (const kv_t*)a;
  const kv_t* y = (const kv_t*)b;
  if (x->key != y->key) return (x->key > y->key) - (x->key < y->key);
  return (x->v > y->v) - (x->v < y->v);
}

typedef struct { int emit; i64* v; i64 n; } task_t;
typedef struct { task_t* t; i64 n, cap; } tstack;
static void tpush(tstack* s, int emit, i64* v, i64 n) {
  if (s->n == s->cap) { s->cap = s->cap ? 2 * s->cap : 64; s->t = (task_t*)realloc(s->t, sizeof(task_t) * s->cap); }
  s->t[s->n].emit = emit; s->t[s->n].v = v; s->t[s->n].n = n; s->n++;
}

/* nested_dissection (ordering.rs:28-98) */
void orc_nested_dissection(i64 n, const double* xyz, const i64* aptr, const i64* aidx, i64* order) {
  i64 k = 0;
  unsigned char* side = (unsigned char*)calloc(n ? n : 1, 1);
  tstack st = {0};
  i64* all = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  for (i64 v = 0; v < n; ++v) all[v] = v;
  tpush(&st, 0, all, n);
  kv_t* kv = (kv_t*)malloc(sizeof(kv_t) * (n ? n : 1));
  while (st.n > 0) {
    task_t t = st.t[--st.n];
    if (t.emit || t.n <= LEAF) {
      for (i64 i = 0; i < t.n; ++i) order[k++] = t.v[i];
      free(t.v);
      continue;
    }
    i64* part = t.v;
    i64 len = t.n;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (i64 i = 0; i < len; ++i)
      for (int d = 0; d < 3; ++d) {
        double p = xyz[3 * part[i] + d];
        lo[d] = fmin(lo[d], p);
        hi[d] = fmax(hi[d], p);
      }
    int axis = 0;
    for (int d = 0; d < 3; ++d)
      if (hi[d] - lo[d] > hi[axis] - lo[axis]) axis = d;
    for (i64 i = 0; i < len; ++i) { kv[i].key = total_key(xyz[3 * part[i] + axis]); kv[i].v = part[i]; }
    qsort(kv, (size_t)len, sizeof(kv_t), cmp_kv);
    for (i64 i = 0; i < len; ++i) part[i] = kv[i].v;
    i64 half = len / 2;
    for (i64 i = 0; i < len; ++i) side[part[i]] = i < half ? 1 : 2;
    i64* lb = (i64*)malloc(sizeof(i64) * (half ? half : 1));
    i64* rb = (i64*)malloc(sizeof(i64) * (len - half ? len - half : 1));
    i64 nl = 0, nr = 0;
    for (i64 i = 0; i < len; ++i) {
      i64 v = part[i];
      unsigned char other = i < half ? 2 : 1;
      int touches = 0;
      for (i64 q = aptr[v]; q < aptr[v + 1]; ++q)
        if (side[aidx[q]] == other) { touches = 1; break; }
      if (touches) { if (i < half) lb[nl++] = v; else rb[nr++] = v; }
    }
    i64* sep; i64 ns;
    if (nl <= nr) { sep = lb; ns = nl; free(rb); } else { sep = rb; ns = nr; free(lb); }
    for (i64 i = 0; i < len; ++i) side[part[i]] = 0;
    for (i64 i = 0; i < ns; ++i) side[sep[i]] = 3;
    i64* a = (i64*)malloc(sizeof(i64) * (half ? half : 1));
    i64* b = (i64*)malloc(sizeof(i64) * (len - half ? len - half : 1));
    i64 na = 0, nb = 0;
    for (i64 i = 0; i < half; ++i) if (side[part[i]] != 3) a[na++] = part[i];
    for (i64 i = half; i < len; ++i) if (side[part[i]] != 3) b[nb++] = part[i];
    for (i64 i = 0; i < ns; ++i) side[sep[i]] = 0;
    free(part);
    tpush(&st, 1, sep, ns);
    tpush(&st, 0, b, nb);
    tpush(&st, 0, a, na);
  }
  free(st.t); free(side); free(kv);
}

/* level_structure (ordering.rs:117-141); levels stored flat with level starts */
static void level_structure(const i64* aptr, const i64* aidx, i64 root, i64 token, const i64* owner,
                            i64* level, vec* nodes, vec* starts) {
  nodes->n = 0; starts->n = 0;
  vpush(starts, 0);
  vpush(nodes, root);
  vpush(starts, 1);
  level[root] = 0;
  for (;;) {
    i64 depth = starts->n - 1;
    i64 s = starts->v[depth - 1], e = starts->v[depth];
    i64 before = nodes->n;
    for (i64 i = s; i < e; ++i) {
      i64 v = nodes->v[i];
      for (i64 q = aptr[v]; q < aptr[v + 1]; ++q) {
        i64 w = aidx[q];
        if (owner[w] == token && level[w] == NONE) {
          level[w] = depth;
          vpush(nodes, w);
        }
      }
    }
    if (nodes->n == before) return;
    vpush(starts, nodes->n);
  }
}
static void clear_levels(i64* level, const vec* nodes) {
  for (i64 i = 0; i < nodes->n; ++i) level[nodes->v[i]] = NONE;
}

/* graph_dissection (ordering.rs:156-254) */
void orc_graph_dissection(i64 n, const i64* aptr, const i64* aidx, i64* order) {
  i64 k = 0;
  i64* owner = (i64*)calloc(n ? n : 1, sizeof(i64));
  i64* level = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  for (i64 v = 0; v < n; ++v) level[v] = NONE;
  i64 token = 0;
  tstack st = {0};
  i64* all = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  for (i64 v = 0; v < n; ++v) all[v] = v;
  tpush(&st, 0, all, n);
  vec sn = {0}, ss = {0}, dn = {0}, ds = {0};
  while (st.n > 0) {
    task_t t = st.t[--st.n];
    if (t.emit) {
      for (i64 i = 0; i < t.n; ++i) order[k++] = t.v[i];
      free(t.v);
      continue;
    }
    i64* part = t.v;
    i64 len = t.n;
    if (len == 0) { free(part); continue; }
    if (len <= LEAF) {
      for (i64 i = 0; i < len; ++i) order[k++] = part[i];
      free(part);
      continue;
    }
    token += 1;
    for (i64 i = 0; i < len; ++i) owner[part[i]] = token;
    level_structure(aptr, aidx, part[0], token, owner, level, &sn, &ss);
    for (int it = 0; it < PERIPHERAL_SEARCHES; ++it) {
      i64 depth = ss.n - 1;
      i64 s = ss.v[depth - 1], e = ss.v[depth];
      i64 cand = NONE, cdeg = 0;
      for (i64 i = s; i < e; ++i) {
        i64 v = sn.v[i], deg = 0;
        for (i64 q = aptr[v]; q < aptr[v + 1]; ++q) deg += owner[aidx[q]] == token;
        if (cand == NONE || deg < cdeg || (deg == cdeg && v < cand)) { cand = v; cdeg = deg; }
      }
      if (cand == NONE) break;
      clear_levels(level, &sn);
      level_structure(aptr, aidx, cand, token, owner, level, &dn, &ds);
      if (ds.n <= ss.n) {
        clear_levels(level, &dn);
        for (i64 d = 0; d + 1 < ss.n; ++d)
          for (i64 i = ss.v[d]; i < ss.v[d + 1]; ++i) level[sn.v[i]] = d;
        break;
      }
      vec tn = sn, ts = ss;
      sn = dn; ss = ds; dn = tn; ds = ts;
    }
    i64 nlev = ss.n - 1;
    i64 reached = sn.n;
    if (reached < len) {
      i64* rest = (i64*)malloc(sizeof(i64) * len);
      i64 nr = 0;
      for (i64 i = 0; i < len; ++i)
        if (level[part[i]] == NONE) rest[nr++] = part[i];
      clear_levels(level, &sn);
      i64* comp = (i64*)malloc(sizeof(i64) * (reached ? reached : 1));
      memcpy(comp, sn.v, sizeof(i64) * reached);
      qsort(comp, (size_t)reached, sizeof(i64), cmp_i64);
      free(part);
      tpush(&st, 0, rest, nr);
      tpush(&st, 0, comp, reached);
      continue;
    }
    if (nlev < 3) {
      clear_levels(level, &sn);
      for (i64 i = 0; i < len; ++i) order[k++] = part[i];
      free(part);
      continue;
    }
    i64 count = 0, middle = 1;
    for (i64 d = 1; d < nlev; ++d) {
      count += ss.v[d + 1] - ss.v[d];
      if (2 * count >= len) {
        middle = d < nlev - 2 ? d : nlev - 2;
        break;
      }
    }
    i64* before = (i64*)malloc(sizeof(i64) * len);
    i64* sep = (i64*)malloc(sizeof(i64) * len);
    i64* after = (i64*)malloc(sizeof(i64) * len);
    i64 nb = 0, nsep = 0, na = 0;
    for (i64 i = ss.v[0]; i < ss.v[middle]; ++i) before[nb++] = sn.v[i];
    for (i64 i = ss.v[middle]; i < ss.v[middle + 1]; ++i) {
      i64 v = sn.v[i];
      int touches = 0;
      for (i64 q = aptr[v]; q < aptr[v + 1]; ++q) {
        i64 w = aidx[q];
        if (owner[w] == token && level[w] == middle + 1) { touches = 1; break; }
      }
      if (touches) sep[nsep++] = v; else before[nb++] = v;
    }
    for (i64 i = ss.v[middle + 1]; i < sn.n; ++i) after[na++] = sn.v[i];
    clear_levels(level, &sn);
    qsort(before, (size_t)nb, sizeof(i64), cmp_i64);
    qsort(sep, (size_t)nsep, sizeof(i64), cmp_i64);
    qsort(after, (size_t)na, sizeof(i64), cmp_i64);
    free(part);
    tpush(&st, 1, sep, nsep);
    tpush(&st, 0, after, na);
    tpush(&st, 0, before, nb);
  }
  vfree(&sn); vfree(&ss); vfree(&dn); vfree(&ds);
  free(st.t); free(owner); free(level);
}

/* fill (ordering.rs:258-281): entries of the Cholesky factor of the node graph in `order` */
i64 orc_fill(i64 n, const i64* aptr, const i64* aidx, const i64* order) {
  i64* position = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  for (i64 k = 0; k < n; ++k) position[order[k]] = k;
  i64* ptr = (i64*)malloc(sizeof(i64) * (n + 1));
  vec idx = {0};
  ptr[0] = 0;
  i64* rows = (i64*)malloc(sizeof(i64) * (n + 1));
  for (i64 k = 0; k < n; ++k) {
    i64 v = order[k], nr = 0;
    for (i64 q = aptr[v]; q < aptr[v + 1]; ++q) {
      i64 r = position[aidx[q]];
      if (r < k) rows[nr++] = r;
    }
    rows[nr++] = k;
    qsort(rows, (size_t)nr, sizeof(i64), cmp_i64);
    i64 prev = NONE;
    for (i64 i = 0; i < nr; ++i)
      if (rows[i] != prev) { vpush(&idx, rows[i]); prev = rows[i]; }
    ptr[k + 1] = idx.n;
  }
  upper_t a = {n, ptr, idx.v ? idx.v : rows, NULL};
  i64* parent = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  i64* counts = (i64*)malloc(sizeof(i64) * (n ? n : 1));
  symbolic(&a, parent, counts, NULL);
  i64 total = 0;
  for (i64 j = 0; j < n; ++j) total += counts[j];
  free(position); free(ptr); vfree(&idx); free(rows); free(parent); free(counts);
  return total;
}

/* fill_reducing_order (ordering.rs:102-110). Returns 1 when the graph's order was kept. */
int orc_fill_reducing_order(i64 n, const double* xyz, const i64* aptr, const i64* aidx, i64* order) {
  orc_nested_dissection(n, xyz, aptr, aidx, order);
  if (n <= LEAF) return 0;
  i64* by_graph = (i64*)malloc(sizeof(i64) * n);
  orc_graph_dissection(n, aptr, aidx, by_graph);
  i64 fg = orc_fill(n, aptr, aidx, by_graph), fp = orc_fill(n, aptr, aidx, order);
  int graph = fg < fp;
  if (graph) memcpy(order, by_graph, sizeof(i64) * n);
  free(by_graph);
  return graph;
}


/* ------------------------------------------------------------------------------------------ */
/* LDL^T with bounded Bunch-Kaufman (rook) pivoting inside supernodes (crates/solver/src/ldlt.rs),
 * on the Cholesky factor's symbolic analysis, and its pivoted solves (cholesky.rs:904-1037).   */

#define ALPHA 0.6403882032022076 /* (1 + sqrt 17) / 8, ldlt.rs:78 */
#define ROOK_STEPS 64            /* ldlt.rs:82 */
enum { PIV_ONE = 0, PIV_FIRST = 1, PIV_SECOND = 2 };

/* times_d (cholesky.rs:460-484): row (entries for columns k0.. of the trapezoid t of m rows)
 * <- row D, with the pivots from column k0 */
static void times_d(double* row, i64 len, const double* t, i64 m, i64 k0, const signed char* piv) {
  i64 k = 0;
  while (k < len) {
    if (piv[k] == PIV_FIRST) {
      double a = row[k], b = k + 1 < len ? row[k + 1] : 0.0;
      double d11 = AT(t, m, k0 + k, k0 + k), d21 = AT(t, m, k0 + k + 1, k0 + k), d22 = AT(t, m, k0 + k + 1, k0 + k + 1);
      row[k] = a * d11 + b * d21;
      if (k + 1 < len) row[k + 1] = a * d21 + b * d22;
      k += 2;
    } else {
      row[k] *= AT(t, m, k0 + k, k0 + k);
      k += 1;
    }
  }
}

/* products + subtract with Y = L D (cholesky.rs:531-686, pivots given): for the columns
 * [k0, k0 + kw) of trapezoid lk (mk rows, pivots pk from k0), rows p.. :
 * own(rel(i), rel(c)) -= sum_k L(p + i, k) Y(p + c, k), i >= c, each sum in k order from 0.0 */
static void ldlt_update(double* own, i64 m, const double* lk, i64 mk, i64 k0, i64 kw, const signed char* pk, i64 p,
                        i64 mu, i64 wu, const i64* rel, int inner) {
  i64 cols = wu < mu ? wu : mu;
  double* y = (double*)malloc(sizeof(double) * (size_t)(cols * kw > 0 ? cols * kw : 1));
  for (i64 c = 0; c < cols; ++c) {
    for (i64 k = 0; k < kw; ++k) y[c * kw + k] = AT(lk, mk, p + c, k0 + k);
    times_d(y + c * kw, kw, lk, mk, k0, pk);
  }
#pragma omp parallel for schedule(dynamic, 1) if (inner && mu * wu * kw >= PARALLEL_WORK)
  for (i64 c = 0; c < cols; ++c) {
    i64 cc = rel[c];
    double* target = own + offset(m, cc);
    const double* yc = y + c * kw;
    for (i64 i = c; i < mu; ++i) {
      double s = 0.0;
      for (i64 k = 0; k < kw; ++k) s += AT(lk, mk, p + i, k0 + k) * yc[k];
      target[rel[i] - cc] -= s;
    }
  }
  free(y);
}

/* interchange (ldlt.rs:131-145): local columns p < q of an m-row trapezoid */
static void interchange(double* v, i64 m, i64 p, i64 q) {
#define SWP(x, y) do { double t_ = (x); (x) = (y); (y) = t_; } while (0)
  for (i64 c = 0; c < p; ++c) SWP(AT(v, m, p, c), AT(v, m, q, c));
  SWP(AT(v, m, p, p), AT(v, m, q, q));
  for (i64 c = p + 1; c < q; ++c) SWP(AT(v, m, c, p), AT(v, m, q, c));
  for (i64 r = q + 1; r < m; ++r) SWP(AT(v, m, r, p), AT(v, m, r, q));
#undef SWP
}

/* largest (ldlt.rs:149-164): the largest |v[i]| over [lo, hi) except skip (first index of the
 * largest; NaN returned at once) */
static void largest(const double* v, i64 lo, i64 hi, i64 skip, i64* at, double* value) {
  i64 bi = lo;
  double bv = 0.0;
  for (i64 i = lo; i < hi; ++i) {
    if (i == skip) continue;
    double x = fabs(v[i]);
    if (isnan(x)) { *at = i; *value = NAN; return; }
    if (x > bv) { bi = i; bv = x; }
  }
  *at = bi;
  *value = bv;
}

typedef struct {
  double* values;
  signed char* piv;
  i64* order;
  i64 m, w;
  double *a, *b, *wrow;
  i64 two_by_two;
} block_t;

/* Block::current (ldlt.rs:176-193): column c (local, >= k) rows k..m, less the products with the
 * panel's factored columns c0..k */
static void current(block_t* B, i64 c0, i64 k, i64 c, double* out) {
  i64 m = B->m;
  for (i64 r = k; r < m; ++r) out[r] = r >= c ? AT(B->values, m, r, c) : AT(B->values, m, c, r);
  i64 len = k - c0;
  for (i64 q = c0; q < k; ++q) B->wrow[q - c0] = AT(B->values, m, c, q);
  times_d(B->wrow, len, B->values, m, c0, B->piv + c0);
  for (i64 q = c0; q < k; ++q) {
    double wq = B->wrow[q - c0];
    for (i64 r = k; r < m; ++r) out[r] -= AT(B->values, m, r, q) * wq;
  }
}

/* Block::swap (ldlt.rs:196-206) */
static void bswap(block_t* B, i64 p, i64 q) {
  if (p == q) return;
  if (p > q) { i64 t = p; p = q; q = t; }
  interchange(B->values, B->m, p, q);
  i64 t = B->order[p]; B->order[p] = B->order[q]; B->order[q] = t;
  double x = B->a[p]; B->a[p] = B->a[q]; B->a[q] = x;
  x = B->b[p]; B->b[p] = B->b[q]; B->b[q] = x;
}

/* Block::one, Block::two (ldlt.rs:207-253) */
static void store_one(block_t* B, i64 k, const double* col) {
  i64 m = B->m, start = offset(m, k);
  double d = col[k];
  B->values[start] = d;
  for (i64 r = k + 1; r < m; ++r) B->values[start + r - k] = col[r] / d;
  B->piv[k] = PIV_ONE;
}
static void store_two(block_t* B, i64 k, const double* a, const double* b) {
  i64 m = B->m;
  double d11 = a[k], d21 = a[k + 1], d22 = b[k + 1];
  double det = d11 * d22 - d21 * d21;
  i64 s0 = offset(m, k), s1 = offset(m, k + 1);
  B->values[s0] = d11;
  B->values[s0 + 1] = d21;
  B->values[s1] = d22;
  for (i64 r = k + 2; r < m; ++r) {
    double x = a[r], y = b[r];
    B->values[s0 + r - k] = (x * d22 - y * d21) / det;
    B->values[s1 + r - k - 1] = (y * d11 - x * d21) / det;
  }
  B->piv[k] = PIV_FIRST;
  B->piv[k + 1] = PIV_SECOND;
  B->two_by_two += 1;
}

/* Block::panel (ldlt.rs:256-335): factors the panel from c0; returns where it ended, or -1 with
 * *bad the local pivot position of a singular column */
static i64 ldlt_panel(block_t* B, i64 c0, const double* scale_local, i64* bad) {
  i64 w = B->w;
  i64 end = c0 + PANEL < w ? c0 + PANEL : w;
  i64 k = c0;
  while (k < end) {
    current(B, c0, k, k, B->a);
    double akk = fabs(B->a[k]);
    i64 r0;
    double omega_k;
    largest(B->a, k + 1, w, -1, &r0, &omega_k);
    if (!isfinite(akk) || !isfinite(omega_k)) { *bad = k; return -1; }
    int two = 0;
    i64 one_at = k, ti = k, tr = k;
    if (akk >= ALPHA * omega_k) {
      one_at = k;
    } else {
      i64 i = k, r = r0;
      double omega_i = omega_k;
      int step = 0;
      for (;;) {
        current(B, c0, k, r, B->b);
        i64 s_at;
        double omega_r;
        largest(B->b, k, w, r, &s_at, &omega_r);
        double arr = fabs(B->b[r]);
        if (!isfinite(arr) || !isfinite(omega_r)) { *bad = k; return -1; }
        step += 1;
        if (arr >= ALPHA * omega_r) {
          double* t = B->a; B->a = B->b; B->b = t;
          one_at = r;
          break;
        }
        if (omega_r <= omega_i || step >= ROOK_STEPS) { two = 1; ti = i; tr = r; break; }
        double* t = B->a; B->a = B->b; B->b = t;
        i = r; omega_i = omega_r; r = s_at;
      }
    }
    if (!two) {
      bswap(B, k, one_at);
      if (!(fabs(B->a[k]) > PIVOT_TOLERANCE * scale_local[B->order[k]])) { *bad = k; return -1; }
      store_one(B, k, B->a);
      k += 1;
    } else {
      bswap(B, k, ti);
      i64 r = tr == k ? ti : tr;
      bswap(B, k + 1, r);
      double off = fabs(B->a[k + 1]);
      double s0 = scale_local[B->order[k]], s1 = scale_local[B->order[k + 1]];
      double size = s0 > s1 ? s0 : s1;
      if (!(off > PIVOT_TOLERANCE * size)) { *bad = k; return -1; }
      store_two(B, k, B->a, B->b);
      k += 2;
    }
  }
  return k;
}

/* supernode (ldlt.rs:338-358): assemble (A's columns, then the descendants' L D L^T), then the
 * panels right-looking. Returns ORC_OK or ORC_SINGULAR with *failed_column (postorder). */
static int ldlt_supernode(orc_factor* f, i64 j, int inner, i64* failed_column) {
  const i64* rows = s_rows(f, j);
  i64 m = s_m(f, j), w = s_w(f, j), fc = s_first(f, j);
  double* own = f->val + f->voff[j];
  memset(own, 0, sizeof(double) * trapezoid(m, w));
  for (i64 c = 0; c < w; ++c) {
    for (i64 t = f->pat.ptr[fc + c]; t < f->pat.ptr[fc + c + 1]; ++t) {
      i64 r = f->pat.idx[t];
      i64 lo = 0, hi = m;
      while (lo < hi) { i64 mid = (lo + hi) / 2; if (rows[mid] < r) lo = mid + 1; else hi = mid; }
      if (lo >= m || rows[lo] != r || lo < c) return ORC_INTERNAL;
      own[offset(m, c) + lo - c] = f->lval[t];
    }
  }
  i64* rel = (i64*)malloc(sizeof(i64) * (m ? m : 1));
  for (i64 q = f->uptr[j]; q < f->uptr[j + 1]; ++q) {
    update_t u = f->upd[q];
    const i64* kr = s_rows(f, u.k);
    i64 km = s_m(f, u.k), kw = s_w(f, u.k);
    i64 mu = km - u.p1, wu = u.p2 - u.p1;
    i64 at = 0;
    for (i64 i = 0; i < mu; ++i) {
      i64 r = kr[u.p1 + i];
      while (at < m && rows[at] < r) ++at;
      if (at >= m || rows[at] != r) { free(rel); return ORC_INTERNAL; }
      rel[i] = at;
    }
    ldlt_update(own, m, f->val + f->voff[u.k], km, 0, kw, f->piv + s_first(f, u.k), u.p1, mu, wu, rel, inner);
  }
  block_t B;
  B.values = own;
  B.piv = f->piv + fc;
  B.order = f->order + fc;
  B.m = m;
  B.w = w;
  B.a = (double*)calloc((size_t)(m ? m : 1), sizeof(double));
  B.b = (double*)calloc((size_t)(m ? m : 1), sizeof(double));
  B.wrow = (double*)malloc(sizeof(double) * (PANEL + 2));
  B.two_by_two = 0;
  for (i64 c = 0; c < w; ++c) { B.piv[c] = PIV_ONE; B.order[c] = c; }
  double* scale_local = f->scale + fc;
  for (i64 i = 0; i < m; ++i) rel[i] = i;
  int status = ORC_OK;
  i64 c0 = 0;
  while (c0 < w) {
    i64 bad = -1;
    i64 c1 = ldlt_panel(&B, c0, scale_local, &bad);
    if (c1 < 0) { *failed_column = fc + B.order[bad]; status = ORC_SINGULAR; break; }
    if (c1 < w) ldlt_update(own, m, own, m, c0, c1 - c0, B.piv + c0, c1, m - c1, w - c1, rel + c1, inner);
    c0 = c1;
  }
  double* pa = B.a < B.b ? B.a : B.b; /* the two work columns, whichever order the swaps left */
  double* pb = B.a < B.b ? B.b : B.a;
  free(pa); free(pb); free(B.wrow); free(rel);
#pragma omp atomic
  f->two_by_two += B.two_by_two;
  return status;
}

/* LdltFactor::from_symbolic (ldlt.rs:377-391) with Symbolic::numeric's scale
 * (cholesky.rs:1390-1416) and the scheduler of Symbolic::factor (cholesky.rs:1443-1529) */
int orc_ldlt_numeric(orc_factor* f, const double* aval, i64* column_out) {
  numeric(f, aval);
  i64 ns = f->ns, n = f->n;
  free(f->scale);
  f->scale = (double*)calloc(n ? n : 1, sizeof(double));
  for (i64 e = 0; e < f->pat.nnz; ++e) {
    double x = fabs(aval[e]);
    i64 r = f->pat.crow[e], c = f->pat.ccol[e];
    /* f64::max: a NaN entry leaves the scale as it is */
    if (x > f->scale[r]) f->scale[r] = x;
    if (x > f->scale[c]) f->scale[c] = x;
  }
  free(f->voff); free(f->val); free(f->piv); free(f->order);
  f->voff = (i64*)malloc(sizeof(i64) * (ns + 1));
  f->voff[0] = 0;
  for (i64 s2 = 0; s2 < ns; ++s2) f->voff[s2 + 1] = f->voff[s2] + trapezoid(s_m(f, s2), s_w(f, s2));
  f->val = (double*)calloc(f->voff[ns] ? f->voff[ns] : 1, sizeof(double));
  f->piv = (signed char*)calloc(n ? n : 1, 1);
  f->order = (i64*)calloc(n ? n : 1, sizeof(i64));
  f->two_by_two = 0;
  f->have_values = 0;
  i64 nt = f->ntasks;
  i64* fail_col = (i64*)malloc(sizeof(i64) * (nt ? nt : 1));
  int* fail_kind = (int*)calloc(nt ? nt : 1, sizeof(int));
  int broken = ORC_OK;
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 t = 0; t < nt; ++t) {
    for (i64 j = f->tasks[2 * t]; j < f->tasks[2 * t + 1]; ++j) {
      i64 col = NONE;
      int e = ldlt_supernode(f, j, 0, &col);
      if (e == ORC_SINGULAR) { fail_kind[t] = e; fail_col[t] = col; break; }
      if (e != ORC_OK) {
#pragma omp critical
        { if (broken == ORC_OK) broken = e; }
        break;
      }
    }
  }
  if (broken != ORC_OK) { free(fail_col); free(fail_kind); return broken; }
  i64 earliest = NONE;
  for (i64 t = 0; t < nt; ++t)
    if (fail_kind[t] && (earliest == NONE || fail_col[t] < earliest)) earliest = fail_col[t];
  free(fail_col); free(fail_kind);
  for (i64 q = 0; q < f->ntop; ++q) {
    i64 j = f->top[q];
    if (earliest != NONE && f->first[j] > earliest) break;
    i64 col = NONE;
    int e = ldlt_supernode(f, j, g_parallel_inner, &col);
    if (e == ORC_SINGULAR) { *column_out = f->perm[col]; return e; }
    if (e != ORC_OK) return e;
  }
  if (earliest != NONE) { *column_out = f->perm[earliest]; return ORC_SINGULAR; }
  f->have_values = 1;
  return ORC_OK;
}

int orc_ldlt_new(i64 n, const i64* ptr, const i64* idx, const double* val, i64 max_entries, orc_factor** out,
                 i64* column_out, i64* entries_out) {
  orc_factor* f = NULL;
  int e = orc_analyze(n, ptr, idx, max_entries, &f, entries_out);
  if (e != ORC_OK) return e;
  e = orc_ldlt_numeric(f, val, column_out);
  if (e != ORC_OK) { orc_free(f); return e; }
  *out = f;
  return ORC_OK;
}

/* solve_group (cholesky.rs:904-1037) with LDL^T's unit L, block diagonal D and pivot order, one
 * right-hand side */
static void ldlt_solve_one(const orc_factor* f, double* b) {
  i64 n = f->n;
  double* x = (double*)malloc(sizeof(double) * (n ? n : 1));
  double* y = (double*)malloc(sizeof(double) * (n ? n : 1));
  for (i64 i = 0; i < n; ++i) x[i] = b[f->perm[i]];
  for (i64 k = 0; k < f->ns; ++k) {
    const i64* rows = s_rows(f, k);
    i64 m = s_m(f, k), fc = s_first(f, k), w = s_w(f, k);
    const double* t = f->val + f->voff[k];
    const signed char* pv = f->piv + fc;
    const i64* od = f->order + fc;
    double* own = x + fc;
    for (i64 c = 0; c < w; ++c) y[c] = own[od[c]];
    for (i64 q = 0; q < w; ++q) {
      i64 end = q + (pv[q] == PIV_FIRST ? 2 : 1);
      double yq = y[q];
      for (i64 r = end; r < w; ++r) y[r] -= AT(t, m, r, q) * yq;
      for (i64 r = w; r < m; ++r) x[rows[r]] -= AT(t, m, r, q) * yq;
    }
    i64 c = 0;
    while (c < w) {
      if (pv[c] == PIV_FIRST) {
        double d11 = AT(t, m, c, c), d21 = AT(t, m, c + 1, c), d22 = AT(t, m, c + 1, c + 1);
        double det = d11 * d22 - d21 * d21;
        double p = y[c], q2 = y[c + 1];
        y[c] = (d22 * p - d21 * q2) / det;
        y[c + 1] = (d11 * q2 - d21 * p) / det;
        c += 2;
      } else {
        y[c] /= AT(t, m, c, c);
        c += 1;
      }
    }
    for (i64 c2 = 0; c2 < w; ++c2) own[c2] = y[c2]; /* left in pivot order */
  }
  for (i64 k = f->ns - 1; k >= 0; --k) {
    const i64* rows = s_rows(f, k);
    i64 m = s_m(f, k), fc = s_first(f, k), w = s_w(f, k);
    const double* t = f->val + f->voff[k];
    const signed char* pv = f->piv + fc;
    const i64* od = f->order + fc;
    double* own = x + fc;
    for (i64 c = 0; c < w; ++c) y[c] = own[c];
    for (i64 c = w - 1; c >= 0; --c) {
      i64 q, end;
      if (pv[c] == PIV_SECOND && c > 0) { q = c - 1; end = c + 1; }
      else if (pv[c] == PIV_FIRST) { q = c + 1; end = c + 2; }
      else { q = c; end = c + 1; }
      double sum = y[q];
      for (i64 r = end; r < w; ++r) sum -= AT(t, m, r, q) * y[r];
      for (i64 r = w; r < m; ++r) sum -= AT(t, m, r, q) * x[rows[r]];
      y[q] = sum;
    }
    for (i64 pos = 0; pos < w; ++pos) own[od[pos]] = y[pos];
  }
  for (i64 i = 0; i < n; ++i) b[f->perm[i]] = x[i];
  free(x);
  free(y);
}

int orc_ldlt_solve(const orc_factor* f, double* bs, i64 nrhs) {
  if (!f->have_values || !f->piv) return ORC_INTERNAL;
#pragma omp parallel for schedule(dynamic, 1) if (nrhs > 1)
  for (i64 r = 0; r < nrhs; ++r) ldlt_solve_one(f, bs + r * f->n);
  return ORC_OK;
}

/* LdltFactor::inertia (ldlt.rs:412-445) and two_by_two */
void orc_ldlt_inertia(const orc_factor* f, i64* positive, i64* negative, i64* two_by_two) {
  i64 pos = 0, neg = 0;
  for (i64 k = 0; k < f->ns; ++k) {
    i64 m = s_m(f, k), fc = s_first(f, k), w = s_w(f, k);
    const double* t = f->val + f->voff[k];
    const signed char* pv = f->piv + fc;
    i64 c = 0;
    while (c < w) {
      if (pv[c] == PIV_FIRST) {
        double d11 = AT(t, m, c, c), d21 = AT(t, m, c + 1, c), d22 = AT(t, m, c + 1, c + 1);
        double det = d11 * d22 - d21 * d21;
        if (det < 0.0) { pos += 1; neg += 1; }
        else if (d11 + d22 > 0.0) pos += 2;
        else neg += 2;
        c += 2;
      } else {
        if (AT(t, m, c, c) > 0.0) pos += 1; else neg += 1;
        c += 1;
      }
    }
  }
  *positive = pos;
  *negative = neg;
  *two_by_two = f->two_by_two;
}

/* ------------------------------------------------------------------------------------------ */
/* Complex-symmetric LDL^T (ldlt.rs over Complex64, scalar.rs:115-131): the same steps with
 * num-complex 0.4's arithmetic (products (ac - bd, ad + bc), quotients by the squared norm,
 * componentwise sums), the modulus hypot(re, im) for the pivot tests, transposes and never
 * conjugates. The factor's values live in f->val as re, im pairs; no inertia. */

typedef struct { double re, im; } zc;
static inline zc zmk(double re, double im) { zc z = {re, im}; return z; }
static inline zc zadd(zc a, zc b) { return zmk(a.re + b.re, a.im + b.im); }
static inline zc zsub(zc a, zc b) { return zmk(a.re - b.re, a.im - b.im); }
static inline zc zmul(zc a, zc b) { return zmk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline zc zdiv(zc a, zc b) {
  double nrm = b.re * b.re + b.im * b.im;
  return zmk((a.re * b.re + a.im * b.im) / nrm, (a.im * b.re - a.re * b.im) / nrm);
}
static inline double zmod(zc a) { return hypot(a.re, a.im); }

static void ztimes_d(zc* row, i64 len, const zc* t, i64 m, i64 k0, const signed char* piv) {
  i64 k = 0;
  while (k < len) {
    if (piv[k] == PIV_FIRST) {
      zc a = row[k], b = k + 1 < len ? row[k + 1] : zmk(0.0, 0.0);
      zc d11 = AT(t, m, k0 + k, k0 + k), d21 = AT(t, m, k0 + k + 1, k0 + k), d22 = AT(t, m, k0 + k + 1, k0 + k + 1);
      row[k] = zadd(zmul(a, d11), zmul(b, d21));
      if (k + 1 < len) row[k + 1] = zadd(zmul(a, d21), zmul(b, d22));
      k += 2;
    } else {
      row[k] = zmul(row[k], AT(t, m, k0 + k, k0 + k));
      k += 1;
    }
  }
}

static void zldlt_update(zc* own, i64 m, const zc* lk, i64 mk, i64 k0, i64 kw, const signed char* pk, i64 p, i64 mu,
                         i64 wu, const i64* rel, int inner) {
  i64 cols = wu < mu ? wu : mu;
  zc* y = (zc*)malloc(sizeof(zc) * (size_t)(cols * kw > 0 ? cols * kw : 1));
  for (i64 c = 0; c < cols; ++c) {
    for (i64 k = 0; k < kw; ++k) y[c * kw + k] = AT(lk, mk, p + c, k0 + k);
    ztimes_d(y + c * kw, kw, lk, mk, k0, pk);
  }
#pragma omp parallel for schedule(dynamic, 1) if (inner && mu * wu * kw >= PARALLEL_WORK)
  for (i64 c = 0; c < cols; ++c) {
    i64 cc = rel[c];
    zc* target = own + offset(m, cc);
    const zc* yc = y + c * kw;
    for (i64 i = c; i < mu; ++i) {
      zc s = zmk(0.0, 0.0);
      for (i64 k = 0; k < kw; ++k) s = zadd(s, zmul(AT(lk, mk, p + i, k0 + k), yc[k]));
      target[rel[i] - cc] = zsub(target[rel[i] - cc], s);
    }
  }
  free(y);
}

static void zinterchange(zc* v, i64 m, i64 p, i64 q) {
#define ZSWP(x, y) do { zc t_ = (x); (x) = (y); (y) = t_; } while (0)
  for (i64 c = 0; c < p; ++c) ZSWP(AT(v, m, p, c), AT(v, m, q, c));
  ZSWP(AT(v, m, p, p), AT(v, m, q, q));
  for (i64 c = p + 1; c < q; ++c) ZSWP(AT(v, m, c, p), AT(v, m, q, c));
  for (i64 r = q + 1; r < m; ++r) ZSWP(AT(v, m, r, p), AT(v, m, r, q));
#undef ZSWP
}

static void zlargest(const zc* v, i64 lo, i64 hi, i64 skip, i64* at, double* value) {
  i64 bi = lo;
  double bv = 0.0;
  for (i64 i = lo; i < hi; ++i) {
    if (i == skip) continue;
    double x = zmod(v[i]);
    if (isnan(x)) { *at = i; *value = NAN; return; }
    if (x > bv) { bi = i; bv = x; }
  }
  *at = bi;
  *value = bv;
}

typedef struct {
  zc* values;
  signed char* piv;
  i64* order;
  i64 m, w;
  zc *a, *b, *wrow;
  i64 two_by_two;
} zblock_t;

static void zcurrent(zblock_t* B, i64 c0, i64 k, i64 c, zc* out) {
  i64 m = B->m;
  for (i64 r = k; r < m; ++r) out[r] = r >= c ? AT(B->values, m, r, c) : AT(B->values, m, c, r);
  i64 len = k - c0;
  for (i64 q = c0; q < k; ++q) B->wrow[q - c0] = AT(B->values, m, c, q);
  ztimes_d(B->wrow, len, B->values, m, c0, B->piv + c0);
  for (i64 q = c0; q < k; ++q) {
    zc wq = B->wrow[q - c0];
    for (i64 r = k; r < m; ++r) out[r] = zsub(out[r], zmul(AT(B->values, m, r, q), wq));
  }
}

static void zbswap(zblock_t* B, i64 p, i64 q) {
  if (p == q) return;
  if (p > q) { i64 t = p; p = q; q = t; }
  zinterchange(B->values, B->m, p, q);
  i64 t = B->order[p]; B->order[p] = B->order[q]; B->order[q] = t;
  zc x = B->a[p]; B->a[p] = B->a[q]; B->a[q] = x;
  x = B->b[p]; B->b[p] = B->b[q]; B->b[q] = x;
}

static void zstore_one(zblock_t* B, i64 k, const zc* col) {
  i64 m = B->m, start = offset(m, k);
  zc d = col[k];
  B->values[start] = d;
  for (i64 r = k + 1; r < m; ++r) B->values[start + r - k] = zdiv(col[r], d);
  B->piv[k] = PIV_ONE;
}
static void zstore_two(zblock_t* B, i64 k, const zc* a, const zc* b) {
  i64 m = B->m;
  zc d11 = a[k], d21 = a[k + 1], d22 = b[k + 1];
  zc det = zsub(zmul(d11, d22), zmul(d21, d21));
  i64 s0 = offset(m, k), s1 = offset(m, k + 1);
  B->values[s0] = d11;
  B->values[s0 + 1] = d21;
  B->values[s1] = d22;
  for (i64 r = k + 2; r < m; ++r) {
    zc x = a[r], y = b[r];
    B->values[s0 + r - k] = zdiv(zsub(zmul(x, d22), zmul(y, d21)), det);
    B->values[s1 + r - k - 1] = zdiv(zsub(zmul(y, d11), zmul(x, d21)), det);
  }
  B->piv[k] = PIV_FIRST;
  B->piv[k + 1] = PIV_SECOND;
  B->two_by_two += 1;
}

static i64 zldlt_panel(zblock_t* B, i64 c0, const double* scale_local, i64* bad) {
  i64 w = B->w;
  i64 end = c0 + PANEL < w ? c0 + PANEL : w;
  i64 k = c0;
  while (k < end) {
    zcurrent(B, c0, k, k, B->a);
    double akk = zmod(B->a[k]);
    i64 r0;
    double omega_k;
    zlargest(B->a, k + 1, w, -1, &r0, &omega_k);
    if (!isfinite(akk) || !isfinite(omega_k)) { *bad = k; return -1; }
    int two = 0;
    i64 one_at = k, ti = k, tr = k;
    if (!(akk >= ALPHA * omega_k)) {
      i64 i = k, r = r0;
      double omega_i = omega_k;
      int step = 0;
      for (;;) {
        zcurrent(B, c0, k, r, B->b);
        i64 s_at;
        double omega_r;
        zlargest(B->b, k, w, r, &s_at, &omega_r);
        double arr = zmod(B->b[r]);
        if (!isfinite(arr) || !isfinite(omega_r)) { *bad = k; return -1; }
        step += 1;
        if (arr >= ALPHA * omega_r) {
          zc* t = B->a; B->a = B->b; B->b = t;
          one_at = r;
          break;
        }
        if (omega_r <= omega_i || step >= ROOK_STEPS) { two = 1; ti = i; tr = r; break; }
        zc* t = B->a; B->a = B->b; B->b = t;
        i = r; omega_i = omega_r; r = s_at;
      }
    }
    if (!two) {
      zbswap(B, k, one_at);
      if (!(zmod(B->a[k]) > PIVOT_TOLERANCE * scale_local[B->order[k]])) { *bad = k; return -1; }
      zstore_one(B, k, B->a);
      k += 1;
    } else {
      zbswap(B, k, ti);
      i64 r = tr == k ? ti : tr;
      zbswap(B, k + 1, r);
      double off = zmod(B->a[k + 1]);
      double s0 = scale_local[B->order[k]], s1 = scale_local[B->order[k + 1]];
      double size = s0 > s1 ? s0 : s1;
      if (!(off > PIVOT_TOLERANCE * size)) { *bad = k; return -1; }
      zstore_two(B, k, B->a, B->b);
      k += 2;
    }
  }
  return k;
}

static int zldlt_supernode(orc_factor* f, const zc* lval, i64 j, int inner, i64* failed_column) {
  const i64* rows = s_rows(f, j);
  i64 m = s_m(f, j), w = s_w(f, j), fc = s_first(f, j);
  zc* val = (zc*)f->val;
  zc* own = val + f->voff[j];
  memset(own, 0, sizeof(zc) * trapezoid(m, w));
  for (i64 c = 0; c < w; ++c) {
    for (i64 t = f->pat.ptr[fc + c]; t < f->pat.ptr[fc + c + 1]; ++t) {
      i64 r = f->pat.idx[t];
      i64 lo = 0, hi = m;
      while (lo < hi) { i64 mid = (lo + hi) / 2; if (rows[mid] < r) lo = mid + 1; else hi = mid; }
      if (lo >= m || rows[lo] != r || lo < c) return ORC_INTERNAL;
      own[offset(m, c) + lo - c] = lval[t];
    }
  }
  i64* rel = (i64*)malloc(sizeof(i64) * (m ? m : 1));
  for (i64 q = f->uptr[j]; q < f->uptr[j + 1]; ++q) {
    update_t u = f->upd[q];
    const i64* kr = s_rows(f, u.k);
    i64 km = s_m(f, u.k), kw = s_w(f, u.k);
    i64 mu = km - u.p1, wu = u.p2 - u.p1;
    i64 at = 0;
    for (i64 i = 0; i < mu; ++i) {
      i64 r = kr[u.p1 + i];
      while (at < m && rows[at] < r) ++at;
      if (at >= m || rows[at] != r) { free(rel); return ORC_INTERNAL; }
      rel[i] = at;
    }
    zldlt_update(own, m, val + f->voff[u.k], km, 0, kw, f->piv + s_first(f, u.k), u.p1, mu, wu, rel, inner);
  }
  zblock_t B;
  B.values = own;
  B.piv = f->piv + fc;
  B.order = f->order + fc;
  B.m = m;
  B.w = w;
  B.a = (zc*)calloc((size_t)(m ? m : 1), sizeof(zc));
  B.b = (zc*)calloc((size_t)(m ? m : 1), sizeof(zc));
  B.wrow = (zc*)malloc(sizeof(zc) * (PANEL + 2));
  B.two_by_two = 0;
  for (i64 c = 0; c < w; ++c) { B.piv[c] = PIV_ONE; B.order[c] = c; }
  double* scale_local = f->scale + fc;
  for (i64 i = 0; i < m; ++i) rel[i] = i;
  int status = ORC_OK;
  i64 c0 = 0;
  while (c0 < w) {
    i64 bad = -1;
    i64 c1 = zldlt_panel(&B, c0, scale_local, &bad);
    if (c1 < 0) { *failed_column = fc + B.order[bad]; status = ORC_SINGULAR; break; }
    if (c1 < w) zldlt_update(own, m, own, m, c0, c1 - c0, B.piv + c0, c1, m - c1, w - c1, rel + c1, inner);
    c0 = c1;
  }
  zc* pa = B.a < B.b ? B.a : B.b;
  zc* pb = B.a < B.b ? B.b : B.a;
  free(pa); free(pb); free(B.wrow); free(rel);
#pragma omp atomic
  f->two_by_two += B.two_by_two;
  return status;
}

/* LdltFactor::<Complex64>::from_symbolic: aval holds nnz (re, im) pairs of the upper columns */
int orc_zldlt_numeric(orc_factor* f, const double* aval_pairs, i64* column_out) {
  const zc* aval = (const zc*)aval_pairs;
  i64 ns = f->ns, n = f->n, nnz = f->pat.nnz;
  zc* lval = (zc*)malloc(sizeof(zc) * (nnz ? nnz : 1));
  for (i64 e = 0; e < nnz; ++e) lval[f->pat.slot[e]] = aval[e];
  free(f->scale);
  f->scale = (double*)calloc(n ? n : 1, sizeof(double));
  for (i64 e = 0; e < nnz; ++e) {
    double x = zmod(aval[e]);
    i64 r = f->pat.crow[e], c = f->pat.ccol[e];
    if (x > f->scale[r]) f->scale[r] = x;
    if (x > f->scale[c]) f->scale[c] = x;
  }
  free(f->voff); free(f->val); free(f->piv); free(f->order);
  f->voff = (i64*)malloc(sizeof(i64) * (ns + 1));
  f->voff[0] = 0;
  for (i64 s2 = 0; s2 < ns; ++s2) f->voff[s2 + 1] = f->voff[s2] + trapezoid(s_m(f, s2), s_w(f, s2));
  f->val = (double*)calloc(2 * (f->voff[ns] ? f->voff[ns] : 1), sizeof(double));
  f->piv = (signed char*)calloc(n ? n : 1, 1);
  f->order = (i64*)calloc(n ? n : 1, sizeof(i64));
  f->two_by_two = 0;
  f->have_values = 0;
  i64 nt = f->ntasks;
  i64* fail_col = (i64*)malloc(sizeof(i64) * (nt ? nt : 1));
  int* fail_kind = (int*)calloc(nt ? nt : 1, sizeof(int));
  int broken = ORC_OK;
#pragma omp parallel for schedule(dynamic, 1)
  for (i64 t = 0; t < nt; ++t) {
    for (i64 j = f->tasks[2 * t]; j < f->tasks[2 * t + 1]; ++j) {
      i64 col = NONE;
      int e = zldlt_supernode(f, lval, j, 0, &col);
      if (e == ORC_SINGULAR) { fail_kind[t] = e; fail_col[t] = col; break; }
      if (e != ORC_OK) {
#pragma omp critical
        { if (broken == ORC_OK) broken = e; }
        break;
      }
    }
  }
  if (broken != ORC_OK) { free(fail_col); free(fail_kind); free(lval); return broken; }
  i64 earliest = NONE;
  for (i64 t = 0; t < nt; ++t)
    if (fail_kind[t] && (earliest == NONE || fail_col[t] < earliest)) earliest = fail_col[t];
  free(fail_col); free(fail_kind);
  for (i64 q = 0; q < f->ntop; ++q) {
    i64 j = f->top[q];
    if (earliest != NONE && f->first[j] > earliest) break;
    i64 col = NONE;
    int e = zldlt_supernode(f, lval, j, g_parallel_inner, &col);
    if (e == ORC_SINGULAR) { free(lval); *column_out = f->perm[col]; return e; }
    if (e != ORC_OK) { free(lval); return e; }
  }
  free(lval);
  if (earliest != NONE) { *column_out = f->perm[earliest]; return ORC_SINGULAR; }
  f->have_values = 1;
  return ORC_OK;
}

int orc_zldlt_new(i64 n, const i64* ptr, const i64* idx, const double* val_pairs, i64 max_entries, orc_factor** out,
                  i64* column_out, i64* entries_out) {
  orc_factor* f = NULL;
  int e = orc_analyze(n, ptr, idx, max_entries, &f, entries_out);
  if (e != ORC_OK) return e;
  e = orc_zldlt_numeric(f, val_pairs, column_out);
  if (e != ORC_OK) { orc_free(f); return e; }
  *out = f;
  return ORC_OK;
}

static void zldlt_solve_one(const orc_factor* f, zc* b) {
  i64 n = f->n;
  const zc* val = (const zc*)f->val;
  zc* x = (zc*)malloc(sizeof(zc) * (n ? n : 1));
  zc* y = (zc*)malloc(sizeof(zc) * (n ? n : 1));
  for (i64 i = 0; i < n; ++i) x[i] = b[f->perm[i]];
  for (i64 k = 0; k < f->ns; ++k) {
    const i64* rows = s_rows(f, k);
    i64 m = s_m(f, k), fc = s_first(f, k), w = s_w(f, k);
    const zc* t = val + f->voff[k];
    const signed char* pv = f->piv + fc;
    const i64* od = f->order + fc;
    zc* own = x + fc;
    for (i64 c = 0; c < w; ++c) y[c] = own[od[c]];
    for (i64 q = 0; q < w; ++q) {
      i64 end = q + (pv[q] == PIV_FIRST ? 2 : 1);
      zc yq = y[q];
      for (i64 r = end; r < w; ++r) y[r] = zsub(y[r], zmul(AT(t, m, r, q), yq));
      for (i64 r = w; r < m; ++r) x[rows[r]] = zsub(x[rows[r]], zmul(AT(t, m, r, q), yq));
    }
    i64 c = 0;
    while (c < w) {
      if (pv[c] == PIV_FIRST) {
        zc d11 = AT(t, m, c, c), d21 = AT(t, m, c + 1, c), d22 = AT(t, m, c + 1, c + 1);
        zc det = zsub(zmul(d11, d22), zmul(d21, d21));
        zc p = y[c], q2 = y[c + 1];
        y[c] = zdiv(zsub(zmul(d22, p), zmul(d21, q2)), det);
        y[c + 1] = zdiv(zsub(zmul(d11, q2), zmul(d21, p)), det);
        c += 2;
      } else {
        y[c] = zdiv(y[c], AT(t, m, c, c));
        c += 1;
      }
    }
    for (i64 c2 = 0; c2 < w; ++c2) own[c2] = y[c2];
  }
  for (i64 k = f->ns - 1; k >= 0; --k) {
    const i64* rows = s_rows(f, k);
    i64 m = s_m(f, k), fc = s_first(f, k), w = s_w(f, k);
    const zc* t = val + f->voff[k];
    const signed char* pv = f->piv + fc;
    const i64* od = f->order + fc;
    zc* own = x + fc;
    for (i64 c = 0; c < w; ++c) y[c] = own[c];
    for (i64 c = w - 1; c >= 0; --c) {
      i64 q, end;
      if (pv[c] == PIV_SECOND && c > 0) { q = c - 1; end = c + 1; }
      else if (pv[c] == PIV_FIRST) { q = c + 1; end = c + 2; }
      else { q = c; end = c + 1; }
      zc sum = y[q];
      for (i64 r = end; r < w; ++r) sum = zsub(sum, zmul(AT(t, m, r, q), y[r]));
      for (i64 r = w; r < m; ++r) sum = zsub(sum, zmul(AT(t, m, r, q), x[rows[r]]));
      y[q] = sum;
    }
    for (i64 pos = 0; pos < w; ++pos) own[od[pos]] = y[pos];
  }
  for (i64 i = 0; i < n; ++i) b[f->perm[i]] = x[i];
  free(x);
  free(y);
}

int orc_zldlt_solve(const orc_factor* f, double* bs_pairs, i64 nrhs) {
  if (!f->have_values || !f->piv) return ORC_INTERNAL;
#pragma omp parallel for schedule(dynamic, 1) if (nrhs > 1)
  for (i64 r = 0; r < nrhs; ++r) zldlt_solve_one(f, (zc*)bs_pairs + r * f->n);
  return ORC_OK;
}

/* the complex factor's trapezoids as (re, im) pairs, 2 * orc_entries(f) doubles */
void orc_zexport_values(const orc_factor* f, double* out) {
  if (f->val) memcpy(out, f->val, sizeof(double) * 2 * (size_t)f->voff[f->ns]);
}
