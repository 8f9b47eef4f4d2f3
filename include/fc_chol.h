/*
 * fc_chol.h — C ABI of the MI355X supernodal Cholesky (numeric phase on one GPU, gfx950).
 *
 * The entry points mirror the factor/solve API of FieldCraft's solver crate
 * (crates/solver/src/cholesky.rs), so a host program that calls Factor::new / solve /
 * solve_many can call these instead. Plain pointers and sizes only; the library owns device
 * memory, the caller owns every array it passes. Every entry point returns an fc_status and
 * fills *err (when err is not NULL); none aborts.
 *
 * Matrix input is the reference's UpperColumns (cholesky.rs:99-108): column k holds the rows
 * i <= k of the upper triangle, sorted, with 64-bit ptr/idx. Column numbers in errors are the
 * caller's original columns, as FactorError names them (cholesky.rs:1506-1513).
 */
#ifndef FC_CHOL_H
#define FC_CHOL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* FactorError (cholesky.rs:346-361), plus argument and device failures. */
typedef enum {
  FC_OK = 0,
  FC_NOT_POSITIVE_DEFINITE = 1, /* FactorError::NotPositiveDefinite { column }            */
  FC_SINGULAR = 2,              /* FactorError::Singular { column } (LDL^T only)            */
  FC_TOO_LARGE = 3,             /* FactorError::TooLarge { entries, limit }                 */
  FC_OUT_OF_MEMORY = 4,         /* FactorError::OutOfMemory { entries }                     */
  FC_INTERNAL = 5,              /* FactorError::Internal(String)                            */
  FC_CANCELLED = 6,             /* FactorError::Cancelled                                   */
  FC_INVALID_ARGUMENT = 7,      /* a malformed pattern, length or option                    */
  FC_DEVICE = 8,                /* a HIP runtime error                                      */
  FC_DEGENERATE_ELEMENT = 9,    /* assembly: an inverted or degenerate element (its index in
                                   element order in fc_error.column), Error::Degenerate of
                                   crates/fem/src/element.rs:326-328                         */
  FC_NOT_CONVERGED = 10         /* Lanczos: Error::NotConverged (crates/solver/src/eigen.rs)  */
} fc_status;

typedef struct {
  int32_t kind;      /* fc_status */
  int32_t pad;
  uint64_t column;   /* failing original column (NOT_POSITIVE_DEFINITE, SINGULAR); the
                        element index (DEGENERATE_ELEMENT)                              */
  uint64_t entries;  /* entries of L (TOO_LARGE, OUT_OF_MEMORY)                         */
  uint64_t limit;    /* the cap (TOO_LARGE)                                             */
  char message[256];
} fc_error;

typedef struct fc_symbolic fc_symbolic;
typedef struct fc_factor fc_factor;
typedef struct fc_matrix fc_matrix;

/* Arithmetic of the numeric phase. */
typedef enum {
  FC_FP64 = 0,  /* the plain method: every step in FP64 on the FP64 matrix cores           */
  FC_MIXED = 1, /* Schur updates in BF16 (power-of-two scaled, FP32 accumulation), the rows
                   below a panel (TRSM) on the FP32 matrix cores, diagonal blocks factored
                   and inverted in FP32 arithmetic (a block with a pivot below 1e-3 |A_jj|
                   is factored again in FP64 arithmetic, where the reference's pivot test
                   decides); the factor is stored in FP32; solutions are refined in FP64   */
  FC_FP32 = 2   /* every step in FP32 (diagonal blocks as FC_MIXED's), factor stored in FP32 */
} fc_precision;

typedef struct {
  int32_t precision;              /* fc_precision                                           */
  int32_t device;                 /* HIP device ordinal                                     */
  void* stream;                   /* hipStream_t to run on; NULL: the library's own stream  */
  const volatile int32_t* cancel; /* polled between levels of the front tree; NULL: none    */
  int32_t bf16_terms;             /* FC_MIXED: 1 = one BF16 product per operand pair,
                                     3 = split FP32 operands (hi*hi + hi*lo + lo*hi)       */
  int32_t use_graph;              /* bit 0: capture the factorization's launches into a hipGraph
                                     on the first call and replay it (the cancel flag is then
                                     polled once per call and no kind is timed); bit 1: the
                                     same for the solves, one graph per right-hand-side count */
  int32_t fallback;               /* FC_MIXED, FC_FP32: when a pivot fails the test in the
                                     reduced precision (an SPD matrix can lose definiteness
                                     there), 0 (the default) refactors in the next precision
                                     (mixed -> FP32 -> FP64, the last on FP64 storage) and
                                     reports FC_NOT_POSITIVE_DEFINITE only when the FP64 factor
                                     fails too, naming its column; 1 reports the first
                                     precision's failure as it is. fc_factor_precision tells
                                     which precision the factor holds                      */
  int32_t pad0;
  int32_t profile_mask;           /* bit k: time every launch of kernel kind k with HIP events
                                     (fc_factor_profile); 0: none                           */
  int32_t pad;
} fc_factor_opts;

/* Kernel kinds of the numeric phase (fc_factor_profile, fc_symbolic_info.kind_*). */
enum {
  FC_KIND_EXTEND_L = 0, /* children's update matrices into the fronts' L parts               */
  FC_KIND_EXTEND_U = 1, /* ... into the fronts' update matrices                               */
  FC_KIND_UPDATE = 2,   /* panel minus its products with the front's earlier columns (GEMM)  */
  FC_KIND_POTRF = 3,    /* diagonal blocks                                                   */
  FC_KIND_TRSM = 4,     /* rows below a panel against its diagonal block                     */
  FC_KIND_SYRK = 5,     /* update matrices U = -L21 L21^T                                    */
  FC_KIND_REDUCE = 6,   /* sums of the pieces of a GEMM split along k                        */
  FC_KIND_SMALL = 7,    /* small fronts (<= 64 columns, <= 768 rows below, an update matrix  */
                        /* of <= 2e6 multiply-adds): diagonal block, rows below and update   */
                        /* matrix of each by one workgroup, one launch a level              */
  FC_KINDS = 8
};

typedef struct {
  uint64_t n, supernodes, entries, rows, updates, tasks, top;
  uint64_t fronts, levels, stored; /* execution plan: fronts, tree levels, stored L entries  */
  double flops;                    /* sum over supernodes of sum_{c<w} (m - c)^2            */
  double front_flops;              /* the same over fronts (padding rows included)          */
  uint64_t l_bytes_fp64, u_bytes_fp64; /* device storage of L and of the update arena      */
  int32_t by_blocks;               /* analysed on the 3x3 node blocks' graph                */
  int32_t pad;
  uint64_t launches;               /* kernel launches of one numeric factorization          */
  double kind_flops[8];            /* algorithmic flops per kernel kind (FC_KIND_*)          */
  double kind_entries[8];          /* update-matrix entries added per extend-add kind       */
  uint64_t solve_bytes;            /* device work bytes of a solve per right-hand side: the
                                      permuted vector, the forward result and the fronts'
                                      contributions (fc_solve_device_noalloc's measure)      */
} fc_symbolic_info;

/* --- analysis ---------------------------------------------------------------------------- */

/* Symbolic::new (cholesky.rs:1213-1378): elimination tree, column counts, postorder,
 * fundamental supernodes, row structures, update lists and subtree tasks, on the host.
 * Refuses with FC_TOO_LARGE before planning any numeric memory when the factor would hold
 * more than max_entries entries (cholesky.rs:1216-1219). */
int fc_analyze(uint64_t n, const uint64_t* ptr, const uint64_t* idx, uint64_t max_entries, fc_symbolic** out,
               fc_error* err);

int fc_symbolic_info_get(const fc_symbolic* s, fc_symbolic_info* info);

/* The reference's symbolic outputs (postorder numbering; arrays sized by fc_symbolic_info):
 * perm[n] (cholesky.rs:1197), first[ns+1], rptr[ns+1], rows[rows] (Structure, :392-397),
 * uptr[ns+1] + updates[3*updates] as {k, p1, p2} (Update, :382-389), tasks[2*tasks] as
 * [start, end) and top[top] (:1203-1205). Any pointer may be NULL. */
int fc_symbolic_export(const fc_symbolic* s, uint64_t* perm, uint64_t* first, uint64_t* rptr, uint64_t* rows,
                       uint64_t* uptr, uint64_t* updates, uint64_t* tasks, uint64_t* top);

void fc_symbolic_free(fc_symbolic* s);

/* --- schedule and tuner ------------------------------------------------------------------ */

/* No counterpart in the reference (its CPU factorization has no launch schedule; its own
 * work-division constants are fixed, cholesky.rs:48-97): the numeric phase's schedule on the
 * GPU, a property of the analysis. Panels of 128 columns are factored left-looking inside
 * each front; while the caller's stream factors a slab of `slab` panels, a second stream
 * subtracts from the next slab its products with the columns before the current one (one
 * product slab x 128 columns wide). GEMM launches with fewer output tiles than a target are
 * split along k (pieces of at least 512) and their pieces summed in a fixed order. The
 * schedule changes launch shapes and the order of floating-point sums, not the algorithm:
 * every schedule computes the same factor to rounding. */
typedef struct {
  int32_t slab;            /* panels per side-stream slab update, 1..16 (1: one-panel lookahead) */
  int32_t split_target;    /* tiles the update-matrix (SYRK) launches are split up to, 1..65536 */
  int32_t split_side_fp64; /* the same for the FP64 factor's side-stream slab updates           */
  int32_t split_side_fp32; /* ... and for the FP32 and mixed factors'                           */
} fc_schedule;

/* The analysis's current schedule (the built-in default until set or tuned). */
int fc_symbolic_schedule_get(const fc_symbolic* s, fc_schedule* out);

/* Rebuilds the analysis's launch program for a new schedule. Factors created afterwards use
 * it; a factor already made keeps the program it was created with. FC_INVALID_ARGUMENT for
 * values out of range. */
int fc_symbolic_schedule_set(fc_symbolic* s, const fc_schedule* sched, fc_error* err);

typedef struct {
  int32_t precision; /* fc_precision the candidates are timed in                           */
  int32_t device;    /* HIP device ordinal                                                  */
  void* stream;      /* hipStream_t; NULL: a stream of the tuner's own                      */
  int32_t repeats;   /* timed factorizations per candidate, the fastest counting (0: 2); one
                        untimed factorization precedes them                                  */
  int32_t pad;
} fc_tune_opts;

#define FC_TUNE_MAX 32
typedef struct {
  fc_schedule chosen;              /* installed in the analysis when the call succeeds          */
  double seconds_start;            /* the schedule the analysis had before the call              */
  double seconds_chosen;
  int32_t candidates;              /* schedules timed (tried[0] is the starting one)             */
  int32_t pad;
  fc_schedule tried[FC_TUNE_MAX];
  double tried_seconds[FC_TUNE_MAX];
} fc_tune_report;

/* No counterpart in the reference: an autotuner for the schedule. It factors the caller's
 * matrix (upper values in device memory, as fc_factor_numeric_device) once per candidate
 * schedule, timed with HIP events, searching one parameter at a time from the analysis's
 * current schedule: slab in {1, 2, 4, 8}, then the side-stream split target of the precision
 * in {128, 256, 384, 512}, then the SYRK split target in {512, 1024, 2048}; each parameter keeps
 * its fastest value. The fastest schedule is installed in the analysis (as
 * fc_symbolic_schedule_set) and reported. Errors of a factorization (FC_NOT_POSITIVE_DEFINITE,
 * FC_OUT_OF_MEMORY, ...) end the search and are returned; the analysis then keeps its schedule. */
int fc_tune(fc_symbolic* s, const double* upper_val_device, uint64_t nnz, const fc_tune_opts* opts,
            fc_tune_report* report, fc_error* err);

/* --- numeric phase ----------------------------------------------------------------------- */

/* Factor::from_symbolic (cholesky.rs:1549-1555): the numeric factorization on the GPU of a
 * matrix with the analysed pattern (values of the upper columns, nnz = ptr[n]). Fails with
 * FC_NOT_POSITIVE_DEFINITE naming the first failing column in postorder, translated to the
 * original column (cholesky.rs:1506-1528), with FC_OUT_OF_MEMORY when device memory is short,
 * and with FC_CANCELLED when *opts->cancel is set. */
int fc_factor_numeric(const fc_symbolic* s, const double* upper_val, uint64_t nnz, const fc_factor_opts* opts,
                      fc_factor** out, fc_error* err);

/* fc_factor_numeric with the values already in device memory (no host transfer). */
int fc_factor_numeric_device(const fc_symbolic* s, const double* upper_val_device, uint64_t nnz,
                             const fc_factor_opts* opts, fc_factor** out, fc_error* err);

/* Factor::new (cholesky.rs:1545-1547): fc_analyze then fc_factor_numeric. */
int fc_factor_new(uint64_t n, const uint64_t* ptr, const uint64_t* idx, const double* val, uint64_t max_entries,
                  const fc_factor_opts* opts, fc_factor** out, fc_error* err);

/* The same factor's values refactored for a new matrix of the same pattern (one Symbolic
 * serving many numeric factorizations, cholesky.rs:1390-1395). */
int fc_refactor(fc_factor* f, const double* upper_val, uint64_t nnz, fc_error* err);
int fc_refactor_device(fc_factor* f, const double* upper_val_device, uint64_t nnz, fc_error* err);

/* Factor::entries (cholesky.rs:1557-1559). */
uint64_t fc_entries(const fc_factor* f);

/* Factor::solve (cholesky.rs:1562-1564): A x = b in place, host memory, FP64 result. An
 * FC_MIXED or FC_FP32 factor solves with its FP32 factor (no refinement: fc_solve_refined). */
int fc_solve(const fc_factor* f, double* b, uint64_t n, fc_error* err);

/* Factor::solve_many (cholesky.rs:1568-1570): each of bs[0..nrhs) solved in place, each the
 * same bit for bit as fc_solve. */
int fc_solve_many(const fc_factor* f, double* const* bs, uint64_t nrhs, uint64_t n, fc_error* err);

/* Device flavour: nrhs right-hand sides in device memory, column-major with leading dimension
 * ldb, solved in place on the factor's stream. */
int fc_solve_device(const fc_factor* f, double* b, uint64_t n, uint64_t nrhs, uint64_t ldb, fc_error* err);

/* No counterpart in the reference: fc_solve_device for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same solution, bit for bit, with no hipMalloc, hipFree or allocating library call: the
 * work vectors live in a fixed 128 MiB device array of the library (one per device it runs on,
 * held until the process ends). A solve needs fc_symbolic_info.solve_bytes of it per right-hand
 * side (about 36 bytes an equation); right-hand sides are taken as many a pass as fit, and a
 * factor needing more than the array for one right-hand side is refused with FC_TOO_LARGE
 * (err->entries: the bytes needed, err->limit: the array's). Calls on one device share the array
 * and are serialised by the library; launches are eager (never a graph), on the factor's stream,
 * which is synchronised before the call returns. */
int fc_solve_device_noalloc(const fc_factor* f, double* b, uint64_t n, uint64_t nrhs, uint64_t ldb, fc_error* err);

/* The factor values in the reference layout (cholesky.rs:393-420): every supernode's lower
 * trapezoid, column by column (column c holds rows c..m of its row structure), supernodes in
 * order; `out` holds `entries` doubles. For parity checks. */
int fc_factor_export(const fc_factor* f, double* out, uint64_t len, fc_error* err);

/* Seconds the last numeric factorization took on the device (events around the launches),
 * fallback refactorizations included. */
double fc_factor_seconds(const fc_factor* f);

/* The precision (fc_precision) of the factor's last successful factorization: the one asked for,
 * or the one a fallback (fc_factor_opts.fallback) reached; -1 for NULL. Each factorization
 * starts from the precision asked for, except that a factor moved to FP64 storage by a fallback
 * stays in FP64. */
int fc_factor_precision(const fc_factor* f);

/* Kernel kind `kind` in the last numeric factorization: launches timed and their summed
 * device seconds (when opts.profile_mask selected the kind), and the kind's algorithmic flops
 * and extend-add entries over the whole factorization. */
int fc_factor_profile(const fc_factor* f, int32_t kind, double* launches, double* seconds, double* flops,
                      double* entries);

void fc_factor_free(fc_factor* f);

/* --- refinement -------------------------------------------------------------------------- */

/* The matrix by full rows (Csr, crates/solver/src/sparse.rs), uploaded once for the FP64
 * residuals of refinement (Csr::mul, sparse.rs:42-50). */
int fc_matrix_new(uint64_t n, const uint64_t* ptr, const uint64_t* idx, const double* val, int32_t device,
                  fc_matrix** out, fc_error* err);
void fc_matrix_free(fc_matrix* a);

/* Declares rows [offsets[s], offsets[s + 1]) (s < count, offsets[0] = 0, offsets[count] = n,
 * nondecreasing) as independent segments: the diagonal blocks of a batch of models factored as
 * one block-diagonal matrix. Refinement then measures each segment's backward error with its
 * own norms and stops when every segment of every column meets the target. No counterpart in
 * the reference, which factors one model at a time; default: one segment. */
int fc_matrix_set_segments(fc_matrix* a, uint64_t count, const uint64_t* offsets, fc_error* err);

typedef struct {
  int32_t method;        /* 0: refinement x += solve(b - A x); 1: conjugate gradients
                            preconditioned by the factor; 2: 0, switching to 1 when a step
                            does not shrink the backward error                                */
  int32_t max_iter;      /* 0: 50                                                            */
  double target_berr;    /* 0: 1e-12; berr = ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) */
} fc_refine_opts;

typedef struct {
  int32_t iterations;    /* solves after the first                                         */
  int32_t converged;
  double berr;           /* final normwise backward error                                   */
  double seconds;        /* device time of the refined solve                                */
  double berr_history[64];
} fc_refine_report;

/* Symmetric::solve_refined (ldlt.rs:524-563) made general: x solves A x = b to the target
 * normwise backward error, with FP64 residuals A x on the device. b and x are host arrays
 * of n doubles (x may alias b). */
int fc_solve_refined(const fc_factor* f, const fc_matrix* a, const double* b, double* x, uint64_t n,
                     const fc_refine_opts* opts, fc_refine_report* report, fc_error* err);

/* fc_solve_refined for nrhs right-hand sides in device memory: B and X are n x nrhs,
 * column-major (leading dimension n), distinct. Columns are refined together, MAX 64 at a
 * time, each with its own scalars; berr (nrhs doubles on the host, or NULL) receives each
 * column's final backward error (the largest over the segments); the report holds the worst.
 * Solve many load cases at once (Factor::solve_many, cholesky.rs:1568-1570, refined). */
int fc_solve_refined_many_device(const fc_factor* f, const fc_matrix* a, const double* B, double* X, uint64_t n,
                                 uint64_t nrhs, const fc_refine_opts* opts, fc_refine_report* report, double* berr,
                                 fc_error* err);

/* fc_solve_refined with b and x in device memory (n doubles each, distinct). */
int fc_solve_refined_device(const fc_factor* f, const fc_matrix* a, const double* b, double* x, uint64_t n,
                            const fc_refine_opts* opts, fc_refine_report* report, fc_error* err);

/* --- assembly of the factor's input ------------------------------------------------------ */

/* The step before the factorization (crates/solver/src/analysis.rs:199-233 assemble,
 * sparse.rs:199-243 add_block / add_scalar, constraints.rs:820-906 reduce, sparse.rs:81-87
 * add_scaled, cholesky.rs:110-144 UpperColumns::from_symmetric) on the device: the solid
 * element matrices of a mesh added in element order into the upper triangle by columns of
 * K_ff, M_ff or K_ff - shift M_ff, the same bit for bit as the reference's, written where
 * fc_refactor_device / fc_factor_numeric_device read the values. */
typedef struct fc_assembly fc_assembly;

enum {
  FC_ELEMENT_HEX8 = 0,  /* 8-node hexahedra, full 2x2x2 integration (fem element.rs:146-155)  */
  FC_ELEMENT_TET10 = 1  /* 10-node tetrahedra, 4-point stiffness, conical 4x4x4 mass (:82-91) */
};

enum {
  FC_ASSEMBLE_K = 0,       /* the stiffness K_ff                                            */
  FC_ASSEMBLE_M = 1,       /* the consistent mass M_ff (10-node tetrahedra, density > 0)    */
  FC_ASSEMBLE_SHIFTED = 2  /* K_ff + (-shift) M_ff (Csr::add_scaled)                        */
};

/* Builds the gather plan of a mesh on the host and uploads the mesh. xyz: nodes x 3; conn:
 * elements x (8 | 10) node numbers in the model's node order (fem shape.rs:3-11); young: one
 * value per element (n_young = elements) or one for all (n_young = 1); isotropic materials
 * (fem material.rs:40-54); density > 0 (10-node tetrahedra) also plans the mass. eq_dof:
 * neq free equations, equation -> mesh dof 3 node + direction (constraints.rs:755-771, a plain
 * selection); ptr/idx: the upper columns of K_ff over those equations (what fc_analyze was
 * given). timing != 0 times each fc_assemble_device's kernels with HIP events
 * (fc_assembly_info). Refuses material constants that are not a stable material
 * (material.rs:22-27) and patterns that lack an entry two elements' nodes couple. */
int fc_assembly_new(int32_t element_kind, uint64_t nodes, const double* xyz, uint64_t elements, const int64_t* conn,
                    const double* young, uint64_t n_young, double poisson, double density, uint64_t neq,
                    const int64_t* eq_dof, const uint64_t* ptr, const uint64_t* idx, int32_t device, int32_t timing,
                    fc_assembly** out, fc_error* err);

/* Assembles `what` into upper_val_device (nnz doubles of device memory, the pattern's entry
 * count) on `stream` (NULL: the assembly's own stream), and waits for it. The element matrices
 * are recomputed on every call. FC_DEGENERATE_ELEMENT names the first degenerate element in
 * element order, as the reference's sequential loop reports it (analysis.rs:203-233); the
 * output is then undefined. */
int fc_assemble_device(fc_assembly* a, int32_t what, double shift, double* upper_val_device, uint64_t nnz,
                       void* stream, fc_error* err);

/* The plan's contribution counts (stiffness, mass) and, with timing on, the device seconds of
 * the last call's element kernels and gather kernel. */
int fc_assembly_info(const fc_assembly* a, uint64_t* contributions_k, uint64_t* contributions_m, double* elements_s,
                     double* gather_s);
void fc_assembly_free(fc_assembly* a);

/* --- LDL^T ------------------------------------------------------------------------------ */

/* LdltFactor::from_symbolic (crates/solver/src/ldlt.rs:378-391): P A P^T = L D L^T with bounded
 * Bunch-Kaufman (rook) pivoting inside supernodes, on the Cholesky analysis `s` (fc_analyze), on
 * the device; LdltFactor::new (ldlt.rs:374-376) is fc_analyze followed by this call. The factor,
 * its pivots and the solutions are the reference's bit for bit (each entry's sums in the
 * reference's order). FC_SINGULAR names the first column in postorder with no acceptable pivot,
 * as an original column (ldlt.rs:300-329). The analysis keeps a host plan built by its first
 * LDL^T factorization for the later ones. */
typedef struct fc_ldlt fc_ldlt;
int fc_ldlt_numeric(const fc_symbolic* s, const double* upper_val, uint64_t nnz, int32_t device, fc_ldlt** out,
                    fc_error* err);
/* LdltFactor::solve (ldlt.rs:401-404): b (host, n doubles) in place. */
int fc_ldlt_solve(const fc_ldlt* f, double* b, uint64_t n, fc_error* err);
/* LdltFactor::inertia and two_by_two (ldlt.rs:393-445). */
int fc_ldlt_inertia(const fc_ldlt* f, uint64_t* positive, uint64_t* negative, uint64_t* two_by_two);
/* The supernodes' trapezoids (L below the pivots, D on them) in the reference's layout: the
 * factor's entry count of doubles, twice that (re, im pairs) for a complex factor. */
int fc_ldlt_export(const fc_ldlt* f, double* out, uint64_t len, fc_error* err);
void fc_ldlt_free(fc_ldlt* f);

/* LdltFactor::<Complex64>::from_symbolic (ldlt.rs:378-391 over scalar.rs:115-131): the
 * complex-symmetric A = A^T (never Hermitian) factored as P A P^T = L D L^T with the same pivoting,
 * the modulus for its tests and num-complex's arithmetic; upper_val holds nnz (re, im) pairs in
 * the analysed pattern. The factor and the solutions are the oracle's restatement bit for bit
 * where the device's hypot gives the host's moduli. fc_ldlt_inertia gives only the 2 x 2 count
 * (no inertia for complex values); fc_ldlt_export and fc_ldlt_free serve both kinds. */
int fc_zldlt_numeric(const fc_symbolic* s, const double* upper_val, uint64_t nnz, int32_t device, fc_ldlt** out,
                     fc_error* err);
/* LdltFactor::<Complex64>::solve (ldlt.rs:404-408): b (host, n (re, im) pairs) in place. */
int fc_zldlt_solve(const fc_ldlt* f, double* b, uint64_t n, fc_error* err);

/* --- shift-invert block Lanczos ----------------------------------------------------------- */

/* What lanczos (crates/solver/src/eigen.rs:281) is given: Wanted (eigen.rs:178-185), the block
 * size, the basis cap and the relative residual tolerance of a converged Ritz pair. */
typedef struct {
  uint64_t count;      /* modes to return with min <= lambda <= max                        */
  double min, max;     /* bounds on lambda (-inf, +inf for none)                             */
  uint32_t block;      /* vectors per block, 1..8 (the reference's BLOCK = 6)               */
  uint32_t pad;
  uint64_t max_basis;  /* most Lanczos vectors (lanczos_basis, analysis.rs:642-645)          */
  double tolerance;    /* ||B_{j+1} s|| <= tolerance theta (TOLERANCE = 1e-9, analysis.rs:653) */
} fc_lanczos_opts;

typedef struct {
  uint64_t found;           /* modes returned                                                */
  uint64_t basis;           /* Lanczos vectors used (Eigenpairs::basis)                      */
  uint64_t negative_count;  /* converged negative Ritz values (Eigenpairs::negative)         */
  double negative[16];      /* their eigenvalues, largest |theta| first (the first 16)       */
  double seconds;           /* device time of the run                                        */
} fc_lanczos_report;

/* The lowest modes of K phi = lambda M phi with lambda >= shift, by the shifted block Lanczos
 * method with full reorthogonalization (eigen.rs:281-466; ShiftInvert, analysis.rs:617-635),
 * entirely on the device: f is the factor of K - shift M (fc_refactor_device after
 * fc_assemble_device's FC_ASSEMBLE_SHIFTED), m the mass matrix by full rows (fc_matrix_new).
 * values (host, opts->count doubles) receives the eigenvalues ascending; vectors_device (n x
 * count column-major doubles of device memory, or NULL) the M-orthonormal eigenvectors, each
 * with its largest component positive. The start block is the reference's seeded
 * xorshift64*. Errors: FC_NOT_CONVERGED (the basis cap, a block that loses rank), FC_CANCELLED
 * (the factor's cancel flag, polled once per block). */
int fc_lanczos_device(const fc_factor* f, const fc_matrix* m, double shift, const fc_lanczos_opts* opts,
                      double* values, double* vectors_device, fc_lanczos_report* report, fc_error* err);
/* The same Lanczos run with an LDL^T factor of K - shift M (fc_ldlt_numeric): a shift between
 * modes, where K - shift M is indefinite (Symmetric with Kind::Auto under ShiftInvert,
 * analysis.rs:291-293, eigen.rs:187-211). report->negative receives the converged negative Ritz
 * values as eigenvalues, largest |theta| first (eigen.rs:413-426): possibly fewer than the modes
 * below the shift. The factor's stream runs the solves; calls on one factor must not overlap. */
int fc_lanczos_device_ldlt(const fc_ldlt* f, const fc_matrix* m, double shift, const fc_lanczos_opts* opts,
                           double* values, double* vectors_device, fc_lanczos_report* report, fc_error* err);

/* Library version string. */
const char* fc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* FC_CHOL_H */
