/*
 * ORACLE — test infrastructure only (tests/, smoke() and bench.py's CPU-baseline leg).
 *
 * Plain-C restatement of the step before the factorization in FieldCraft's static analysis:
 * the solid element matrices, their assembly in element order onto the node graph, the
 * reduction to the free equations, K - sigma*M, and the upper triangle by columns that the
 * factor reads. Only 8-node hexahedra (full 2x2x2 integration) and 10-node tetrahedra
 * (4-point stiffness, Stroud conical 4x4x4 consistent mass) with isotropic materials: the
 * element kinds of the factor's workloads.
 *
 * Followed, step for step and in the same order of floating-point operations (no fused
 * multiply-add: built with -ffp-contract=off):
 *   crates/fem/src/quadrature.rs:27-45 (Gauss-Legendre), 85-113 (Hex, Tet4, TetConical points)
 *   crates/fem/src/shape.rs:60-73 (tet4), 75-93 (tet10), 95-104 (hex8)
 *   crates/model/src/mesh.rs:185 (tet10 edges), 254-269 (normalized Jacobian, admissibility)
 *   crates/fem/src/dense.rs:91-116 (inverse3, positive_inverse), 169-172 (det3)
 *   crates/fem/src/material.rs:40-54 (isotropic elasticity)
 *   crates/fem/src/element.rs:301-324 (Jacobian, gradients), 416-441 (stiffness, mass),
 *                              723-774 (b_block, accumulate_symmetric)
 *   crates/solver/src/analysis.rs:199-233 (element matrices added in element order)
 *   crates/solver/src/sparse.rs:199-243 (add_block, add_scalar), 81-87 (add_scaled)
 *   crates/solver/src/constraints.rs:820-906 (reduce: plain selection, each sum from 0.0)
 *   crates/solver/src/cholesky.rs:110-144 (UpperColumns::from_symmetric: row k's entries with
 *                                          column <= k are column k)
 * Rust's float sums (`Iterator::sum`) start from -0.0 or 0.0 depending on the toolchain; the
 * sums here start from 0.0. The two differ only in the sign of an all-zero sum, and every such
 * value is added into a matrix entry that starts at +0.0 before anything else reads it, so
 * the assembled matrices are the same bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef int64_t i64;

/* quadrature.rs:28-33 */
static const double G2X[2] = {-0.5773502691896258, 0.5773502691896258};
static const double G2W[2] = {1.0, 1.0};
static const double G4X[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
static const double G4W[4] = {0.3478548451374539, 0.6521451548625461, 0.6521451548625461, 0.3478548451374539};

/* shape.rs:50-51 */
static const double HEX_C[8][3] = {{-1., -1., -1.}, {1., -1., -1.}, {1., 1., -1.}, {-1., 1., -1.},
                                   {-1., -1., 1.},  {1., -1., 1.},  {1., 1., 1.},  {-1., 1., 1.}};
/* mesh.rs:185: (corner, corner, mid-side node) */
static const int TET10_EDGES[6][3] = {{0, 1, 4}, {1, 2, 5}, {2, 0, 6}, {0, 3, 7}, {1, 3, 8}, {2, 3, 9}};

typedef struct {
  double xi[3], w;
} Point;

/* quadrature.rs:87-99: xi fastest, then eta, then zeta */
static int hex2_points(Point* out) {
  int k = 0;
  for (int z = 0; z < 2; ++z)
    for (int y = 0; y < 2; ++y)
      for (int x = 0; x < 2; ++x) {
        out[k].xi[0] = G2X[x];
        out[k].xi[1] = G2X[y];
        out[k].xi[2] = G2X[z];
        out[k].w = G2W[x] * G2W[y] * G2W[z];
        ++k;
      }
  return k;
}

/* quadrature.rs:100-106 */
static int tet4_points(Point* out) {
  const double a = 0.5854101966249685, b = 0.1381966011250105;
  const double p[4][3] = {{b, b, b}, {a, b, b}, {b, a, b}, {b, b, a}};
  for (int k = 0; k < 4; ++k) {
    for (int c = 0; c < 3; ++c) out[k].xi[c] = p[k][c];
    out[k].w = 1.0 / 24.0;
  }
  return 4;
}

/* quadrature.rs:107-118: Stroud's conical product, 4 points per direction */
static int tet_conical4_points(Point* out) {
  double g[4], gw[4];
  for (int i = 0; i < 4; ++i) {
    g[i] = (G4X[i] + 1.0) / 2.0;
    gw[i] = G4W[i] / 2.0;
  }
  int k = 0;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {
        double u = g[a], v = g[b], w = g[c];
        out[k].xi[0] = u;
        out[k].xi[1] = v * (1.0 - u);
        out[k].xi[2] = w * (1.0 - u) * (1.0 - v);
        out[k].w = gw[a] * gw[b] * gw[c] * (1.0 - u) * (1.0 - u) * (1.0 - v);
        ++k;
      }
  return k;
}

/* shape.rs:95-104 */
static void hex8_shape(const double xi[3], double n[8], double dn[8][3]) {
  for (int i = 0; i < 8; ++i) {
    const double* c = HEX_C[i];
    double f0 = 1.0 + xi[0] * c[0], f1 = 1.0 + xi[1] * c[1], f2 = 1.0 + xi[2] * c[2];
    n[i] = f0 * f1 * f2 / 8.0;
    dn[i][0] = c[0] * f1 * f2 / 8.0;
    dn[i][1] = c[1] * f0 * f2 / 8.0;
    dn[i][2] = c[2] * f0 * f1 / 8.0;
  }
}

/* shape.rs:60-93 */
static void tet10_shape(const double xi[3], double n[10], double dn[10][3]) {
  const double l[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
  const double dl[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
  for (int i = 0; i < 4; ++i) {
    n[i] = l[i] * (2.0 * l[i] - 1.0);
    for (int c = 0; c < 3; ++c) dn[i][c] = (4.0 * l[i] - 1.0) * dl[i][c];
  }
  for (int e = 0; e < 6; ++e) {
    int a = TET10_EDGES[e][0], b = TET10_EDGES[e][1], m = TET10_EDGES[e][2];
    n[m] = 4.0 * l[a] * l[b];
    for (int c = 0; c < 3; ++c) dn[m][c] = 4.0 * (dl[a][c] * l[b] + l[a] * dl[b][c]);
  }
}

/* mesh.rs:257-269 */
static int admissible(const double j[3][3]) {
  const double *a = j[0], *b = j[1], *c = j[2];
  double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0]);
  double norms = 1.0; /* Iterator::product starts from 1.0 */
  for (int r = 0; r < 3; ++r) norms = norms * sqrt(j[r][0] * j[r][0] + j[r][1] * j[r][1] + j[r][2] * j[r][2]);
  double q = det / norms;
  if (isfinite(q)) {
    if (q < -1.0) q = -1.0;
    if (q > 1.0) q = 1.0;
  } else {
    q = 0.0;
  }
  return q > 1e-12;
}

/* dense.rs:91-116, 169-172; 0 when not admissible or singular */
static int positive_inverse(const double m[3][3], double inv[3][3], double* det_out) {
  if (!admissible(m)) return 0;
  double det = m[0][0] * (m[1][1] * m[2][2] - m[1][2] * m[2][1]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
               m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
  if (!(fabs(det) > 0.0 && isfinite(det))) return 0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      int j1 = (j + 1) % 3, j2 = (j + 2) % 3, i1 = (i + 1) % 3, i2 = (i + 2) % 3;
      inv[i][j] = (m[j1][i1] * m[j2][i2] - m[j1][i2] * m[j2][i1]) / det;
    }
  *det_out = det;
  return 1;
}

/* element.rs:301-324: shape at xi, Jacobian J[a][b] = sum_i dN_i/dxi_a x_ib, physical gradients */
static int gradients(int nn, const double (*x)[3], const double n[], const double (*dn)[3], double (*g)[3],
                     double* det) {
  (void)n;
  double j[3][3] = {{0.0}};
  for (int i = 0; i < nn; ++i)
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) j[a][b] += dn[i][a] * x[i][b];
  double inv[3][3];
  if (!positive_inverse(j, inv, det)) return 0;
  if (g)
    for (int i = 0; i < nn; ++i)
      for (int b = 0; b < 3; ++b) g[i][b] = inv[b][0] * dn[i][0] + inv[b][1] * dn[i][1] + inv[b][2] * dn[i][2];
  return 1;
}

/* material.rs:40-54 */
static void isotropic(double young, double poisson, double d[6][6]) {
  double lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  double mu = young / (2.0 * (1.0 + poisson));
  memset(d, 0, 36 * sizeof(double));
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) d[i][j] = lambda;
    d[i][i] = lambda + 2.0 * mu;
    d[i + 3][i + 3] = mu;
  }
}

/* element.rs:723-725 */
static void b_block(const double g[3], double b[6][3]) {
  const double v[6][3] = {{g[0], 0.0, 0.0}, {0.0, g[1], 0.0}, {0.0, 0.0, g[2]},
                          {g[1], g[0], 0.0}, {g[2], 0.0, g[0]}, {0.0, g[2], g[1]}};
  memcpy(b, v, sizeof(v));
}

/* element.rs:745-774: the blocks j >= i, upper triangle of the diagonal blocks, mirrored */
static void accumulate_symmetric(double* out, int nn, const double (*g)[3], const double d[6][6], double w) {
  const int cols = 3 * nn;
  for (int j = 0; j < nn; ++j) {
    double bj[6][3], db[6][3];
    b_block(g[j], bj);
    for (int s = 0; s < 6; ++s)
      for (int c = 0; c < 3; ++c) {
        double sum = 0.0;
        for (int t = 0; t < 6; ++t) sum += d[s][t] * bj[t][c];
        db[s][c] = sum;
      }
    for (int i = 0; i <= j; ++i) {
      double bi[6][3];
      b_block(g[i], bi);
      for (int r = 0; r < 3; ++r)
        for (int c = (i == j ? r : 0); c < 3; ++c) {
          double sum = 0.0;
          for (int s = 0; s < 6; ++s) sum += bi[s][r] * db[s][c];
          double v = w * sum;
          int row = 3 * i + r, col = 3 * j + c;
          out[row * cols + col] += v;
          if (row != col) out[col * cols + row] += v;
        }
    }
  }
}

/* element.rs:416-422: the stiffness, 3n x 3n row-major; 0 on a degenerate element */
int orc_solid_stiffness(int kind, const double* xyz, double young, double poisson, double* ke) {
  Point pts[8];
  int np, nn;
  if (kind == 0) {
    np = hex2_points(pts);
    nn = 8;
  } else {
    np = tet4_points(pts);
    nn = 10;
  }
  double d[6][6];
  isotropic(young, poisson, d);
  memset(ke, 0, sizeof(double) * 9 * nn * nn);
  for (int p = 0; p < np; ++p) {
    double n[10], dn[10][3], g[10][3], det;
    if (kind == 0)
      hex8_shape(pts[p].xi, n, dn);
    else
      tet10_shape(pts[p].xi, n, dn);
    if (!gradients(nn, (const double(*)[3])xyz, n, (const double(*)[3])dn, g, &det)) return 0;
    accumulate_symmetric(ke, nn, (const double(*)[3])g, (const double(*)[6])d, det * pts[p].w);
  }
  return 1;
}

/* element.rs:427-441: the consistent mass per component, n x n row-major (tet10 only) */
int orc_tet10_mass(const double* xyz, double density, double* me) {
  Point pts[64];
  int np = tet_conical4_points(pts);
  memset(me, 0, sizeof(double) * 100);
  for (int p = 0; p < np; ++p) {
    double n[10], dn[10][3], det;
    tet10_shape(pts[p].xi, n, dn);
    if (!gradients(10, (const double(*)[3])xyz, n, (const double(*)[3])dn, NULL, &det)) return 0;
    double w = density * det * pts[p].w;
    for (int i = 0; i < 10; ++i)
      for (int j = 0; j < 10; ++j) me[i * 10 + j] += w * n[i] * n[j];
  }
  return 1;
}

/*
 * The upper triangle by columns of K_ff (what = 0), M_ff (1) or K_ff - shift*M_ff (2) of a mesh
 * of one element kind (0 = hex8, 1 = tet10; the mass needs tet10 and density > 0), into
 * out_val aligned with the caller's pattern (uptr, uidx over neq equations).
 *   xyz[nodes][3], conn[elements][npe] (the model's node order), young[n_young] (one per
 *   element, or one for all), eq_dof[neq] (equation -> mesh dof 3*node + direction).
 * Returns -1 on success, the first degenerate element's index (element order) when one is
 * found, -2 when the pattern differs from the assembled one, -3 on an argument error.
 */
i64 orc_assemble_upper(int kind, i64 nodes, const double* xyz, i64 elements, const i64* conn, const double* young,
                       i64 n_young, double poisson, double density, double shift, int what, i64 neq,
                       const i64* eq_dof, const i64* uptr, const i64* uidx, double* out_val) {
  const int npe = kind == 0 ? 8 : 10, W = 3 * npe;
  const int mass = what != 0;
  if ((kind != 0 && kind != 1) || (mass && (kind != 1 || !(density > 0.0))) || n_young < 1) return -3;
  /* node graph (sparse.rs:104-120): each node's sorted neighbours, itself included */
  i64* deg = calloc((size_t)nodes + 1, sizeof(i64));
  for (i64 e = 0; e < elements; ++e)
    for (int a = 0; a < npe; ++a) deg[conn[e * npe + a] + 1] += npe;
  for (i64 v = 0; v < nodes; ++v) deg[v + 1] += deg[v];
  i64* nb = malloc(sizeof(i64) * (size_t)(deg[nodes] ? deg[nodes] : 1));
  i64* fillp = malloc(sizeof(i64) * (size_t)(nodes + 1));
  memcpy(fillp, deg, sizeof(i64) * (size_t)(nodes + 1));
  for (i64 e = 0; e < elements; ++e)
    for (int a = 0; a < npe; ++a)
      for (int b = 0; b < npe; ++b) nb[fillp[conn[e * npe + a]]++] = conn[e * npe + b];
  i64* gptr = calloc((size_t)nodes + 1, sizeof(i64));
  for (i64 v = 0; v < nodes; ++v) {
    i64 s = deg[v], t = deg[v + 1];
    /* sort and unique (insertion sort: a node has at most a few hundred entries) */
    for (i64 i = s + 1; i < t; ++i) {
      i64 x = nb[i], k = i - 1;
      while (k >= s && nb[k] > x) {
        nb[k + 1] = nb[k];
        --k;
      }
      nb[k + 1] = x;
    }
    i64 u = s;
    for (i64 i = s; i < t; ++i)
      if (i == s || nb[i] != nb[u - 1]) nb[u++] = nb[i];
    gptr[v + 1] = u - s;
  }
  for (i64 v = 0; v < nodes; ++v) gptr[v + 1] += gptr[v];
  i64* gidx = malloc(sizeof(i64) * (size_t)(gptr[nodes] ? gptr[nodes] : 1));
  for (i64 v = 0; v < nodes; ++v) memcpy(gidx + gptr[v], nb + deg[v], sizeof(i64) * (size_t)(gptr[v + 1] - gptr[v]));
  free(nb);
  free(fillp);
  free(deg);
  /* K (and M) over the mesh dofs by full rows, zero (NodeGraph::zero_matrix): dof row 3v + r
     holds columns 3u + c for the neighbours u of v, ascending */
  const i64 rows = 3 * nodes;
  i64* rptr = malloc(sizeof(i64) * (size_t)(rows + 1));
  rptr[0] = 0;
  for (i64 v = 0; v < nodes; ++v)
    for (int r = 0; r < 3; ++r) rptr[3 * v + r + 1] = rptr[3 * v + r] + 3 * (gptr[v + 1] - gptr[v]);
  const i64 total = rptr[rows];
  double* kv = calloc((size_t)(total ? total : 1), sizeof(double));
  double* mv = mass ? calloc((size_t)(total ? total : 1), sizeof(double)) : NULL;
  /* element matrices computed in parallel rounds, added in element order (analysis.rs:203-233) */
  const i64 ROUND = 4096;
  double* ke = malloc(sizeof(double) * (size_t)(ROUND * W * W));
  double* me = mass ? malloc(sizeof(double) * (size_t)(ROUND * 100)) : NULL;
  int* ok = malloc(sizeof(int) * (size_t)ROUND);
  i64 bad = -1;
  for (i64 e0 = 0; e0 < elements && bad < 0; e0 += ROUND) {
    i64 cnt = elements - e0 < ROUND ? elements - e0 : ROUND;
#pragma omp parallel for schedule(dynamic, 16)
    for (i64 q = 0; q < cnt; ++q) {
      i64 e = e0 + q;
      double x[10][3];
      for (int a = 0; a < npe; ++a)
        for (int c = 0; c < 3; ++c) x[a][c] = xyz[conn[e * npe + a] * 3 + c];
      double ye = young[n_young == 1 ? 0 : e];
      int good = orc_solid_stiffness(kind, &x[0][0], ye, poisson, ke + q * W * W);
      if (good && mass) good = orc_tet10_mass(&x[0][0], density, me + q * 100);
      ok[q] = good;
    }
    for (i64 q = 0; q < cnt && bad < 0; ++q) {
      i64 e = e0 + q;
      if (!ok[q]) {
        bad = e;
        break;
      }
      const i64* en = conn + e * npe;
      for (int i = 0; i < npe; ++i)
        for (int j = 0; j < npe; ++j) {
          /* Block::new: the position of node en[j] among node en[i]'s neighbours */
          i64 a = en[i], b = en[j], lo = gptr[a], hi = gptr[a + 1];
          while (lo < hi) {
            i64 mid = (lo + hi) / 2;
            if (gidx[mid] < b)
              lo = mid + 1;
            else
              hi = mid;
          }
          i64 slot = 3 * (lo - gptr[a]);
          for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) kv[rptr[3 * a + r] + slot + c] += ke[q * W * W + (3 * i + r) * W + 3 * j + c];
          if (mass)
            for (int c = 0; c < 3; ++c) mv[rptr[3 * a + c] + slot + c] += me[q * 100 + i * 10 + j];
        }
    }
  }
  free(ke);
  free(me);
  free(ok);
  i64 result = bad;
  if (bad < 0) {
    /* reduce (constraints.rs:865-897) then from_symmetric (cholesky.rs:110-144), per equation
       row k: the entries of dof row eq_dof[k] whose column is a free equation f <= k, sorted by
       f, each 0.0 + (1.0 * v) * 1.0 */
    i64* eq = malloc(sizeof(i64) * (size_t)(rows ? rows : 1));
    for (i64 d = 0; d < rows; ++d) eq[d] = -1;
    for (i64 k = 0; k < neq; ++k) eq[eq_dof[k]] = k;
    int mismatch = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : mismatch)
    for (i64 k = 0; k < neq; ++k) {
      i64 d = eq_dof[k];
      i64 at = uptr[k];
      /* the columns of a dof row ascend with the node, and equations follow the node order
         with x, y, z within a node; sort the kept entries by equation */
      i64 s = rptr[d], t = rptr[d + 1];
      i64 buf_f[1024];
      double buf_k[1024], buf_m[1024];
      i64 cntk = 0;
      for (i64 p = s; p < t; ++p) {
        i64 nbr = gidx[gptr[d / 3] + (p - s) / 3];
        i64 f = eq[3 * nbr + (p - s) % 3];
        if (f < 0 || f > k) continue;
        if (cntk >= 1024) {
          mismatch = 1;
          break;
        }
        buf_f[cntk] = f;
        buf_k[cntk] = 0.0 + (1.0 * kv[p]) * 1.0;
        buf_m[cntk] = mass ? 0.0 + (1.0 * mv[p]) * 1.0 : 0.0;
        ++cntk;
      }
      for (i64 i = 1; i < cntk; ++i) {
        i64 f = buf_f[i], q = i - 1;
        double vk = buf_k[i], vm = buf_m[i];
        while (q >= 0 && buf_f[q] > f) {
          buf_f[q + 1] = buf_f[q];
          buf_k[q + 1] = buf_k[q];
          buf_m[q + 1] = buf_m[q];
          --q;
        }
        buf_f[q + 1] = f;
        buf_k[q + 1] = vk;
        buf_m[q + 1] = vm;
      }
      if (uptr[k + 1] - at != cntk) {
        mismatch = 1;
        continue;
      }
      for (i64 i = 0; i < cntk; ++i) {
        if (uidx[at + i] != buf_f[i]) mismatch = 1;
        double v = what == 0 ? buf_k[i] : what == 1 ? buf_m[i] : buf_k[i] + (-shift) * buf_m[i];
        out_val[at + i] = v;
      }
    }
    free(eq);
    if (mismatch) result = -2;
  }
  free(kv);
  free(mv);
  free(rptr);
  free(gptr);
  free(gidx);
  return result;
}
