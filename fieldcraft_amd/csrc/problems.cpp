// Synthetic inputs for the supernodal Cholesky: structured solid meshes, their element
// matrices, the assembled stiffness (and consistent mass), the fill-reducing node order and
// the free-equation matrix K_ff that the reference's static analysis factors.
//
// This is the input side of the factorization, in the reference's own order of work:
//   mesh nodes in label order, node graph (sparse.rs:104-120), fill-reducing order
//   (ordering.rs:102-110), element matrices added in element label order, row-major
//   (analysis.rs:199-233, sparse.rs:207-224), free equations numbered node by node in the
//   order, x, y, z (constraints.rs:755-766), and K_ff = T^T K T for a plain selection, each row
//   sorted, starting each sum from 0.0 (constraints.rs:820-906).
// Element kernels are written from the mathematics: trilinear hex8 with 2x2x2 Gauss points
// (the reference's default full integration, fem element.rs:123-133), quadratic tet10 with the
// 4-point rule for stiffness (element.rs:90-99) and an exact consistent mass from the
// barycentric monomial integral  int L1^a L2^b L3^c L4^d dV = 6V a!b!c!d!/(a+b+c+d+3)!.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "ordering.hpp"

namespace {

using i64 = int64_t;

struct Problem {
  i64 nodes = 0;
  std::vector<double> xyz;             // 3 per node
  std::vector<i64> conn;               // npe per element
  int npe = 0;
  std::vector<i64> adj_ptr, adj_idx;   // node graph, self included, sorted
  std::vector<i64> order;              // fill-reducing node order
  int graph_chosen = 0;
  std::vector<uint8_t> fixed;          // per mesh dof
  std::vector<i64> eq, eq_dof;         // mesh dof -> equation, equation -> mesh dof
  // K_ff (and M_ff) by full rows in equation order
  std::vector<i64> ptr, idx;
  std::vector<double> kval, mval;
};

// ------------------------------------------------------------------------------------------
// Element matrices

struct Iso {
  double lambda, mu;
};
Iso iso(double young, double poisson) {
  // fem material.rs:40-54
  return {young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson)), young / (2.0 * (1.0 + poisson))};
}

// Inverse and determinant of a 3x3 matrix.
bool inv3(const double j[3][3], double inv[3][3], double* det) {
  double d = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) - j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
             j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
  if (!(d > 0.0)) return false;
  inv[0][0] = (j[1][1] * j[2][2] - j[1][2] * j[2][1]) / d;
  inv[0][1] = (j[0][2] * j[2][1] - j[0][1] * j[2][2]) / d;
  inv[0][2] = (j[0][1] * j[1][2] - j[0][2] * j[1][1]) / d;
  inv[1][0] = (j[1][2] * j[2][0] - j[1][0] * j[2][2]) / d;
  inv[1][1] = (j[0][0] * j[2][2] - j[0][2] * j[2][0]) / d;
  inv[1][2] = (j[0][2] * j[1][0] - j[0][0] * j[1][2]) / d;
  inv[2][0] = (j[1][0] * j[2][1] - j[1][1] * j[2][0]) / d;
  inv[2][1] = (j[0][1] * j[2][0] - j[0][0] * j[2][1]) / d;
  inv[2][2] = (j[0][0] * j[1][1] - j[0][1] * j[1][0]) / d;
  *det = d;
  return true;
}

// Adds w * B^T D B for isotropic D, given the physical gradients g[a][3] of nn shape functions:
// block (a, b) is lambda g_a g_b^T + mu g_b g_a^T + mu (g_a . g_b) I.
void add_isotropic(int nn, const double (*g)[3], Iso m, double w, double* ke) {
  int width = 3 * nn;
  for (int a = 0; a < nn; ++a)
    for (int b = 0; b < nn; ++b) {
      double dot = g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2];
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
          double v = m.lambda * g[a][r] * g[b][c] + m.mu * g[a][c] * g[b][r];
          if (r == c) v += m.mu * dot;
          ke[(3 * a + r) * width + 3 * b + c] += w * v;
        }
    }
}

const double HEX_CORNERS[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                  {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// Hex8 stiffness, 2x2x2 Gauss points (24 x 24, node-major, x y z per node).
bool hex8_stiffness(const double x[8][3], Iso m, double* ke) {
  std::fill(ke, ke + 24 * 24, 0.0);
  const double q = 1.0 / std::sqrt(3.0);
  for (int gz = 0; gz < 2; ++gz)
    for (int gy = 0; gy < 2; ++gy)
      for (int gx = 0; gx < 2; ++gx) {
        double xi[3] = {gx ? q : -q, gy ? q : -q, gz ? q : -q};
        double dn[8][3];
        for (int i = 0; i < 8; ++i) {
          double f[3];
          for (int k = 0; k < 3; ++k) f[k] = 1.0 + xi[k] * HEX_CORNERS[i][k];
          dn[i][0] = HEX_CORNERS[i][0] * f[1] * f[2] / 8.0;
          dn[i][1] = HEX_CORNERS[i][1] * f[0] * f[2] / 8.0;
          dn[i][2] = HEX_CORNERS[i][2] * f[0] * f[1] / 8.0;
        }
        double j[3][3] = {{0}}, inv[3][3], det;
        for (int i = 0; i < 8; ++i)
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) j[a][b] += dn[i][a] * x[i][b];
        if (!inv3(j, inv, &det)) return false;
        double g[8][3];
        for (int i = 0; i < 8; ++i)
          for (int b = 0; b < 3; ++b) g[i][b] = inv[b][0] * dn[i][0] + inv[b][1] * dn[i][1] + inv[b][2] * dn[i][2];
        add_isotropic(8, g, m, det, ke);
      }
  return true;
}

// Tet10 edges as (corner, corner, mid-side node), model mesh.rs:176-186.
const int TET10_EDGES[6][3] = {{0, 1, 4}, {1, 2, 5}, {2, 0, 6}, {0, 3, 7}, {1, 3, 8}, {2, 3, 9}};

// Tet10 shape derivatives with respect to the barycentric-based natural coordinates
// (xi = L1, L2, L3; L0 = 1 - xi sum).
void tet10_dn(const double l[4], double dn[10][3]) {
  const double dl[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  for (int i = 0; i < 4; ++i)
    for (int c = 0; c < 3; ++c) dn[i][c] = (4.0 * l[i] - 1.0) * dl[i][c];
  for (auto& e : TET10_EDGES)
    for (int c = 0; c < 3; ++c) dn[e[2]][c] = 4.0 * (dl[e[0]][c] * l[e[1]] + l[e[0]] * dl[e[1]][c]);
}

// Tet10 stiffness with the 4-point rule (30 x 30).
bool tet10_stiffness(const double x[10][3], Iso m, double* ke) {
  std::fill(ke, ke + 30 * 30, 0.0);
  const double a = 0.5854101966249685, b = 0.1381966011250105;
  for (int p = 0; p < 4; ++p) {
    double l[4] = {b, b, b, b};
    l[p] = a;
    double dn[10][3];
    tet10_dn(l, dn);
    double j[3][3] = {{0}}, inv[3][3], det;
    for (int i = 0; i < 10; ++i)
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) j[r][c] += dn[i][r] * x[i][c];
    if (!inv3(j, inv, &det)) return false;
    double g[10][3];
    for (int i = 0; i < 10; ++i)
      for (int c = 0; c < 3; ++c) g[i][c] = inv[c][0] * dn[i][0] + inv[c][1] * dn[i][1] + inv[c][2] * dn[i][2];
    add_isotropic(10, g, m, det / 24.0, ke);  // reference tet volume 1/6, four points of weight 1/4
  }
  return true;
}

// Exact consistent mass of an affine tet10 (10 x 10 scalar matrix; the same on each direction).
// Shape functions as barycentric polynomials: corner i: 2 Li^2 - Li, mid (a,b): 4 La Lb.
void tet10_mass(const double x[10][3], double rho, double* me) {
  double e1[3], e2[3], e3[3];
  for (int c = 0; c < 3; ++c) {
    e1[c] = x[1][c] - x[0][c];
    e2[c] = x[2][c] - x[0][c];
    e3[c] = x[3][c] - x[0][c];
  }
  double vol = (e1[0] * (e2[1] * e3[2] - e2[2] * e3[1]) - e1[1] * (e2[0] * e3[2] - e2[2] * e3[0]) +
                e1[2] * (e2[0] * e3[1] - e2[1] * e3[0])) / 6.0;
  // monomials: exponent vectors with coefficients
  struct Term { int e[4]; double c; };
  std::vector<std::vector<Term>> shape(10);
  for (int i = 0; i < 4; ++i) {
    Term t2{{0, 0, 0, 0}, 2.0}, t1{{0, 0, 0, 0}, -1.0};
    t2.e[i] = 2;
    t1.e[i] = 1;
    shape[i] = {t2, t1};
  }
  for (auto& ed : TET10_EDGES) {
    Term t{{0, 0, 0, 0}, 4.0};
    t.e[ed[0]] += 1;
    t.e[ed[1]] += 1;
    shape[ed[2]] = {t};
  }
  auto fact = [](int k) { double f = 1; for (int i = 2; i <= k; ++i) f *= i; return f; };
  for (int i = 0; i < 10; ++i)
    for (int j = 0; j < 10; ++j) {
      double s = 0.0;
      for (auto& a : shape[i])
        for (auto& b : shape[j]) {
          int e[4], tot = 0;
          double num = 1.0;
          for (int k = 0; k < 4; ++k) { e[k] = a.e[k] + b.e[k]; tot += e[k]; num *= fact(e[k]); }
          s += a.c * b.c * 6.0 * num / fact(tot + 3);
        }
      me[i * 10 + j] = rho * vol * s;
    }
}

// ------------------------------------------------------------------------------------------
// Assembly

void node_graph(Problem& p) {
  // NodeGraph::new (sparse.rs:104-120): every element couples all its nodes, self included
  std::vector<std::vector<i64>> rows(p.nodes);
  i64 ne = (i64)p.conn.size() / p.npe;
  for (i64 e = 0; e < ne; ++e) {
    const i64* c = &p.conn[e * p.npe];
    for (int a = 0; a < p.npe; ++a) rows[c[a]].insert(rows[c[a]].end(), c, c + p.npe);
  }
  p.adj_ptr.assign(p.nodes + 1, 0);
  for (i64 v = 0; v < p.nodes; ++v) {
    auto& r = rows[v];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    p.adj_ptr[v + 1] = p.adj_ptr[v] + (i64)r.size();
  }
  p.adj_idx.resize(p.adj_ptr[p.nodes]);
  for (i64 v = 0; v < p.nodes; ++v) std::copy(rows[v].begin(), rows[v].end(), p.adj_idx.begin() + p.adj_ptr[v]);
}

// Full K (and M) over the mesh dofs by rows (3 dofs a node), values added in element order,
// then reduced to the free equations.
struct Mesh3 {
  std::vector<i64> rptr;   // per mesh dof row start
  std::vector<double> k, m;
};

i64 entry(const Problem& p, const Mesh3& g, i64 a, int r, i64 b, int c) {
  const i64* s = &p.adj_idx[p.adj_ptr[a]];
  const i64* e = &p.adj_idx[p.adj_ptr[a + 1]];
  i64 pos = std::lower_bound(s, e, b) - s;
  return g.rptr[3 * a + r] + 3 * pos + c;
}

void reduce(Problem& p, const Mesh3& g, bool mass) {
  // equation numbering (constraints.rs:755-766)
  i64 ndof = 3 * p.nodes;
  p.eq.assign(ndof, -1);
  p.eq_dof.clear();
  for (i64 v : p.order)
    for (int a = 0; a < 3; ++a) {
      i64 d = 3 * v + a;
      if (!p.fixed[d]) { p.eq[d] = (i64)p.eq_dof.size(); p.eq_dof.push_back(d); }
    }
  i64 n = (i64)p.eq_dof.size();
  p.ptr.assign(n + 1, 0);
  for (i64 e = 0; e < n; ++e) {
    i64 d = p.eq_dof[e], v = d / 3;
    i64 cnt = 0;
    for (i64 q = p.adj_ptr[v]; q < p.adj_ptr[v + 1]; ++q)
      for (int b = 0; b < 3; ++b) cnt += !p.fixed[3 * p.adj_idx[q] + b];
    p.ptr[e + 1] = p.ptr[e] + cnt;
  }
  p.idx.resize(p.ptr[n]);
  p.kval.resize(p.ptr[n]);
  if (mass) p.mval.resize(p.ptr[n]);
  std::vector<std::pair<i64, i64>> row;
  for (i64 e = 0; e < n; ++e) {
    i64 d = p.eq_dof[e], v = d / 3;
    int a = (int)(d % 3);
    row.clear();
    i64 at = g.rptr[d];
    for (i64 q = p.adj_ptr[v]; q < p.adj_ptr[v + 1]; ++q)
      for (int b = 0; b < 3; ++b, ++at) {
        i64 col = 3 * p.adj_idx[q] + b;
        if (!p.fixed[col]) row.emplace_back(p.eq[col], at);
      }
    (void)a;
    std::sort(row.begin(), row.end());
    i64 o = p.ptr[e];
    for (auto& rc : row) {
      p.idx[o] = rc.first;
      p.kval[o] = 0.0 + g.k[rc.second];
      if (mass) p.mval[o] = 0.0 + g.m[rc.second];
      ++o;
    }
  }
}

Mesh3 zero_matrix(const Problem& p, bool mass) {
  Mesh3 g;
  g.rptr.assign(3 * p.nodes + 1, 0);
  for (i64 v = 0; v < p.nodes; ++v) {
    i64 deg = p.adj_ptr[v + 1] - p.adj_ptr[v];
    for (int a = 0; a < 3; ++a) g.rptr[3 * v + a + 1] = g.rptr[3 * v + a] + 3 * deg;
  }
  g.k.assign(g.rptr.back(), 0.0);
  if (mass) g.m.assign(g.rptr.back(), 0.0);
  return g;
}

// add_block (sparse.rs:207-224)
void add_block(const Problem& p, Mesh3& g, const i64* nodes, int nn, const double* ke) {
  int width = 3 * nn;
  for (int i = 0; i < nn; ++i)
    for (int j = 0; j < nn; ++j)
      for (int r = 0; r < 3; ++r) {
        i64 base = entry(p, g, nodes[i], r, nodes[j], 0);
        for (int c = 0; c < 3; ++c) g.k[base + c] += ke[(3 * i + r) * width + 3 * j + c];
      }
}

// add_scalar (sparse.rs:228-243): an nn x nn matrix on each direction
void add_scalar(const Problem& p, Mesh3& g, const i64* nodes, int nn, const double* me) {
  for (int i = 0; i < nn; ++i)
    for (int j = 0; j < nn; ++j)
      for (int c = 0; c < 3; ++c) g.m[entry(p, g, nodes[i], c, nodes[j], c)] += me[i * nn + j];
}

}  // namespace

// ------------------------------------------------------------------------------------------
// C interface

extern "C" {

struct fcp_problem;

// A structured block of nx x ny x nz cells on [0, lx] x [0, ly] x [0, lz].
//   kind        0 = hex8, 1 = tet10 (six Kuhn tetrahedra a cell, mid-side nodes on the doubled
//               lattice, as the reference's test grids, solver tests/common/mod.rs:23-98)
//   young       per element (n_young == number of elements) or one value (n_young == 1)
//   clamp_face  0..5 = x=0, x=L, y=0, y=L, z=0, z=L: every dof of its nodes is fixed; -1: none
//   jitter      fraction of the cell size by which corner nodes off the clamped face move
//               (uniform in +-jitter per direction, seeded); mid-side nodes stay at the
//               midpoints of their corners, so elements stay straight-sided
//   rho         > 0 also assembles the consistent mass (tet10 only)
// Returns NULL on a degenerate element.
fcp_problem* fcp_block(int kind, i64 nx, i64 ny, i64 nz, double lx, double ly, double lz, const double* young,
                       i64 n_young, double poisson, int clamp_face, double jitter, uint64_t seed, double rho) {
  Problem* p = new Problem();
  const bool quad = kind == 1;
  const i64 sx = quad ? 2 * nx + 1 : nx + 1, sy = quad ? 2 * ny + 1 : ny + 1, sz = quad ? 2 * nz + 1 : nz + 1;
  const i64 dx = quad ? 2 * nx : nx, dy = quad ? 2 * ny : ny, dz = quad ? 2 * nz : nz;
  p->nodes = sx * sy * sz;
  p->xyz.resize(3 * p->nodes);
  auto id = [&](i64 i, i64 j, i64 k) { return i + sx * (j + sy * k); };
  for (i64 k = 0; k < sz; ++k)
    for (i64 j = 0; j < sy; ++j)
      for (i64 i = 0; i < sx; ++i) {
        double* x = &p->xyz[3 * id(i, j, k)];
        x[0] = (double)i * lx / (double)dx;
        x[1] = (double)j * ly / (double)dy;
        x[2] = (double)k * lz / (double)dz;
      }
  auto on_face = [&](i64 i, i64 j, i64 k, int face) {
    switch (face) {
      case -1: return false;  // no supports
      case 0: return i == 0;
      case 1: return i == sx - 1;
      case 2: return j == 0;
      case 3: return j == sy - 1;
      case 4: return k == 0;
      default: return k == sz - 1;
    }
  };
  if (jitter > 0.0) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-jitter, jitter);
    double h[3] = {lx / nx, ly / ny, lz / nz};
    int step = quad ? 2 : 1;
    for (i64 k = 0; k < sz; k += step)
      for (i64 j = 0; j < sy; j += step)
        for (i64 i = 0; i < sx; i += step) {
          double* x = &p->xyz[3 * id(i, j, k)];
          double d[3] = {u(rng), u(rng), u(rng)};
          if (on_face(i, j, k, clamp_face)) continue;
          for (int c = 0; c < 3; ++c) x[c] += d[c] * h[c];
        }
    if (quad)  // mid-side nodes at the midpoints of their corners
      for (i64 k = 0; k < sz; ++k)
        for (i64 j = 0; j < sy; ++j)
          for (i64 i = 0; i < sx; ++i) {
            if (i % 2 == 0 && j % 2 == 0 && k % 2 == 0) continue;
            // the lattice point is the midpoint of the Kuhn edge through it: its odd
            // coordinates come from both neighbours along the same diagonal
            i64 i0 = i - (i % 2), j0 = j - (j % 2), k0 = k - (k % 2);
            i64 i1 = i + (i % 2), j1 = j + (j % 2), k1 = k + (k % 2);
            const double* a = &p->xyz[3 * id(i0, j0, k0)];
            const double* b = &p->xyz[3 * id(i1, j1, k1)];
            double* x = &p->xyz[3 * id(i, j, k)];
            for (int c = 0; c < 3; ++c) x[c] = 0.5 * (a[c] + b[c]);
          }
  }
  // elements in label order: cells z-major, then y, then x
  if (!quad) {
    p->npe = 8;
    const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
    p->conn.reserve(8 * nx * ny * nz);
    for (i64 k = 0; k < nz; ++k)
      for (i64 j = 0; j < ny; ++j)
        for (i64 i = 0; i < nx; ++i)
          for (auto& c : corner) p->conn.push_back(id(i + c[0], j + c[1], k + c[2]));
  } else {
    p->npe = 10;
    const int perms[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    p->conn.reserve(60 * nx * ny * nz);
    for (i64 cz = 0; cz < nz; ++cz)
      for (i64 cy = 0; cy < ny; ++cy)
        for (i64 cx = 0; cx < nx; ++cx)
          for (auto& pm : perms) {
            int v[4][3] = {{0, 0, 0}};
            for (int s = 0; s < 3; ++s) {
              for (int c = 0; c < 3; ++c) v[s + 1][c] = v[s][c];
              v[s + 1][pm[s]] = 2;
            }
            double e[3][3];
            for (int q = 0; q < 3; ++q)
              for (int c = 0; c < 3; ++c) e[q][c] = (double)(v[q + 1][c] - v[0][c]);
            double vol = e[0][0] * (e[1][1] * e[2][2] - e[1][2] * e[2][1]) - e[0][1] * (e[1][0] * e[2][2] - e[1][2] * e[2][0]) +
                         e[0][2] * (e[1][0] * e[2][1] - e[1][1] * e[2][0]);
            if (vol < 0)
              for (int c = 0; c < 3; ++c) std::swap(v[1][c], v[2][c]);
            i64 base[3] = {2 * cx, 2 * cy, 2 * cz};
            for (int q = 0; q < 4; ++q) p->conn.push_back(id(base[0] + v[q][0], base[1] + v[q][1], base[2] + v[q][2]));
            for (auto& ed : TET10_EDGES)
              p->conn.push_back(id(base[0] + (v[ed[0]][0] + v[ed[1]][0]) / 2, base[1] + (v[ed[0]][1] + v[ed[1]][1]) / 2,
                                   base[2] + (v[ed[0]][2] + v[ed[1]][2]) / 2));
          }
  }
  node_graph(*p);
  std::vector<fc_order::i64> order;
  p->graph_chosen = fc_order::fill_reducing_order(p->nodes, p->xyz.data(), p->adj_ptr.data(), p->adj_idx.data(), order);
  p->order.assign(order.begin(), order.end());
  p->fixed.assign(3 * p->nodes, 0);
  for (i64 k = 0; k < sz; ++k)
    for (i64 j = 0; j < sy; ++j)
      for (i64 i = 0; i < sx; ++i)
        if (on_face(i, j, k, clamp_face))
          for (int a = 0; a < 3; ++a) p->fixed[3 * id(i, j, k) + a] = 1;
  const bool mass = rho > 0.0 && quad;
  Mesh3 g = zero_matrix(*p, mass);
  i64 ne = (i64)p->conn.size() / p->npe;
  std::vector<double> ke(30 * 30), me(100);
  double cached_e = std::nan("");
  for (i64 e = 0; e < ne; ++e) {
    const i64* c = &p->conn[e * p->npe];
    double ye = n_young == 1 ? young[0] : young[e];
    Iso m = iso(ye, poisson);
    if (!quad) {
      // every hex of a block has the same shape: the matrix depends on E alone
      if (!(ye == cached_e) || jitter > 0.0) {
        double x[8][3];
        for (int a = 0; a < 8; ++a)
          for (int d = 0; d < 3; ++d) x[a][d] = p->xyz[3 * c[a] + d];
        if (!hex8_stiffness(x, m, ke.data())) { delete p; return nullptr; }
        cached_e = ye;
      }
      add_block(*p, g, c, 8, ke.data());
    } else {
      double x[10][3];
      for (int a = 0; a < 10; ++a)
        for (int d = 0; d < 3; ++d) x[a][d] = p->xyz[3 * c[a] + d];
      if (!tet10_stiffness(x, m, ke.data())) { delete p; return nullptr; }
      add_block(*p, g, c, 10, ke.data());
      if (mass) {
        tet10_mass(x, rho, me.data());
        add_scalar(*p, g, c, 10, me.data());
      }
    }
  }
  reduce(*p, g, mass);
  return reinterpret_cast<fcp_problem*>(p);
}

void fcp_free(fcp_problem* h) { delete reinterpret_cast<Problem*>(h); }

i64 fcp_nodes(const fcp_problem* h) { return reinterpret_cast<const Problem*>(h)->nodes; }
i64 fcp_elements(const fcp_problem* h) {
  auto p = reinterpret_cast<const Problem*>(h);
  return (i64)p->conn.size() / p->npe;
}
int fcp_nodes_per_element(const fcp_problem* h) { return reinterpret_cast<const Problem*>(h)->npe; }
i64 fcp_neq(const fcp_problem* h) { return (i64)reinterpret_cast<const Problem*>(h)->eq_dof.size(); }
i64 fcp_nnz(const fcp_problem* h) { return (i64)reinterpret_cast<const Problem*>(h)->idx.size(); }
i64 fcp_adjacency_nnz(const fcp_problem* h) { return (i64)reinterpret_cast<const Problem*>(h)->adj_idx.size(); }
int fcp_graph_chosen(const fcp_problem* h) { return reinterpret_cast<const Problem*>(h)->graph_chosen; }
int fcp_has_mass(const fcp_problem* h) { return !reinterpret_cast<const Problem*>(h)->mval.empty(); }

void fcp_get_mesh(const fcp_problem* h, double* xyz, i64* conn, i64* adj_ptr, i64* adj_idx, i64* order, i64* eq_dof) {
  auto p = reinterpret_cast<const Problem*>(h);
  if (xyz) std::copy(p->xyz.begin(), p->xyz.end(), xyz);
  if (conn) std::copy(p->conn.begin(), p->conn.end(), conn);
  if (adj_ptr) std::copy(p->adj_ptr.begin(), p->adj_ptr.end(), adj_ptr);
  if (adj_idx) std::copy(p->adj_idx.begin(), p->adj_idx.end(), adj_idx);
  if (order) std::copy(p->order.begin(), p->order.end(), order);
  if (eq_dof) std::copy(p->eq_dof.begin(), p->eq_dof.end(), eq_dof);
}

// K_ff (and M_ff) by full rows (both triangles), equation order.
void fcp_get_csr(const fcp_problem* h, i64* ptr, i64* idx, double* kval, double* mval) {
  auto p = reinterpret_cast<const Problem*>(h);
  if (ptr) std::copy(p->ptr.begin(), p->ptr.end(), ptr);
  if (idx) std::copy(p->idx.begin(), p->idx.end(), idx);
  if (kval) std::copy(p->kval.begin(), p->kval.end(), kval);
  if (mval && !p->mval.empty()) std::copy(p->mval.begin(), p->mval.end(), mval);
}

// The fill-reducing order of any node graph (ordering.rs:102-110); returns 1 when the graph
// dissection was kept.
int fcp_fill_reducing_order(i64 n, const double* xyz, const i64* adj_ptr, const i64* adj_idx, i64* order) {
  std::vector<fc_order::i64> o;
  int g = fc_order::fill_reducing_order(n, xyz, adj_ptr, adj_idx, o);
  std::copy(o.begin(), o.end(), order);
  return g;
}
void fcp_nested_dissection(i64 n, const double* xyz, const i64* adj_ptr, const i64* adj_idx, i64* order) {
  auto o = fc_order::nested_dissection(n, xyz, adj_ptr, adj_idx);
  std::copy(o.begin(), o.end(), order);
}
void fcp_graph_dissection(i64 n, const i64* adj_ptr, const i64* adj_idx, i64* order) {
  auto o = fc_order::graph_dissection(n, adj_ptr, adj_idx);
  std::copy(o.begin(), o.end(), order);
}
i64 fcp_fill(i64 n, const i64* adj_ptr, const i64* adj_idx, const i64* order) {
  return fc_order::fill(n, adj_ptr, adj_idx, order);
}

}  // extern "C"
