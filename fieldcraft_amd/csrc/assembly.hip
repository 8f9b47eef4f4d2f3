// Assembly of the factor's input on the GPU: the solid element matrices of a mesh, added in
// element order into the upper triangle by columns of K_ff, M_ff or K_ff - shift * M_ff, in
// device memory, where fc_refactor_device reads it.
//
// Reference path replaced (FieldCraft 0.8.1):
//   crates/solver/src/analysis.rs:199-233   assemble: element matrices added in element order
//   crates/solver/src/sparse.rs:199-243     add_block / add_scalar onto the node graph
//   crates/solver/src/constraints.rs:820-906 reduce: K_ff = T^T K T (a plain selection here)
//   crates/solver/src/sparse.rs:81-87       add_scaled: K + (-shift) M
//   crates/solver/src/cholesky.rs:110-144   UpperColumns::from_symmetric
// and the element kernels of crates/fem: element.rs:301-324 (Jacobian, gradients), 416-441
// (stiffness, consistent mass), 723-774 (b_block, accumulate_symmetric), shape.rs:60-104,
// quadrature.rs:27-118, material.rs:40-54, dense.rs:91-116 and 169-172, model mesh.rs:254-269.
//
// Every value is the reference's bit for bit: the element matrices follow the reference's
// operations in the same order (this file is compiled without floating-point contraction),
// and each entry of K_ff sums its elements' contributions in element order from 0.0.
//
// Design (MI355X):
//  - A gather plan, built once on the host from the pattern: for each upper entry, the
//    positions of its contributions in the element-matrix buffer, in element order (no atomics,
//    so the sums are deterministic and the reference's).
//  - Element kernels: one workgroup takes several elements; the (element, integration point)
//    threads form the Jacobians and physical gradients into LDS, then one thread per 3 x 3 node
//    block (blocks j >= i, as the reference forms them) accumulates over the points. The
//    stiffness is stored as its packed upper triangle (the reference's matrix is exactly
//    symmetric: each mirrored entry receives the same values in the same order).
//  - A gather kernel, one thread per entry of the upper triangle, streams the plan and writes
//    the result (K, M or the shifted matrix).
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "fc_chol.h"

namespace {

using i64 = int64_t;

struct DeviceError {
  hipError_t e;
  const char* what;
};
struct ArgError {
  std::string msg;
};

#define ASM_CHECK(x)                                 \
  do {                                               \
    hipError_t e_ = (x);                             \
    if (e_ != hipSuccess) throw DeviceError{e_, #x}; \
  } while (0)

void put_error(fc_error* err, int kind, const std::string& msg, uint64_t column = 0) {
  if (!err) return;
  err->kind = kind;
  err->pad = 0;
  err->column = column;
  err->entries = 0;
  err->limit = 0;
  std::snprintf(err->message, sizeof err->message, "%s", msg.c_str());
}

template <typename T>
T* dalloc(size_t count) {
  if (count == 0) count = 1;
  void* p = nullptr;
  if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    throw std::bad_alloc();
  }
  return (T*)p;
}

template <typename T>
T* dupload(const T* src, size_t count, hipStream_t st) {
  T* p = dalloc<T>(count);
  if (count) ASM_CHECK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, st));
  return p;
}

// ---------------------------------------------------------------------------------------------
// The element kernels' arithmetic, as the reference writes it (device and host alike).

// quadrature.rs:28-33
__constant__ double G2X[2] = {-0.5773502691896258, 0.5773502691896258};
__constant__ double G2W[2] = {1.0, 1.0};
__constant__ double G4X[4] = {-0.8611363115940526, -0.3399810435848563, 0.3399810435848563, 0.8611363115940526};
__constant__ double G4W[4] = {0.3478548451374539, 0.6521451548625461, 0.6521451548625461, 0.3478548451374539};
// shape.rs:50-51: hexahedron corners
__constant__ double HEX_C[8][3] = {{-1., -1., -1.}, {1., -1., -1.}, {1., 1., -1.}, {-1., 1., -1.},
                                   {-1., -1., 1.},  {1., -1., 1.},  {1., 1., 1.},  {-1., 1., 1.}};
// model mesh.rs:185: tet10 edges (corner, corner, mid-side node)
__constant__ int TET10_EDGES[6][3] = {{0, 1, 4}, {1, 2, 5}, {2, 0, 6}, {0, 3, 7}, {1, 3, 8}, {2, 3, 9}};

// shape.rs:95-104 (derivatives only: the stiffness needs no values)
__device__ inline void hex8_dn(const double xi[3], double dn[8][3]) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const double c0 = HEX_C[i][0], c1 = HEX_C[i][1], c2 = HEX_C[i][2];
    const double f0 = 1.0 + xi[0] * c0, f1 = 1.0 + xi[1] * c1, f2 = 1.0 + xi[2] * c2;
    dn[i][0] = c0 * f1 * f2 / 8.0;
    dn[i][1] = c1 * f0 * f2 / 8.0;
    dn[i][2] = c2 * f0 * f1 / 8.0;
  }
}

// shape.rs:60-93
__device__ inline void tet10_shape(const double xi[3], double n[10], double dn[10][3]) {
  const double l[4] = {1.0 - xi[0] - xi[1] - xi[2], xi[0], xi[1], xi[2]};
  const double dl[4][3] = {{-1.0, -1.0, -1.0}, {1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}};
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    n[i] = l[i] * (2.0 * l[i] - 1.0);
#pragma unroll
    for (int c = 0; c < 3; ++c) dn[i][c] = (4.0 * l[i] - 1.0) * dl[i][c];
  }
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const int a = TET10_EDGES[e][0], b = TET10_EDGES[e][1], m = TET10_EDGES[e][2];
    n[m] = 4.0 * l[a] * l[b];
#pragma unroll
    for (int c = 0; c < 3; ++c) dn[m][c] = 4.0 * (dl[a][c] * l[b] + l[a] * dl[b][c]);
  }
}

// Jacobian J[a][b] = sum_i dN_i/dxi_a x_ib (element.rs:301-311), the admissibility gate
// (mesh.rs:257-269), inverse3 and det3 (dense.rs:91-116, 169-172), physical gradients
// (element.rs:313-324). False on an inadmissible or singular Jacobian.
template <int NN>
__device__ inline bool gradients(const double (&x)[NN][3], const double (&dn)[NN][3], double (*g)[3], double* det_out) {
  double j[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
  for (int i = 0; i < NN; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b < 3; ++b) j[a][b] += dn[i][a] * x[i][b];
  {
    const double* a = j[0];
    const double* b = j[1];
    const double* c = j[2];
    const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) +
                       a[2] * (b[0] * c[1] - b[1] * c[0]);
    double norms = 1.0;
#pragma unroll
    for (int r = 0; r < 3; ++r) norms = norms * sqrt(j[r][0] * j[r][0] + j[r][1] * j[r][1] + j[r][2] * j[r][2]);
    double q = det / norms;
    q = isfinite(q) ? fmin(fmax(q, -1.0), 1.0) : 0.0;
    if (!(q > 1e-12)) return false;
  }
  const double det = j[0][0] * (j[1][1] * j[2][2] - j[1][2] * j[2][1]) - j[0][1] * (j[1][0] * j[2][2] - j[1][2] * j[2][0]) +
                     j[0][2] * (j[1][0] * j[2][1] - j[1][1] * j[2][0]);
  if (!(fabs(det) > 0.0 && isfinite(det))) return false;
  double inv[3][3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const int c1 = (c + 1) % 3, c2 = (c + 2) % 3, r1 = (r + 1) % 3, r2 = (r + 2) % 3;
      inv[r][c] = (j[c1][r1] * j[c2][r2] - j[c1][r2] * j[c2][r1]) / det;
    }
  if (g)
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
      for (int b = 0; b < 3; ++b) g[i][b] = inv[b][0] * dn[i][0] + inv[b][1] * dn[i][1] + inv[b][2] * dn[i][2];
  *det_out = det;
  return true;
}

// One 3 x 3 node block (i, j), i <= j, of accumulate_symmetric (element.rs:745-774) at one
// point: acc[r][c] += w * sum_s B_i[s][r] (D B_j)[s][c], upper triangle only on the diagonal.
// The terms whose factor is a structural zero of b_block (element.rs:723-725) or of the
// isotropic D (material.rs:44-52) are left out: each only adds +-0.0 to its sum, which leaves
// a nonzero sum as it is, and a sum that is zero ends as +0.0 in K_ff either way (every entry
// of K starts at +0.0 and +0.0 + -0.0 = +0.0). The terms kept are summed in the reference's
// order, so the assembled values are the same bit for bit.
__device__ constexpr bool B_NZ[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1}};
__device__ constexpr bool D_NZ[6][6] = {{1, 1, 1, 0, 0, 0}, {1, 1, 1, 0, 0, 0}, {1, 1, 1, 0, 0, 0},
                                        {0, 0, 0, 1, 0, 0}, {0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 1}};
__device__ constexpr bool db_nz(int s, int c) {
  bool nz = false;
  for (int t = 0; t < 6; ++t) nz = nz || (D_NZ[s][t] && B_NZ[t][c]);
  return nz;
}

__device__ inline void block_at_point(const double* gi, const double* gj, const double (&d)[6][6], double w, bool diag,
                                      double (&acc)[3][3]) {
  const double bj[6][3] = {{gj[0], 0.0, 0.0}, {0.0, gj[1], 0.0}, {0.0, 0.0, gj[2]},
                           {gj[1], gj[0], 0.0}, {gj[2], 0.0, gj[0]}, {0.0, gj[2], gj[1]}};
  const double bi[6][3] = {{gi[0], 0.0, 0.0}, {0.0, gi[1], 0.0}, {0.0, 0.0, gi[2]},
                           {gi[1], gi[0], 0.0}, {gi[2], 0.0, gi[0]}, {0.0, gi[2], gi[1]}};
  double db[6][3];
#pragma unroll
  for (int s = 0; s < 6; ++s)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double sum = 0.0;
#pragma unroll
      for (int t = 0; t < 6; ++t)
        if (D_NZ[s][t] && B_NZ[t][c]) sum += d[s][t] * bj[t][c];
      db[s][c] = sum;
    }
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (diag && c < r) continue;
      double sum = 0.0;
#pragma unroll
      for (int s = 0; s < 6; ++s)
        if (B_NZ[s][r] && db_nz(s, c)) sum += bi[s][r] * db[s][c];
      acc[r][c] += w * sum;
    }
}

// material.rs:40-54
__device__ inline void isotropic(double young, double poisson, double (&d)[6][6]) {
  const double lambda = young * poisson / ((1.0 + poisson) * (1.0 - 2.0 * poisson));
  const double mu = young / (2.0 * (1.0 + poisson));
#pragma unroll
  for (int i = 0; i < 6; ++i)
#pragma unroll
    for (int j = 0; j < 6; ++j) d[i][j] = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
#pragma unroll
    for (int j = 0; j < 3; ++j) d[i][j] = lambda;
    d[i][i] = lambda + 2.0 * mu;
    d[i + 3][i + 3] = mu;
  }
}

// Packed upper triangle of a w x w matrix: (p, q), p <= q.
__host__ __device__ inline int packed(int w, int p, int q) { return p * w - p * (p - 1) / 2 + (q - p); }

// Node block pairs (i, j), i <= j, in the order j = 0.., i = 0..j.
__host__ __device__ inline void pair_of(int b, int& i, int& j) {
  j = 0;
  while ((j + 1) * (j + 2) / 2 <= b) ++j;
  i = b - j * (j + 1) / 2;
}

// Elements per workgroup of 256 threads: one (element, point) per thread for the Jacobians,
// then the (element, node block) or (element, entry) tasks strided over the threads.
constexpr int HEX_EPB = 32;   // 32 x 8 points; 32 x 36 node blocks
constexpr int TET_EPB = 64;   // 64 x 4 points; 64 x 55 node blocks
constexpr int MASS_EPB = 4;   // 4 x 64 points; 4 x 100 entries
constexpr int HEX_KS = 24 * 25 / 2;
constexpr int TET_KS = 30 * 31 / 2;

__device__ inline void mark_bad(unsigned long long* fail, i64 e) { atomicMin(fail, (unsigned long long)e); }

// Stiffness of 8-node hexahedra, 2 x 2 x 2 Gauss points (element.rs:416-422 with Rule::Hex(2)).
__global__ __launch_bounds__(256) void k_hex8_stiffness(const double* __restrict__ xyz, const int32_t* __restrict__ conn,
                                                        const int32_t* __restrict__ slot_elem,
                                                        const double* __restrict__ young, int one_young, double poisson,
                                                        i64 elements, double* __restrict__ ke,
                                                        unsigned long long* __restrict__ fail) {
  __shared__ double G[HEX_EPB][8][8][3];
  __shared__ double Wt[HEX_EPB][8];
  __shared__ int bad[HEX_EPB];
  const int t = threadIdx.x;
  const i64 e0 = (i64)blockIdx.x * HEX_EPB;
  if (t < HEX_EPB) bad[t] = 0;
  __syncthreads();
  if (t < HEX_EPB * 8) {
    const int el = t >> 3, p = t & 7;
    const i64 slot = e0 + el;
    if (slot < elements) {
      const i64 e = slot_elem[slot];
      double x[8][3];
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const i64 v = conn[e * 8 + a];
#pragma unroll
        for (int c = 0; c < 3; ++c) x[a][c] = xyz[v * 3 + c];
      }
      // quadrature.rs:87-99: xi fastest, then eta, then zeta
      const int px = p & 1, py = (p >> 1) & 1, pz = p >> 2;
      const double xi[3] = {G2X[px], G2X[py], G2X[pz]};
      const double wp = G2W[px] * G2W[py] * G2W[pz];
      double dn[8][3], g[8][3], det;
      hex8_dn(xi, dn);
      if (gradients<8>(x, dn, g, &det)) {
#pragma unroll
        for (int i = 0; i < 8; ++i)
#pragma unroll
          for (int c = 0; c < 3; ++c) G[el][p][i][c] = g[i][c];
        Wt[el][p] = det * wp;
      } else {
        bad[el] = 1;
        mark_bad(fail, e);
      }
    }
  }
  __syncthreads();
  for (int task = t; task < HEX_EPB * 36; task += 256) {
    const int el = task / 36, b = task % 36;
    const i64 slot = e0 + el;
    if (slot < elements && !bad[el]) {
      const i64 e = slot_elem[slot];
      int i, j;
      pair_of(b, i, j);
      double d[6][6];
      isotropic(young[one_young ? 0 : e], poisson, d);
      double acc[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll 1
      for (int p = 0; p < 8; ++p) block_at_point(G[el][p][i], G[el][p][j], d, Wt[el][p], i == j, acc);
      double* out = ke + slot * HEX_KS;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          if (i != j || c >= r) out[packed(24, 3 * i + r, 3 * j + c)] = acc[r][c];
    }
  }
}

// Stiffness of 10-node tetrahedra, the symmetric 4-point rule (element.rs:416-422, Rule::Tet4).
__global__ __launch_bounds__(256) void k_tet10_stiffness(const double* __restrict__ xyz, const int32_t* __restrict__ conn,
                                                         const int32_t* __restrict__ slot_elem,
                                                         const double* __restrict__ young, int one_young, double poisson,
                                                         i64 elements, double* __restrict__ ke,
                                                         unsigned long long* __restrict__ fail) {
  __shared__ double G[TET_EPB][4][10][3];
  __shared__ double Wt[TET_EPB][4];
  __shared__ int bad[TET_EPB];
  const int t = threadIdx.x;
  const i64 e0 = (i64)blockIdx.x * TET_EPB;
  if (t < TET_EPB) bad[t] = 0;
  __syncthreads();
  if (t < TET_EPB * 4) {
    const int el = t >> 2, p = t & 3;
    const i64 slot = e0 + el;
    if (slot < elements) {
      const i64 e = slot_elem[slot];
      double x[10][3];
#pragma unroll
      for (int a = 0; a < 10; ++a) {
        const i64 v = conn[e * 10 + a];
#pragma unroll
        for (int c = 0; c < 3; ++c) x[a][c] = xyz[v * 3 + c];
      }
      // quadrature.rs:100-106: point k near corner k
      const double pa = 0.5854101966249685, pb = 0.1381966011250105;
      const double xi[3] = {p == 1 ? pa : pb, p == 2 ? pa : pb, p == 3 ? pa : pb};
      double n[10], dn[10][3], g[10][3], det;
      tet10_shape(xi, n, dn);
      if (gradients<10>(x, dn, g, &det)) {
#pragma unroll
        for (int i = 0; i < 10; ++i)
#pragma unroll
          for (int c = 0; c < 3; ++c) G[el][p][i][c] = g[i][c];
        Wt[el][p] = det * (1.0 / 24.0);
      } else {
        bad[el] = 1;
        mark_bad(fail, e);
      }
    }
  }
  __syncthreads();
  for (int task = t; task < TET_EPB * 55; task += 256) {
    const int el = task / 55, b = task % 55;
    const i64 slot = e0 + el;
    if (slot < elements && !bad[el]) {
      const i64 e = slot_elem[slot];
      int i, j;
      pair_of(b, i, j);
      double d[6][6];
      isotropic(young[one_young ? 0 : e], poisson, d);
      double acc[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll 1
      for (int p = 0; p < 4; ++p) block_at_point(G[el][p][i], G[el][p][j], d, Wt[el][p], i == j, acc);
      double* out = ke + slot * TET_KS;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
          if (i != j || c >= r) out[packed(30, 3 * i + r, 3 * j + c)] = acc[r][c];
    }
  }
}

// Consistent mass of 10-node tetrahedra per displacement component, Stroud's conical product
// with 4 x 4 x 4 points (element.rs:427-441, Rule::TetConical(4), quadrature.rs:107-118).
__global__ __launch_bounds__(256) void k_tet10_mass(const double* __restrict__ xyz, const int32_t* __restrict__ conn,
                                                    const int32_t* __restrict__ slot_elem, double density, i64 elements,
                                                    double* __restrict__ me, unsigned long long* __restrict__ fail) {
  __shared__ double N[MASS_EPB][64][10];
  __shared__ double Wm[MASS_EPB][64];
  __shared__ int bad[MASS_EPB];
  const int t = threadIdx.x;
  const i64 e0 = (i64)blockIdx.x * MASS_EPB;
  if (t < MASS_EPB) bad[t] = 0;
  __syncthreads();
  if (t < MASS_EPB * 64) {
    const int el = t >> 6, p = t & 63;
    const i64 slot = e0 + el;
    if (slot < elements) {
      const i64 e = slot_elem[slot];
      double x[10][3];
#pragma unroll
      for (int a = 0; a < 10; ++a) {
        const i64 v = conn[e * 10 + a];
#pragma unroll
        for (int c = 0; c < 3; ++c) x[a][c] = xyz[v * 3 + c];
      }
      const int ia = p >> 4, ib = (p >> 2) & 3, ic = p & 3;
      const double u = (G4X[ia] + 1.0) / 2.0, v = (G4X[ib] + 1.0) / 2.0, w = (G4X[ic] + 1.0) / 2.0;
      const double wu = G4W[ia] / 2.0, wv = G4W[ib] / 2.0, ww = G4W[ic] / 2.0;
      const double xi[3] = {u, v * (1.0 - u), w * (1.0 - u) * (1.0 - v)};
      const double wp = wu * wv * ww * (1.0 - u) * (1.0 - u) * (1.0 - v);
      double n[10], dn[10][3], det;
      tet10_shape(xi, n, dn);
      if (gradients<10>(x, dn, nullptr, &det)) {
#pragma unroll
        for (int i = 0; i < 10; ++i) N[el][p][i] = n[i];
        Wm[el][p] = density * det * wp;
      } else {
        bad[el] = 1;
        mark_bad(fail, e);
      }
    }
  }
  __syncthreads();
  for (int task = t; task < MASS_EPB * 100; task += 256) {
    const int el = task / 100, ij = task % 100, i = ij / 10, j = ij % 10;
    const i64 slot = e0 + el;
    if (slot < elements && !bad[el]) {
      double acc = 0.0;
      for (int p = 0; p < 64; ++p) acc += Wm[el][p] * N[el][p][i] * N[el][p][j];
      me[slot * 100 + ij] = acc;
    }
  }
}

// The contributions q in [q, q1) of one entry, summed in order from 0.0: loaded B at a time as
// independent loads (positions, then values), so a thread waits for two loads per batch rather
// than two per contribution.
template <int B>
__device__ inline double gather_sum(i64 q, const i64 q1, const uint32_t* __restrict__ src,
                                    const double* __restrict__ buf) {
  double v = 0.0;
  while (q < q1) {
    const int n = (int)(q1 - q < B ? q1 - q : B);
    uint32_t ix[B];
    double x[B];
#pragma unroll
    for (int i = 0; i < B; ++i) ix[i] = i < n ? src[q + i] : 0u;
#pragma unroll
    for (int i = 0; i < B; ++i) x[i] = i < n ? buf[ix[i]] : 0.0;
#pragma unroll
    for (int i = 0; i < B; ++i)
      if (i < n) v += x[i];
    q += n;
  }
  return v;
}

// Each entry of the upper triangle: its contributions summed in element order from 0.0
// (sparse.rs:199-243 onto a zero matrix), reduced as 0.0 + (1.0 v) 1.0 = v (constraints.rs:
// 888-890), then K + (-shift) M (sparse.rs:85) for the shifted matrix.
__global__ __launch_bounds__(256) void k_assemble_gather(i64 nnz, const i64* __restrict__ kptr,
                                                         const uint32_t* __restrict__ ksrc, const double* __restrict__ ke,
                                                         const i64* __restrict__ mptr, const uint32_t* __restrict__ msrc,
                                                         const double* __restrict__ me, int what, double neg_shift,
                                                         double* __restrict__ out) {
  // Workgroups are dealt to the 8 XCDs in turn; the grid is a multiple of 8 and XCD x takes
  // the x-th contiguous range of entries, so neighbouring entries, which read the same element
  // matrices, share one L2.
  const unsigned per = gridDim.x / 8;
  const i64 t = ((i64)(blockIdx.x % 8) * per + blockIdx.x / 8) * 256 + threadIdx.x;
  if (t >= nnz) return;
  double vk = 0.0, vm = 0.0;
  if (what != FC_ASSEMBLE_M) vk = gather_sum<8>(kptr[t], kptr[t + 1], ksrc, ke);
  if (what != FC_ASSEMBLE_K) vm = gather_sum<8>(mptr[t], mptr[t + 1], msrc, me);
  out[t] = what == FC_ASSEMBLE_K ? vk : what == FC_ASSEMBLE_M ? vm : vk + neg_shift * vm;
}

__global__ void k_fill_u64(unsigned long long* p, unsigned long long v) { *p = v; }

}  // namespace

// ---------------------------------------------------------------------------------------------

struct fc_assembly {
  int kind = 0, npe = 0, W = 0, KS = 0, device = 0;
  bool mass = false;
  i64 nodes = 0, elements = 0, neq = 0, nnz = 0;
  double poisson = 0.0, density = 0.0;
  int one_young = 1;
  double* xyz = nullptr;
  int32_t* conn = nullptr;
  int32_t* slot_elem = nullptr;  // element stored at each slot of the element-matrix buffers
  double* young = nullptr;
  i64* kptr = nullptr;
  uint32_t* ksrc = nullptr;
  i64* mptr = nullptr;
  uint32_t* msrc = nullptr;
  double* ke = nullptr;
  double* me = nullptr;
  unsigned long long* fail = nullptr;
  unsigned long long* fail_host = nullptr;
  hipStream_t own = nullptr;
  bool timing = false;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double last_elements_s = 0.0, last_gather_s = 0.0;
  uint64_t contributions_k = 0, contributions_m = 0;

  void release() {
    int cur = 0;
    (void)hipGetDevice(&cur);
    (void)hipSetDevice(device);
    for (void* p : {(void*)xyz, (void*)conn, (void*)slot_elem, (void*)young, (void*)kptr, (void*)ksrc, (void*)mptr, (void*)msrc,
                    (void*)ke, (void*)me, (void*)fail})
      if (p) (void)hipFree(p);
    if (fail_host) (void)hipHostFree(fail_host);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (own) (void)hipStreamDestroy(own);
    (void)hipSetDevice(cur);
  }
};

namespace {

// The gather plan over the caller's pattern. For upper entry t of column k (equation k, mesh
// dof d = eq_dof[k] at node v, direction r): every element e containing v (ascending e, local
// node a), every local node b and direction c whose equation j = eq(3 conn[e][b] + c) is
// <= k contributes Ke[3a + r][3b + c] (from_symmetric reads row k, columns j <= k); the mass
// only where c == r, Me[a][b] (add_scalar puts the scalar matrix on each direction).
//
// Element matrices are stored by slot: elements in the order of the first equation they touch
// (ties by element), so the columns the gather walks in order read element matrices that lie
// together. Only where a matrix is stored changes; each entry still sums its contributions in
// element order.
void build_plan(fc_assembly& A, const i64* conn, const i64* eq_dof, const uint64_t* ptr, const uint64_t* idx,
                std::vector<i64>& kptr, std::vector<uint32_t>& ksrc, std::vector<i64>& mptr, std::vector<uint32_t>& msrc,
                std::vector<int32_t>& slot_elem) {
  const i64 nodes = A.nodes, E = A.elements, neq = A.neq;
  const int npe = A.npe, W = A.W, KS = A.KS;
  std::vector<i64> eq(3 * nodes, -1);
  for (i64 k = 0; k < neq; ++k) {
    const i64 d = eq_dof[k];
    if (d < 0 || d >= 3 * nodes || eq[d] >= 0) throw ArgError{"eq_dof is not a one-to-one map onto mesh dofs"};
    eq[d] = k;
  }
  // node -> (element, local index), elements ascending
  std::vector<i64> nptr(nodes + 1, 0);
  for (i64 e = 0; e < E; ++e)
    for (int a = 0; a < npe; ++a) {
      const i64 v = conn[e * npe + a];
      if (v < 0 || v >= nodes) throw ArgError{"an element names a node outside the mesh"};
      ++nptr[v + 1];
    }
  for (i64 v = 0; v < nodes; ++v) nptr[v + 1] += nptr[v];
  {
    std::vector<std::pair<i64, i64>> key(E);
    for (i64 e = 0; e < E; ++e) {
      i64 first = INT64_MAX;
      for (int a = 0; a < npe; ++a)
        for (int c = 0; c < 3; ++c) {
          const i64 j = eq[3 * conn[e * npe + a] + c];
          if (j >= 0) first = std::min(first, j);
        }
      key[e] = {first, e};
    }
    std::sort(key.begin(), key.end());
    slot_elem.resize(E);
    for (i64 s = 0; s < E; ++s) slot_elem[s] = (int32_t)key[s].second;
  }
  std::vector<i64> slot_of(E);
  for (i64 s = 0; s < E; ++s) slot_of[slot_elem[s]] = s;
  std::vector<i64> nel(nptr[nodes]);
  {
    std::vector<i64> cur(nptr.begin(), nptr.end() - 1);
    for (i64 e = 0; e < E; ++e)
      for (int a = 0; a < npe; ++a) nel[cur[conn[e * npe + a]]++] = e * npe + a;
  }
  const i64 nnz = A.nnz;
  kptr.assign(nnz + 1, 0);
  if (A.mass) mptr.assign(nnz + 1, 0);
  unsigned hw = std::thread::hardware_concurrency();
  const int nt = (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
  std::vector<std::string> errors(nt);
  // pass 0 counts, pass 1 fills
  for (int pass = 0; pass < 2; ++pass) {
    if (pass == 1) {
      for (i64 t = 0; t < nnz; ++t) kptr[t + 1] += kptr[t];
      ksrc.assign(kptr[nnz], 0);
      if (A.mass) {
        for (i64 t = 0; t < nnz; ++t) mptr[t + 1] += mptr[t];
        msrc.assign(mptr[nnz], 0);
      }
    }
    std::vector<std::thread> pool;
    for (int w = 0; w < nt; ++w)
      pool.emplace_back([&, w] {
        std::vector<i64> curk, curm;
        for (i64 k = w; k < neq; k += nt) {
          const i64 d = eq_dof[k], v = d / 3;
          const int r = (int)(d % 3);
          const i64 c0 = (i64)ptr[k], c1 = (i64)ptr[k + 1];
          if (pass == 1) {
            curk.assign(kptr.begin() + c0, kptr.begin() + c1);
            if (A.mass) curm.assign(mptr.begin() + c0, mptr.begin() + c1);
          }
          for (i64 q = nptr[v]; q < nptr[v + 1]; ++q) {
            const i64 e = nel[q] / npe;
            const int a = (int)(nel[q] % npe);
            for (int b = 0; b < npe; ++b)
              for (int c = 0; c < 3; ++c) {
                const i64 j = eq[3 * conn[e * npe + b] + c];
                if (j < 0 || j > k) continue;
                const uint64_t* lo = idx + c0;
                const uint64_t* hi = idx + c1;
                const uint64_t* at = std::lower_bound(lo, hi, (uint64_t)j);
                if (at == hi || *at != (uint64_t)j) {
                  errors[w] = "the pattern lacks an entry that an element couples";
                  return;
                }
                const i64 tl = at - lo;
                const int p = 3 * a + r, s = 3 * b + c;
                if (pass == 0) {
                  ++kptr[c0 + tl + 1];
                  if (A.mass && c == r) ++mptr[c0 + tl + 1];
                } else {
                  const i64 at_e = slot_of[e];
                  ksrc[curk[tl]++] = (uint32_t)(at_e * KS + (p <= s ? packed(W, p, s) : packed(W, s, p)));
                  if (A.mass && c == r) msrc[curm[tl]++] = (uint32_t)(at_e * 100 + a * 10 + b);
                }
              }
          }
        }
      });
    for (auto& th : pool) th.join();
    for (auto& m : errors)
      if (!m.empty()) throw ArgError{m};
  }
}

}  // namespace

extern "C" {

int fc_assembly_new(int32_t element_kind, uint64_t nodes, const double* xyz, uint64_t elements, const int64_t* conn,
                    const double* young, uint64_t n_young, double poisson, double density, uint64_t neq,
                    const int64_t* eq_dof, const uint64_t* ptr, const uint64_t* idx, int32_t device, int32_t timing,
                    fc_assembly** out, fc_error* err) {
  if (out) *out = nullptr;
  if (!out || (element_kind != FC_ELEMENT_HEX8 && element_kind != FC_ELEMENT_TET10) || (nodes && !xyz) ||
      (elements && !conn) || !young || (n_young != 1 && n_young != elements) || (neq && (!eq_dof || !idx)) || !ptr) {
    put_error(err, FC_INVALID_ARGUMENT, "bad arguments to fc_assembly_new");
    return FC_INVALID_ARGUMENT;
  }
  // material.rs:22-27: E > 0 and -1 < nu < 0.5
  for (uint64_t i = 0; i < n_young; ++i)
    if (!(young[i] > 0.0 && std::isfinite(young[i]) && poisson > -1.0 && poisson < 0.5)) {
      put_error(err, FC_INVALID_ARGUMENT, "isotropic constants are not a stable material");
      return FC_INVALID_ARGUMENT;
    }
  if (nodes > (uint64_t)INT32_MAX) {
    put_error(err, FC_INVALID_ARGUMENT, "more nodes than 32-bit node numbers hold");
    return FC_INVALID_ARGUMENT;
  }
  auto* A = new (std::nothrow) fc_assembly();
  if (!A) {
    put_error(err, FC_OUT_OF_MEMORY, "host memory");
    return FC_OUT_OF_MEMORY;
  }
  A->kind = element_kind;
  A->npe = element_kind == FC_ELEMENT_HEX8 ? 8 : 10;
  A->W = 3 * A->npe;
  A->KS = A->W * (A->W + 1) / 2;
  A->mass = element_kind == FC_ELEMENT_TET10 && density > 0.0 && std::isfinite(density);
  A->nodes = (i64)nodes;
  A->elements = (i64)elements;
  A->neq = (i64)neq;
  A->nnz = (i64)ptr[neq];
  A->poisson = poisson;
  A->density = density;
  A->one_young = n_young == 1 ? 1 : 0;
  A->device = device;
  A->timing = timing != 0;
  if ((double)elements * (double)A->KS >= 4294967296.0) {
    delete A;
    put_error(err, FC_INVALID_ARGUMENT, "too many elements for one assembly (32-bit element-matrix positions)");
    return FC_INVALID_ARGUMENT;
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  try {
    std::vector<i64> kptr, mptr;
    std::vector<uint32_t> ksrc, msrc;
    std::vector<int32_t> slots;
    build_plan(*A, conn, eq_dof, ptr, idx, kptr, ksrc, mptr, msrc, slots);
    A->contributions_k = ksrc.size();
    A->contributions_m = msrc.size();
    std::vector<int32_t> conn32(elements * A->npe);
    for (size_t i = 0; i < conn32.size(); ++i) conn32[i] = (int32_t)conn[i];
    ASM_CHECK(hipSetDevice(device));
    ASM_CHECK(hipStreamCreateWithFlags(&A->own, hipStreamNonBlocking));
    hipStream_t st = A->own;
    A->xyz = dupload(xyz, 3 * nodes, st);
    A->conn = dupload(conn32.data(), conn32.size(), st);
    A->slot_elem = dupload(slots.data(), slots.size(), st);
    A->young = dupload(young, n_young, st);
    A->kptr = dupload(kptr.data(), kptr.size(), st);
    A->ksrc = dupload(ksrc.data(), ksrc.size(), st);
    if (A->mass) {
      A->mptr = dupload(mptr.data(), mptr.size(), st);
      A->msrc = dupload(msrc.data(), msrc.size(), st);
    }
    A->fail = dalloc<unsigned long long>(1);
    ASM_CHECK(hipHostMalloc((void**)&A->fail_host, sizeof(unsigned long long), hipHostMallocDefault));
    if (A->timing)
      for (auto& e : A->ev) ASM_CHECK(hipEventCreate(&e));
    ASM_CHECK(hipStreamSynchronize(st));
  } catch (const ArgError& a) {
    A->release();
    delete A;
    (void)hipSetDevice(cur);
    put_error(err, FC_INVALID_ARGUMENT, a.msg);
    return FC_INVALID_ARGUMENT;
  } catch (const std::bad_alloc&) {
    A->release();
    delete A;
    (void)hipSetDevice(cur);
    put_error(err, FC_OUT_OF_MEMORY, "not enough memory for the assembly plan");
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    A->release();
    delete A;
    (void)hipSetDevice(cur);
    put_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  (void)hipSetDevice(cur);
  *out = A;
  put_error(err, FC_OK, "");
  return FC_OK;
}

int fc_assemble_device(fc_assembly* A, int32_t what, double shift, double* upper_val_device, uint64_t nnz,
                       void* stream, fc_error* err) {
  if (!A || !upper_val_device || (what != FC_ASSEMBLE_K && what != FC_ASSEMBLE_M && what != FC_ASSEMBLE_SHIFTED) ||
      nnz != (uint64_t)A->nnz || !std::isfinite(shift)) {
    put_error(err, FC_INVALID_ARGUMENT, "bad arguments to fc_assemble_device");
    return FC_INVALID_ARGUMENT;
  }
  if (what != FC_ASSEMBLE_K && !A->mass) {
    put_error(err, FC_INVALID_ARGUMENT, "the mass needs 10-node tetrahedra and a positive density");
    return FC_INVALID_ARGUMENT;
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  try {
    ASM_CHECK(hipSetDevice(A->device));
    hipStream_t st = stream ? (hipStream_t)stream : A->own;
    const bool stiff = what != FC_ASSEMBLE_M, mass = what != FC_ASSEMBLE_K;
    if (stiff && !A->ke) A->ke = dalloc<double>((size_t)A->elements * A->KS);
    if (mass && !A->me) A->me = dalloc<double>((size_t)A->elements * 100);
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(1), 0, st, A->fail, ~0ull);
    if (A->timing) ASM_CHECK(hipEventRecord(A->ev[0], st));
    if (A->elements > 0) {
      if (stiff && A->kind == FC_ELEMENT_HEX8)
        hipLaunchKernelGGL(k_hex8_stiffness, dim3((unsigned)((A->elements + HEX_EPB - 1) / HEX_EPB)), dim3(256), 0, st,
                           A->xyz, A->conn, A->slot_elem, A->young, A->one_young, A->poisson, A->elements, A->ke,
                           A->fail);
      if (stiff && A->kind == FC_ELEMENT_TET10)
        hipLaunchKernelGGL(k_tet10_stiffness, dim3((unsigned)((A->elements + TET_EPB - 1) / TET_EPB)), dim3(256), 0, st,
                           A->xyz, A->conn, A->slot_elem, A->young, A->one_young, A->poisson, A->elements, A->ke,
                           A->fail);
      if (mass)
        hipLaunchKernelGGL(k_tet10_mass, dim3((unsigned)((A->elements + MASS_EPB - 1) / MASS_EPB)), dim3(256), 0, st,
                           A->xyz, A->conn, A->slot_elem, A->density, A->elements, A->me, A->fail);
    }
    if (A->timing) ASM_CHECK(hipEventRecord(A->ev[1], st));
    if (A->nnz > 0)
      hipLaunchKernelGGL(k_assemble_gather, dim3((unsigned)(((A->nnz + 255) / 256 + 7) / 8 * 8)), dim3(256), 0, st,
                         A->nnz, A->kptr,
                         A->ksrc, A->ke, A->mptr, A->msrc, A->me, what, -shift, upper_val_device);
    if (A->timing) ASM_CHECK(hipEventRecord(A->ev[2], st));
    ASM_CHECK(hipGetLastError());
    ASM_CHECK(hipMemcpyAsync(A->fail_host, A->fail, sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
    ASM_CHECK(hipStreamSynchronize(st));
    if (A->timing) {
      float a = 0.f, b = 0.f;
      ASM_CHECK(hipEventElapsedTime(&a, A->ev[0], A->ev[1]));
      ASM_CHECK(hipEventElapsedTime(&b, A->ev[1], A->ev[2]));
      A->last_elements_s = a * 1e-3;
      A->last_gather_s = b * 1e-3;
    }
    (void)hipSetDevice(cur);
    if (*A->fail_host != ~0ull) {
      char msg[160];
      std::snprintf(msg, sizeof msg, "element %llu is inverted or degenerate (its Jacobian is not positive)",
                    (unsigned long long)*A->fail_host);
      put_error(err, FC_DEGENERATE_ELEMENT, msg, *A->fail_host);
      return FC_DEGENERATE_ELEMENT;
    }
  } catch (const std::bad_alloc&) {
    (void)hipSetDevice(cur);
    put_error(err, FC_OUT_OF_MEMORY, "not enough device memory for the element matrices");
    return FC_OUT_OF_MEMORY;
  } catch (const DeviceError& d) {
    (void)hipSetDevice(cur);
    put_error(err, FC_DEVICE, std::string(d.what) + ": " + hipGetErrorString(d.e));
    return FC_DEVICE;
  }
  put_error(err, FC_OK, "");
  return FC_OK;
}

int fc_assembly_info(const fc_assembly* A, uint64_t* contributions_k, uint64_t* contributions_m, double* elements_s,
                     double* gather_s) {
  if (!A) return FC_INVALID_ARGUMENT;
  if (contributions_k) *contributions_k = A->contributions_k;
  if (contributions_m) *contributions_m = A->contributions_m;
  if (elements_s) *elements_s = A->last_elements_s;
  if (gather_s) *gather_s = A->last_gather_s;
  return FC_OK;
}

void fc_assembly_free(fc_assembly* A) {
  if (!A) return;
  A->release();
  delete A;
}

}  // extern "C"
