// Kernels of the multifrontal numeric phase (gfx950). All are work-list kernels: workgroup b
// runs item items[item0 + b] of its launch; one launch covers one level of the front tree.
//
// Per level: the children's update matrices are added into the fronts' L parts (extend-add);
// each front is factored by panels of 64 columns, left-looking inside the front like the
// reference's panel loop (cholesky.rs:1138-1186): the panel minus its products with the
// front's earlier columns (MFMA GEMM), its diagonal block factored with the pivot test of
// cholesky.rs:1161, the rows below solved against it; then the front's update matrix
// U = -L21 L21^T (MFMA SYRK) and the children's update matrices added into it.
#pragma once

#include "device.hpp"

namespace fc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr double PIVOT_TOLERANCE = 1e-10;  // cholesky.rs:48

// ------------------------------------------------------------------------------------------
// Matrix entries into the factor storage

template <typename T>
__global__ void k_scatter_a(const double* __restrict__ aval, const int64_t* __restrict__ a_pos, int64_t nnz,
                            T* __restrict__ L) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
    L[a_pos[e]] = (T)aval[e];
}

__global__ void k_gather_diag(const double* __restrict__ aval, const int64_t* __restrict__ diag_src, int64_t n,
                              double* __restrict__ diag) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x)
    diag[c] = diag_src[c] >= 0 ? aval[diag_src[c]] : 0.0;
}

// Per front: the exponent e of the smallest power of two with sqrt(A_ii) <= 2^e over the
// front's rows. For an SPD matrix every factor entry satisfies |L_ij| <= sqrt(A_ii) (row i of L
// has squared norm A_ii), so 2^-e scales the front's columns into [-1, 1]. 0 when no row has a
// positive finite diagonal.
__global__ __launch_bounds__(256) void k_front_scale(const FrontDev* __restrict__ fronts, const int32_t* __restrict__ rows,
                                                     const double* __restrict__ diag, int64_t nf,
                                                     int32_t* __restrict__ fexp) {
  __shared__ double red[4];
  for (int64_t f = blockIdx.x; f < nf; f += gridDim.x) {
    const FrontDev F = fronts[f];
    const int32_t* fr = rows + F.rows_off;
    double mx = 0.0;
    for (int i = threadIdx.x; i < F.m; i += blockDim.x) {
      const double d = diag[fr[i]];
      if (isfinite(d) && d > mx) mx = d;
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
      mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
      int e = 0;
      if (mx > 0.0) (void)frexp(sqrt(mx), &e);
      fexp[f] = e;
    }
    __syncthreads();
  }
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t len, T v) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x) p[e] = v;
}

// ------------------------------------------------------------------------------------------
// Extend-add: the workgroup owns the parent's columns [a, b) and adds every child's update
// matrix entries that land there, child after child (in increasing order), so each entry
// receives its contributions in a fixed order and no two workgroups write the same entry.

template <typename T>
__global__ __launch_bounds__(256) void k_extend_add(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                    const int32_t* __restrict__ kids, const int32_t* __restrict__ rel,
                                                    T* __restrict__ L, const T* U, T* Uw) {
  // U (children, read) and Uw (this front, written) are disjoint blocks of one arena
  const Item it = items[blockIdx.x];
  const FrontDev P = fronts[it.f];
  const int j0 = it.a, j1 = it.b;
  const int pw = P.w;
  const int64_t pmu = P.m - P.w;
  for (int q = P.kid0; q < P.kid1; ++q) {
    const FrontDev C = fronts[kids[q]];
    const int cmu = C.m - C.w;
    if (cmu == 0) continue;
    const int32_t* r = rel + C.rel_off;
    // child columns whose parent column lies in [j0, j1): a contiguous run (rel rises)
    int lo = 0, hi = cmu;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (r[mid] < j0) lo = mid + 1; else hi = mid; }
    int jc0 = lo;
    lo = jc0; hi = cmu;
    while (lo < hi) { int mid = (lo + hi) >> 1; if (r[mid] < j1) lo = mid + 1; else hi = mid; }
    const int jc1 = lo;
    const T* cu = U + C.upos;
    for (int jc = jc0; jc < jc1; ++jc) {
      const int jp = r[jc];
      const T* src = cu + (int64_t)jc * cmu;
      if (jp < pw) {
        T* dst = L + P.lpos + (int64_t)jp * P.ld;
        for (int ic = jc + threadIdx.x; ic < cmu; ic += blockDim.x) dst[r[ic]] += src[ic];
      } else {
        T* dst = Uw + P.upos + (int64_t)(jp - pw) * pmu - pw;
        for (int ic = jc + threadIdx.x; ic < cmu; ic += blockDim.x) dst[r[ic]] += src[ic];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------
// GEMM tiles on the matrix cores. A 64 x 64 output tile per workgroup of 4 waves, each wave a
// 32 x 32 quarter as 2 x 2 blocks of the 16x16x4 MFMA. Operands are rows of the front's L part:
// C(i, j) = sum_k L(ia + i, k) L(jb + j, k), k in [k0, k1). Lane l of a 16x16x4 MFMA holds
// A(l % 16, l / 16) and B(l / 16, l % 16); its four outputs v = 0..3 are D(row(l, v), l % 16)
// with row = 4 (l / 16) + v for FP32 and row = l / 16 + 4 v for FP64 (measured on gfx950).

template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mfma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

constexpr int KB = 32;        // k columns staged per step
constexpr int LDS_PAD = 4;    // entries of padding per staged column

template <typename T>
__device__ __forceinline__ void gemm_tile(const T* __restrict__ L, int64_t lpos, int ld, int m, int ia, int jb,
                                          int k0, int k1, typename Mfma<T>::acc_t acc[2][2]) {
  __shared__ T As[KB][TILE_M + LDS_PAD];
  __shared__ T Bs[KB][TILE_N + LDS_PAD];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = typename Mfma<T>::acc_t{0, 0, 0, 0};
  const int row = tid & 63, kq = tid >> 6;  // loader: row within the tile, k quarter
  const bool arow = ia + row < m, brow = jb + row < m;
  for (int ks = k0; ks < k1; ks += KB) {
    const int kn = min(KB, k1 - ks);
#pragma unroll
    for (int r = 0; r < KB / 4; ++r) {
      const int k = kq + 4 * r;
      const int64_t col = lpos + (int64_t)(ks + k) * ld;
      As[k][row] = (k < kn && arow) ? L[col + ia + row] : T(0);
      Bs[k][row] = (k < kn && brow) ? L[col + jb + row] : T(0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KB; kk += 4) {
      T a[2], b[2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi) a[bi] = As[kk + (lane >> 4)][wi + 16 * bi + (lane & 15)];
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) b[bj] = Bs[kk + (lane >> 4)][wj + 16 * bj + (lane & 15)];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) acc[bi][bj] = Mfma<T>::step(a[bi], b[bj], acc[bi][bj]);
    }
    __syncthreads();
  }
}

// Panel update: L[i0 + i, c0 + j] -= sum_{k < c0} L(i0 + i, k) L(c0 + j, k) for rows
// i0 + i < M and panel columns c0 + j < c1, lower part (row >= column) only.
// Item: f, a = i0, b = c0, c = c1.
template <typename T>
__global__ __launch_bounds__(256) void k_update(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                T* __restrict__ L) {
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int i0 = it.a, c0 = it.b, c1 = it.c;
  typename Mfma<T>::acc_t acc[2][2];
  gemm_tile<T>(L, F.lpos, F.ld, F.m, i0, c0, 0, c0, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + 16 * bi + Mfma<T>::row(lane, v);
        const int j = c0 + wj + 16 * bj + (lane & 15);
        if (i < F.m && j < c1 && i >= j) L[F.lpos + (int64_t)j * F.ld + i] -= acc[bi][bj][v];
      }
}

// Update matrix: U(i, j) = -sum_{k < W} L(W + i, k) L(W + j, k), lower tiles, lower entries.
// Item: f, a = i0, b = j0 (update-matrix coordinates).
template <typename T>
__global__ __launch_bounds__(256) void k_syrk(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                              const T* __restrict__ L, T* __restrict__ U) {
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int i0 = it.a, j0 = it.b;
  const int mu = F.m - F.w;
  typename Mfma<T>::acc_t acc[2][2];
  gemm_tile<T>(L, F.lpos, F.ld, F.m, F.w + i0, F.w + j0, 0, F.w, acc);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wi = (wave >> 1) * 32, wj = (wave & 1) * 32;
  for (int bi = 0; bi < 2; ++bi)
    for (int bj = 0; bj < 2; ++bj)
      for (int v = 0; v < 4; ++v) {
        const int i = i0 + wi + 16 * bi + Mfma<T>::row(lane, v);
        const int j = j0 + wj + 16 * bj + (lane & 15);
        if (i < mu && j < mu && i >= j) U[F.upos + (int64_t)j * mu + i] = -acc[bi][bj][v];
      }
}

// ------------------------------------------------------------------------------------------
// Diagonal block of a panel, factored column by column in LDS with the reference's pivot test
// (cholesky.rs:1146-1171): a pivot not above 1e-10 |A_jj| (or not finite) fails the column.
// The smallest failing column (postorder) is kept in *fail. Item: f, a = c0.

template <typename T>
__global__ __launch_bounds__(256) void k_potrf(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                               T* __restrict__ L, const double* __restrict__ diag,
                                               unsigned long long* __restrict__ fail) {
  __shared__ T S[PANEL][PANEL + 1];
  __shared__ int bad;
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(PANEL, F.w - c0);
  T* base = L + F.lpos + (int64_t)c0 * F.ld + c0;
  for (int e = threadIdx.x; e < PANEL * PANEL; e += blockDim.x) {
    const int i = e % PANEL, j = e / PANEL;
    S[i][j] = (i < pw && j < pw && i >= j) ? base[(int64_t)j * F.ld + i] : T(0);
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int c = 0; c < pw; ++c) {
    if (threadIdx.x == 0) {
      const double d = (double)S[c][c];
      const double a = diag[F.col0 + c0 + c];
      if (!(d > PIVOT_TOLERANCE * fabs(a)) || !isfinite(d)) {
        if (!bad) atomicMin(fail, (unsigned long long)(F.col0 + c0 + c));
        bad = 1;
      }
      S[c][c] = (T)sqrt(d);
    }
    __syncthreads();
    const T d = S[c][c];
    for (int i = c + 1 + threadIdx.x; i < pw; i += blockDim.x) S[i][c] /= d;
    __syncthreads();
    const int rest = pw - c - 1;
    for (int e = threadIdx.x; e < rest * rest; e += blockDim.x) {
      const int i = c + 1 + e % rest, j = c + 1 + e / rest;
      if (i >= j) S[i][j] -= S[i][c] * S[j][c];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < pw * pw; e += blockDim.x) {
    const int i = e % pw, j = e / pw;
    if (i >= j) base[(int64_t)j * F.ld + i] = S[i][j];
  }
}

// Rows [r0, r0 + 64) (below the panel) solved against the panel's factored diagonal block:
// X L11^T = B, row by row, column after column (the operations of solve_below,
// cholesky.rs:691-720). Four threads share a row. Item: f, a = c0, b = r0.
template <typename T>
__global__ __launch_bounds__(256) void k_trsm(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                              T* __restrict__ L) {
  __shared__ T D[PANEL][PANEL + 1];
  __shared__ T X[TRSM_ROWS][PANEL + 1];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, r0 = it.b, pw = min(PANEL, F.w - c0);
  const int h = min(TRSM_ROWS, F.m - r0);
  const T* dbase = L + F.lpos + (int64_t)c0 * F.ld + c0;
  T* xbase = L + F.lpos + (int64_t)c0 * F.ld + r0;
  for (int e = threadIdx.x; e < PANEL * PANEL; e += blockDim.x) {
    const int i = e % PANEL, j = e / PANEL;
    D[i][j] = (i < pw && j < pw && i >= j) ? dbase[(int64_t)j * F.ld + i] : T(0);
  }
  for (int e = threadIdx.x; e < TRSM_ROWS * PANEL; e += blockDim.x) {
    const int i = e % TRSM_ROWS, j = e / TRSM_ROWS;
    X[i][j] = (i < h && j < pw) ? xbase[(int64_t)j * F.ld + i] : T(0);
  }
  __syncthreads();
  const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
  for (int c = 0; c < pw; ++c) {
    T s = 0;
    for (int k = q; k < c; k += 4) s += X[r][k] * D[c][k];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (q == 0) X[r][c] = (X[r][c] - s) / D[c][c];
    __syncthreads();
  }
  for (int e = threadIdx.x; e < TRSM_ROWS * PANEL; e += blockDim.x) {
    const int i = e % TRSM_ROWS, j = e / TRSM_ROWS;
    if (i < h && j < pw) xbase[(int64_t)j * F.ld + i] = X[i][j];
  }
}

}  // namespace fc
