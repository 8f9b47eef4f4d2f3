// Dense kernels of a front's partial factorization on the matrix cores (gfx950).
//
// Every GEMM-shaped step of the numeric phase is one item of one kernel, k_gemm:
//   panel update   L[r.., c0..c1) -= L[r.., 0..c0) L[c0..c1, 0..c0)^T      (G_SUB)
//   rows below     L[r.., c0..c1)  = L[r.., c0..c1) Linv^T                   (G_SET, in place)
//   update matrix  U[i.., j..)     = -L[W+i.., 0..W) L[W+j.., 0..W)^T        (G_NEG)
// where Linv is the inverse of the panel's factored diagonal block (k_potrf_inv). A launch
// with few output tiles is split along k: each piece writes its partial tile (G_PART) and
// k_reduce_parts adds the pieces in a fixed order, so results do not depend on scheduling.
//
// Tile: 128 x 128 per workgroup of 4 waves, each wave a 64 x 64 quarter as 4 x 4 blocks of the
// 16x16x4 MFMA; k staged 16 at a time in LDS, the next stage prefetched into registers while
// the current one is multiplied (one barrier a stage). The MFMA's first operand is taken from
// the j side so that a lane's output column index runs along C's rows: 16 lanes store 16
// consecutive rows of a column (coalesced, C is column-major).
#pragma once

#include "device.hpp"

namespace fc {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int GB = 128;  // tile edge
constexpr int GK = 16;   // k per LDS stage
constexpr int NB = 128;  // panel width of the factorization (columns of a diagonal block)

enum GemmMode : int32_t { G_SUB = 0, G_NEG = 1, G_SET = 2, G_PART = 3 };
enum GemmSrc : int8_t { S_L = 0, S_U = 1, S_INV = 2, S_PART = 3 };

struct GemmItem {
  int64_t a, b, c;        // element offsets: A(i, k) = A[a + k lda + i], B(j, k), C(i, j) = C[c + j ldc + i]
  int32_t lda, ldb, ldc;
  int32_t m, n, k;        // rows of C (and A), columns of C (rows of B), depth
  int32_t mode;           // GemmMode
  int32_t diag;           // store C(i, j) only when i - j + diag >= 0
  int8_t asrc, bsrc, csrc, pad0;
  int32_t pad1;
};
static_assert(sizeof(GemmItem) == 64, "GemmItem layout");

struct ReduceItem {
  int64_t c;              // output offset in csrc
  int64_t part0;          // first partial tile (GB x GB, column-major) in the split-K workspace
  int32_t ldc, m, n, mode, diag, parts;
  int8_t csrc, pad0, pad1, pad2;
  int32_t pad3;
};
static_assert(sizeof(ReduceItem) == 48, "ReduceItem layout");

template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef v4d acc_t;
  static __device__ __forceinline__ acc_t step(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output v of lane l is D(row(l, v), l % 16) (measured on gfx950)
  static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <>
struct Mma<float> {
  typedef v4f acc_t;
  static __device__ __forceinline__ acc_t step(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

template <typename T>
__global__ __launch_bounds__(256, 2) void k_gemm(const GemmItem* __restrict__ items, T* L, T* U, T* Inv, T* Part) {
  constexpr int PAD = 4;
  __shared__ T As[2][GK][GB + PAD];
  __shared__ T Bs[2][GK][GB + PAD];
  const GemmItem it = items[blockIdx.x];
  auto base = [&](int8_t src) { return src == S_L ? L : (src == S_U ? U : (src == S_INV ? Inv : Part)); };
  const T* __restrict__ A = base(it.asrc) + it.a;
  const T* __restrict__ B = base(it.bsrc) + it.b;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
  const int lrow = tid & (GB - 1), lk = tid >> 7;  // loader: one row, every other k
  const bool arow = lrow < it.m, brow = lrow < it.n;
  T ra[GK / 2], rb[GK / 2];
  auto load = [&](int ks) {
#pragma unroll
    for (int r = 0; r < GK / 2; ++r) {
      const int k = ks + lk + 2 * r;
      const bool kin = k < it.k;
      ra[r] = (arow && kin) ? A[(int64_t)k * it.lda + lrow] : T(0);
      rb[r] = (brow && kin) ? B[(int64_t)k * it.ldb + lrow] : T(0);
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int r = 0; r < GK / 2; ++r) {
      As[buf][lk + 2 * r][lrow] = ra[r];
      Bs[buf][lk + 2 * r][lrow] = rb[r];
    }
  };
  typename Mma<T>::acc_t acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = typename Mma<T>::acc_t{0, 0, 0, 0};
  const int stages = (it.k + GK - 1) / GK;
  load(0);
  store(0);
  __syncthreads();
  for (int s = 0; s < stages; ++s) {
    const int cur = s & 1;
    if (s + 1 < stages) load((s + 1) * GK);
#pragma unroll
    for (int kk = 0; kk < GK; kk += 4) {
      T fj[4], fi[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        fj[b] = Bs[cur][kk + (lane >> 4)][wj + 16 * b + (lane & 15)];
        fi[b] = As[cur][kk + (lane >> 4)][wi + 16 * b + (lane & 15)];
      }
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = Mma<T>::step(fj[bj], fi[bi], acc[bi][bj]);
    }
    if (s + 1 < stages) store(cur ^ 1);
    __syncthreads();
  }
  // acc[bi][bj][v] = C(wi + 16 bi + lane % 16, wj + 16 bj + row(lane, v))
  const int il = lane & 15;
  if (it.mode == G_PART) {
    T* P = Part + it.c;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<T>::row(lane, v);
          P[(int64_t)j * GB + i] = acc[bi][bj][v];
        }
    return;
  }
  T* C = base(it.csrc) + it.c;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<T>::row(lane, v);
        if (i < it.m && j < it.n && i - j + it.diag >= 0) {
          T* p = C + (int64_t)j * it.ldc + i;
          const T x = acc[bi][bj][v];
          if (it.mode == G_SUB) *p -= x;
          else if (it.mode == G_NEG) *p = -x;
          else *p = x;
        }
      }
}

// Sum of the partial tiles of a split product, in order, applied to C as its mode says.
template <typename T>
__global__ __launch_bounds__(256) void k_reduce_parts(const ReduceItem* __restrict__ items, T* L, T* U,
                                                      const T* __restrict__ Part) {
  const ReduceItem it = items[blockIdx.x];
  T* C = (it.csrc == S_L ? L : U) + it.c;
  for (int e = threadIdx.x; e < GB * GB; e += blockDim.x) {
    const int i = e % GB, j = e / GB;
    if (i >= it.m || j >= it.n || i - j + it.diag < 0) continue;
    T s = 0;
    for (int q = 0; q < it.parts; ++q) s += Part[it.part0 + (int64_t)q * GB * GB + e];
    T* p = C + (int64_t)j * it.ldc + i;
    if (it.mode == G_SUB) *p -= s;
    else *p = -s;
  }
}

// ------------------------------------------------------------------------------------------
// The diagonal block of a panel: factored column by column with the reference's pivot test
// (cholesky.rs:1146-1171: a pivot not above 1e-10 |A_jj|, or not finite, fails the column),
// then inverted, Linv = L11^-1, for the rows below (a GEMM). The arithmetic is FP64 whatever
// the storage. Item: f, a = c0, b = slot of Linv (NB x NB, column-major) in Inv.

constexpr double PIVOT_TOL = 1e-10;  // cholesky.rs:48

template <typename T>
__global__ __launch_bounds__(256, 1) void k_potrf_inv(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                      T* __restrict__ L, const double* __restrict__ diag,
                                                      unsigned long long* __restrict__ fail, T* __restrict__ Inv) {
  __shared__ double S[NB][NB + 1];  // lower: L11; strict upper S[j][i] (j < i): Linv(i, j)
  __shared__ double dinv[NB];
  __shared__ int bad;
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(NB, F.w - c0);
  T* base = L + F.lpos + (int64_t)c0 * F.ld + c0;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int i = e % NB, j = e / NB;
    if (i >= j) S[i][j] = (i < pw && j < pw) ? (double)base[(int64_t)j * F.ld + i] : 0.0;
  }
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (int c = 0; c < pw; ++c) {
    if (threadIdx.x == 0) {
      const double d = S[c][c];
      const double a = diag[F.col0 + c0 + c];
      if (!(d > PIVOT_TOL * fabs(a)) || !isfinite(d)) {
        if (!bad) atomicMin(fail, (unsigned long long)(F.col0 + c0 + c));
        bad = 1;
      }
      S[c][c] = sqrt(d);
    }
    __syncthreads();
    const double d = S[c][c];
    for (int i = c + 1 + threadIdx.x; i < pw; i += blockDim.x) S[i][c] /= d;
    __syncthreads();
    const int rest = pw - c - 1;
    for (int e = threadIdx.x; e < rest * rest; e += blockDim.x) {
      const int i = c + 1 + e % rest, j = c + 1 + e / rest;
      if (i >= j) S[i][j] -= S[i][c] * S[j][c];
    }
    __syncthreads();
  }
  // Linv row by row: Linv(i, j) = -(sum_{k=j}^{i-1} L(i, k) Linv(k, j)) / L(i, i), j < i
  for (int i = threadIdx.x; i < pw; i += blockDim.x) dinv[i] = 1.0 / S[i][i];
  __syncthreads();
  for (int i = 1; i < pw; ++i) {
    for (int j = threadIdx.x; j < i; j += blockDim.x) {
      double s = S[i][j] * dinv[j];
      for (int k = j + 1; k < i; ++k) s += S[i][k] * S[j][k];  // S[j][k] = Linv(k, j)
      S[j][i] = -s * dinv[i];
    }
    __syncthreads();
  }
  T* inv = Inv + (int64_t)it.b * NB * NB;
  for (int e = threadIdx.x; e < NB * NB; e += blockDim.x) {
    const int i = e % NB, j = e / NB;
    double v = 0.0;
    if (i < pw && j < pw) v = i == j ? dinv[i] : (i > j ? S[j][i] : 0.0);
    inv[e] = (T)v;
    if (i < pw && j < pw && i >= j) base[(int64_t)j * F.ld + i] = (T)S[i][j];
  }
}

}  // namespace fc
