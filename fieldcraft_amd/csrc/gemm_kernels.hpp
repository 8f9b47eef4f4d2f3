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
  int32_t front;          // the front both operands come from (its scale in the mixed form)
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

// ------------------------------------------------------------------------------------------
// The Schur products of the mixed form: FP32 operands in HBM, multiplied as BF16 on the
// 16x16x32 BF16 matrix cores with FP32 accumulation. Both operands of a product come from one
// front and are scaled by that front's power of two, 2^-fexp[front], which bounds them by 1
// (|L_ij| <= sqrt(A_ii) for an SPD matrix); the result is scaled back exactly. With TERMS = 3
// each scaled operand x is split into hi = bf16(x) and lo = bf16(x - hi) and the product is
// hi*hi + hi*lo + lo*hi, which carries about 16 significant bits; TERMS = 1 keeps hi*hi.
//
// Same tile and item layout as k_gemm (128 x 128, 4 waves of 64 x 64, modes and split-K
// pieces alike). k is staged BK = 32 at a time: a loader thread reads 16 consecutive k of one
// row (coalesced across threads along the rows), converts and writes them to LDS as k-major
// rows, from which each lane reads its 8 k of a fragment as one 16-byte load.
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
constexpr int BK = 32;
constexpr int BROW = BK + 8;  // LDS row stride (80 bytes: 16-byte aligned, rows spread over the banks)

template <int TERMS>
__global__ __launch_bounds__(256, 2) void k_gemm_bf16(const GemmItem* __restrict__ items, float* L, float* U,
                                                      float* Part, const int32_t* __restrict__ fexp) {
  static_assert(TERMS == 1 || TERMS == 3, "one or three BF16 products");
  constexpr int NS = TERMS == 3 ? 2 : 1;  // hi (and lo) copies of each operand
  __shared__ __attribute__((aligned(16))) __bf16 As[2][NS][GB][BROW];
  __shared__ __attribute__((aligned(16))) __bf16 Bs[2][NS][GB][BROW];
  const GemmItem it = items[blockIdx.x];
  const int e = fexp[it.front];
  const float sc = ldexpf(1.0f, -e);  // exact power of two
  auto base = [&](int8_t src) { return src == S_L ? L : (src == S_U ? U : Part); };
  const float* __restrict__ A = base(it.asrc) + it.a;
  const float* __restrict__ B = base(it.bsrc) + it.b;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
  const int lrow = tid & (GB - 1), lk = (tid >> 7) * 16;  // loader: one row, 16 consecutive k
  const bool arow = lrow < it.m, brow = lrow < it.n;
  float ra[16], rb[16];
  auto load = [&](int ks) {
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int k = ks + lk + r;
      const bool kin = k < it.k;
      ra[r] = (arow && kin) ? A[(int64_t)k * it.lda + lrow] : 0.0f;
      rb[r] = (brow && kin) ? B[(int64_t)k * it.ldb + lrow] : 0.0f;
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      v8bf ah, bh, al, bl;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float xa = ra[8 * h + q] * sc, xb = rb[8 * h + q] * sc;
        ah[q] = (__bf16)xa;
        bh[q] = (__bf16)xb;
        if (TERMS == 3) {
          al[q] = (__bf16)(xa - (float)ah[q]);
          bl[q] = (__bf16)(xb - (float)bh[q]);
        }
      }
      *(v8bf*)&As[buf][0][lrow][lk + 8 * h] = ah;
      *(v8bf*)&Bs[buf][0][lrow][lk + 8 * h] = bh;
      if (TERMS == 3) {
        *(v8bf*)&As[buf][NS - 1][lrow][lk + 8 * h] = al;
        *(v8bf*)&Bs[buf][NS - 1][lrow][lk + 8 * h] = bl;
      }
    }
  };
  v4f acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = v4f{0, 0, 0, 0};
  const int stages = (it.k + BK - 1) / BK;
  const int fr = lane & 15, fk = 8 * (lane >> 4);  // a lane's fragment row and first k
  load(0);
  store(0);
  __syncthreads();
  for (int s = 0; s < stages; ++s) {
    const int cur = s & 1;
    if (s + 1 < stages) load((s + 1) * BK);
    v8bf fj[4], fi[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      fj[b] = *(const v8bf*)&Bs[cur][0][wj + 16 * b + fr][fk];
      fi[b] = *(const v8bf*)&As[cur][0][wi + 16 * b + fr][fk];
    }
    if (TERMS == 3) {
      v8bf gj[4], gi[4];
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        gj[b] = *(const v8bf*)&Bs[cur][NS - 1][wj + 16 * b + fr][fk];
        gi[b] = *(const v8bf*)&As[cur][NS - 1][wi + 16 * b + fr][fk];
      }
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj) {
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gj[bj], fi[bi], acc[bi][bj], 0, 0, 0);
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fj[bj], gi[bi], acc[bi][bj], 0, 0, 0);
        }
    }
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fj[bj], fi[bi], acc[bi][bj], 0, 0, 0);
    if (s + 1 < stages) store(cur ^ 1);
    __syncthreads();
  }
  // acc[bi][bj][v] = C(wi + 16 bi + lane % 16, wj + 16 bj + 4 (lane / 16) + v) / sc^2
  const float unscale = ldexpf(1.0f, 2 * e);
  const int il = lane & 15;
  if (it.mode == G_PART) {
    float* P = Part + it.c;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<float>::row(lane, v);
          P[(int64_t)j * GB + i] = acc[bi][bj][v] * unscale;
        }
    return;
  }
  float* C = base(it.csrc) + it.c;
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<float>::row(lane, v);
        if (i < it.m && j < it.n && i - j + it.diag >= 0) {
          float* p = C + (int64_t)j * it.ldc + i;
          const float x = acc[bi][bj][v] * unscale;
          if (it.mode == G_SUB) *p -= x;
          else if (it.mode == G_NEG) *p = -x;
          else *p = x;
        }
      }
}

// Sum of the partial tiles of a split product, in order, applied to C as its mode says. Each
// tile is shared by RSPLIT workgroups (GB * GB / RSPLIT entries each).
constexpr int RSPLIT = 16;

template <typename T>
__global__ __launch_bounds__(256) void k_reduce_parts(const ReduceItem* __restrict__ items, T* L, T* U,
                                                      const T* __restrict__ Part) {
  const ReduceItem it = items[blockIdx.x / RSPLIT];
  T* C = (it.csrc == S_L ? L : U) + it.c;
  constexpr int CHUNK = GB * GB / RSPLIT;
  const int e0 = (blockIdx.x % RSPLIT) * CHUNK;
  for (int e = e0 + threadIdx.x; e < e0 + CHUNK; e += blockDim.x) {
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

// K: the widest block the variant takes (32, 64 or NB); the host sends each panel to the
// narrowest variant that holds it. One barrier per column: column c's trailing update reads
// column c unscaled (times 1 / d_c^2) while column c - 1 is scaled in the same phase.
template <typename T, int K>
__global__ __launch_bounds__(K == NB ? 256 : 64 * (K / 32), 1) void k_potrf_inv(
    const Item* __restrict__ items, const FrontDev* __restrict__ fronts, T* __restrict__ L,
    const double* __restrict__ diag, unsigned long long* __restrict__ fail, T* __restrict__ Inv) {
  __shared__ double S[K][K + 1];  // lower: L11; strict upper S[j][i] (j < i): Linv(i, j)
  __shared__ double dv[K];        // the pivots' square roots, then their inverses
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(K, F.w - c0);
  T* base = L + F.lpos + (int64_t)c0 * F.ld + c0;
  for (int e = threadIdx.x; e < K * K; e += blockDim.x) {
    const int i = e % K, j = e / K;
    if (i >= j) S[i][j] = (i < pw && j < pw) ? (double)base[(int64_t)j * F.ld + i] : 0.0;
  }
  __syncthreads();
  bool bad = false;
  for (int c = 0; c <= pw; ++c) {
    if (c > 0) {  // scale the previous column
      const double d = dv[c - 1];
      for (int i = c + threadIdx.x; i < pw; i += blockDim.x) S[i][c - 1] /= d;
      if (threadIdx.x == 0) S[c - 1][c - 1] = d;
    }
    if (c < pw) {
      const double p = S[c][c];
      const double d = sqrt(p);
      if (threadIdx.x == 0) {
        dv[c] = d;
        const double a = diag[F.col0 + c0 + c];
        if (!bad && (!(p > PIVOT_TOL * fabs(a)) || !isfinite(p))) {
          atomicMin(fail, (unsigned long long)(F.col0 + c0 + c));
          bad = true;
        }
      }
      const double q = 1.0 / (d * d);
      const int rest = pw - c - 1;
      for (int e = threadIdx.x; e < rest * rest; e += blockDim.x) {
        const int i = c + 1 + e % rest, j = c + 1 + e / rest;
        if (i >= j) S[i][j] -= S[i][c] * S[j][c] * q;
      }
    }
    __syncthreads();
  }
  // Linv row by row: Linv(i, j) = -(sum_{k=j}^{i-1} L(i, k) Linv(k, j)) / L(i, i), j < i
  for (int i = threadIdx.x; i < pw; i += blockDim.x) dv[i] = 1.0 / S[i][i];
  __syncthreads();
  for (int i = 1; i < pw; ++i) {
    for (int j = threadIdx.x; j < i; j += blockDim.x) {
      double s = S[i][j] * dv[j];
      for (int k = j + 1; k < i; ++k) s += S[i][k] * S[j][k];  // S[j][k] = Linv(k, j)
      S[j][i] = -s * dv[i];
    }
    __syncthreads();
  }
  T* inv = Inv + F.inv0 + (int64_t)(c0 / NB) * NB * NB;  // ld = pw
  for (int e = threadIdx.x; e < K * K; e += blockDim.x) {
    const int i = e % K, j = e / K;
    if (i < pw && j < pw) {
      inv[(int64_t)j * pw + i] = (T)(i == j ? dv[i] : (i > j ? S[j][i] : 0.0));
      if (i >= j) base[(int64_t)j * F.ld + i] = (T)S[i][j];
    }
  }
}

// The same for blocks up to NB = 128 wide, with the block in registers: 256 threads as a
// 16 x 16 grid, thread (ti, tj) holding entries (ti + 16 a, tj + 16 b), a, b < 8. A column
// step publishes only its pivot and its column through LDS, and every thread updates its own
// 64 entries, so the rank-1 updates cost no LDS traffic. The inverse is formed the same way
// (right-looking on L X = I) with L read from LDS. Two barriers per column in each part.
template <typename T>
__global__ __launch_bounds__(256, 1) void k_potrf_inv128(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                         T* __restrict__ L, const double* __restrict__ diag,
                                                         unsigned long long* __restrict__ fail, T* __restrict__ Inv) {
  constexpr int G = 16, R = NB / G;
  __shared__ double Ls[NB][NB + 1];
  __shared__ double col[NB], row[NB], piv[NB + 1];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(NB, F.w - c0);
  T* base = L + F.lpos + (int64_t)c0 * F.ld + c0;
  const int ti = threadIdx.x % G, tj = threadIdx.x / G;
  double s[R][R];
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int i = ti + G * a, j = tj + G * b;
      s[a][b] = (i < pw && j < pw && i >= j) ? (double)base[(int64_t)j * F.ld + i] : (i == j ? 1.0 : 0.0);
    }
  if (threadIdx.x == 0) piv[0] = s[0][0];
  __syncthreads();
  bool bad = false;
#pragma unroll
  for (int cb = 0; cb < R; ++cb) {
    for (int cc = 0; cc < G; ++cc) {
      const int c = G * cb + cc;
      if (c >= pw) break;
      const double p = piv[c];
      const double d = sqrt(p);
      if (threadIdx.x == 0 && !bad) {
        const double a0 = diag[F.col0 + c0 + c];
        if (!(p > PIVOT_TOL * fabs(a0)) || !isfinite(p)) {
          atomicMin(fail, (unsigned long long)(F.col0 + c0 + c));
          bad = true;
        }
      }
      if (tj == cc) {  // the owners of column c scale it and publish it
#pragma unroll
        for (int a = 0; a < R; ++a) {
          const int i = ti + G * a;
          if (i > c) {
            s[a][cb] /= d;
            col[i] = s[a][cb];
          } else if (i == c) {
            s[a][cb] = d;
          }
        }
      }
      __syncthreads();
      double li[R], lj[R];
#pragma unroll
      for (int a = 0; a < R; ++a) li[a] = (ti + G * a > c) ? col[ti + G * a] : 0.0;
#pragma unroll
      for (int b = 0; b < R; ++b) lj[b] = (tj + G * b > c) ? col[tj + G * b] : 0.0;
#pragma unroll
      for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b)
          if (ti + G * a >= tj + G * b) s[a][b] -= li[a] * lj[b];
      // the next pivot, from its owner
      const int n1 = c + 1;
      if (n1 < pw && ti == n1 % G && tj == n1 % G) piv[n1] = (cc + 1 < G) ? s[cb][cb] : s[(cb + 1) % R][(cb + 1) % R];
      __syncthreads();
    }
  }
  // L11 to LDS (and back to the front), then X = L11^-1 in registers
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int i = ti + G * a, j = tj + G * b;
      if (i >= j) {
        Ls[i][j] = s[a][b];
        if (i < pw && j < pw) base[(int64_t)j * F.ld + i] = (T)s[a][b];
      }
    }
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) s[a][b] = (ti + G * a == tj + G * b) ? 1.0 : 0.0;  // X = I
  __syncthreads();
#pragma unroll
  for (int kb = 0; kb < R; ++kb) {
    for (int kc = 0; kc < G; ++kc) {
      const int k = G * kb + kc;
      if (k >= pw) break;
      const double dk = Ls[k][k];
      if (ti == kc) {  // the owners of row k of X finish it and publish it
#pragma unroll
        for (int b = 0; b < R; ++b) {
          s[kb][b] /= dk;
          row[tj + G * b] = s[kb][b];
        }
      }
      __syncthreads();
      double lk[R], xk[R];
#pragma unroll
      for (int a = 0; a < R; ++a) lk[a] = (ti + G * a > k) ? Ls[ti + G * a][k] : 0.0;
#pragma unroll
      for (int b = 0; b < R; ++b) xk[b] = row[tj + G * b];
#pragma unroll
      for (int a = 0; a < R; ++a)
#pragma unroll
        for (int b = 0; b < R; ++b) s[a][b] -= lk[a] * xk[b];
      __syncthreads();
    }
  }
  T* inv = Inv + F.inv0 + (int64_t)(c0 / NB) * NB * NB;  // ld = pw
#pragma unroll
  for (int a = 0; a < R; ++a)
#pragma unroll
    for (int b = 0; b < R; ++b) {
      const int i = ti + G * a, j = tj + G * b;
      if (i < pw && j < pw) inv[(int64_t)j * pw + i] = (T)(i >= j ? s[a][b] : 0.0);
    }
}

}  // namespace fc
