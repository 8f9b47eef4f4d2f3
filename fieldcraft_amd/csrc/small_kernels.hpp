// Small fronts in one launch per level (gfx950).
//
// The tree's leaf and near-leaf fronts are many and tiny (at hex8 80^3, 14 755 fronts of median
// width 18 and 180 rows at the bottom level): too small for the tiled kernels' 128 x 128 tiles
// and 128-column diagonal blocks, where each of a level's dense steps is a launch of its own.
// The small fronts of a level (no wider than SMALL_W columns, their update matrices at most
// SMALL_SYRK multiply-adds) take one launch between the level's two extend-add launches, in
// which each workgroup does one front's dense steps:
//   1. its diagonal block, factored and inverted by one wave with the wave functions of
//      k_potrf_blk (64-column register layout; FP32 storage: FP32 arithmetic, the block redone
//      in FP64 below the pivot threshold; FP64: the reference's pivot test), L11 and the
//      inverse written out;
//   2. the rows below, X = B Linv^T with FP64 sums (the FP64 factor refines the product once,
//      as the tiled path does);
//   3. its update matrix U = -L21 L21^T (lower triangle), 64 x 64 tiles through LDS, FP64 sums.
// The values end where the tiled path puts them (L, Inv, U); the FP32-storage forms write no
// BF16 copies for them, since no product of the tiled path reads a small front.
#pragma once

#include "gemm_kernels.hpp"

namespace fc {

constexpr int SMALL_W = 64;            // widest small front (one wave's diagonal block)
constexpr int SMALL_MU = 768;          // most rows below a small front's columns
constexpr double SMALL_SYRK = 2.0e6;   // most multiply-adds of a small front's update matrix

// The rows below the diagonal block, X = B Linv^T (thread t takes rows w + t + 256 k), from the
// block's L11 and inverse in LDS: L(k, c) = S[k][c] for k >= c, Linv(c, k) = S[k][c] for k < c
// and xd[c] for k == c. Each x_c is a dot product of its own (no dependent chain across the
// columns) in FP64, stored as it is made; the FP64 factor then refines the product once, as
// the tiled path does (an explicit-inverse product alone is not backward stable): the row's B
// stays in registers, r = B - X L11^T column by column of X (read back), X += r Linv^T.
// Through a buffer resource (the front's L part is far below 2 GiB): a column's offset is a
// scalar and the lane's part one VGPR. WB >= w columns (a multiple of 16); the columns past w
// take the lane offset 2^31, beyond the resource's 2^31 - 1 bytes, so their loads return 0 and
// their stores are dropped by the buffer unit's range check, with no branch (branches around
// them made the compiler keep the block's LDS reads live across the row).
template <typename T, int WB>
__device__ __forceinline__ void small_rows(const FrontDev& F, T* __restrict__ base, const double (*S)[NB + 1],
                                           const double* xd, int tid) {
  const int w = F.w;
  const auto rs = raw_rsrc(base);
  const uint32_t ldb = (uint32_t)F.ld * sizeof(T);
  constexpr uint32_t OUT = 0x80000000u;
  for (int i = w + tid; i < F.m; i += 256) {
    const uint32_t lo = (uint32_t)i * sizeof(T);
    double b[WB];
#pragma unroll
    for (int k = 0; k < WB; ++k) b[k] = (double)buf_load<T>(rs, k < w ? lo : OUT, k * ldb);
#pragma unroll
    for (int c = 0; c < WB; ++c) {  // x_c = sum_{k <= c} b_k Linv(c, k)
      double t = b[c] * xd[c];
#pragma unroll
      for (int k = 0; k < c; ++k) {
        t = fma(b[k], S[k][c], t);
        if ((k & 15) == 15) __builtin_amdgcn_sched_barrier(0);  // LDS reads in flight capped at 16
      }
      buf_store<T>(rs, (T)t, c < w ? lo : OUT, c * ldb);
      __builtin_amdgcn_sched_barrier(0);
    }
    if constexpr (sizeof(T) == 8) {
#pragma unroll
      for (int c = 0; c < WB; ++c) {  // r_k = b_k - sum_{c <= k} x_c L(k, c), by columns of X
        const double xc = buf_load<T>(rs, c < w ? lo : OUT, c * ldb);
#pragma unroll
        for (int k = c; k < WB; ++k) {
          b[k] = fma(-xc, S[k][c], b[k]);
          if ((k & 15) == 15) __builtin_amdgcn_sched_barrier(0);
        }
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int c = WB - 1; c >= 0; --c) {  // corrections d_c = sum_{k <= c} r_k Linv(c, k), in place
        double t = b[c] * xd[c];
#pragma unroll
        for (int k = 0; k < c; ++k) {
          t = fma(b[k], S[k][c], t);
          if ((k & 15) == 15) __builtin_amdgcn_sched_barrier(0);
        }
        b[c] = t;
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int c = 0; c < WB; ++c) {
        const uint32_t o = c < w ? lo : OUT;
        buf_store<T>(rs, buf_load<T>(rs, o, c * ldb) + b[c], o, c * ldb);
      }
    }
  }
}

// Workgroup b takes the front items[b].f (the launch's items are costliest first, so the
// hardware's dispatch in order balances the CUs).
template <typename T>
__global__ __launch_bounds__(256, 2) void k_small_fronts(const Item* __restrict__ items,
                                                         const FrontDev* __restrict__ fronts, T* __restrict__ L,
                                                         T* __restrict__ U, const double* __restrict__ diag,
                                                         unsigned long long* __restrict__ fail, T* __restrict__ Inv) {
  // the diagonal block in wave_potrf_inv's layout, then the update matrix's staging tiles
  __shared__ double S[SMALL_W][NB + 1];
  __shared__ double xd[NB];
  __shared__ int redo;
  static_assert(2 * SMALL_W * SMALL_W <= SMALL_W * (NB + 1), "two 64 x 64 tiles fit the block's LDS");
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int w = F.w, mu = F.m - F.w;
  T* base = L + F.lpos;
  const double* dg = diag + F.col0;
  const int64_t col = F.col0;
  // 1. the diagonal block: every entry of S's first 64 x 64 written (zeros outside the lower
  // triangle of the w x w block: the row products read them past w), and xd cleared
  // (thread t: row t % 64, columns t / 64 + 4 q; the 16 loads issued before any is used, those
  // outside the triangle reading 0 through the buffer unit's range check, as in small_rows)
  const auto rs = raw_rsrc(base);
  const uint32_t ldb = (uint32_t)F.ld * sizeof(T);
  auto load = [&]() {
    const int i = tid & (SMALL_W - 1), j0 = tid >> 6;
    T v[SMALL_W / 4];
#pragma unroll
    for (int q = 0; q < SMALL_W / 4; ++q) {
      const int j = j0 + 4 * q;
      v[q] = buf_load<T>(rs, (i < w && j <= i) ? (uint32_t)(i * sizeof(T)) : 0x80000000u, j * ldb);
    }
#pragma unroll
    for (int q = 0; q < SMALL_W / 4; ++q) S[i][j0 + 4 * q] = (double)v[q];
    if (tid < NB) xd[tid] = 0.0;
  };
  load();
  __syncthreads();
  bool bad = false;
  if constexpr (sizeof(T) == 4) {
    bool fb = false;
    if (wave == 0) wave_potrf_inv32(S, xd, 0, w, lane, dg, fb);
    if (wave == 0 && lane == 0) redo = fb ? 1 : 0;
    __syncthreads();
    if (redo) {  // uniform
      __syncthreads();  // every thread has read redo
      load();
      __syncthreads();
      if (wave == 0) wave_potrf_inv<false>(S, xd, 0, w, lane, dg, col, fail, bad);
      __syncthreads();
    }
  } else {
    if (wave == 0) wave_potrf_inv<true>(S, xd, 0, w, lane, dg, col, fail, bad);
    __syncthreads();
  }
  // L11 out, and the inverse (column-major, ld = w, zeros above the diagonal):
  // S[i][j] = L(i, j) for i >= j, S[k][i] = Linv(i, k) for i > k, xd[i] = Linv(i, i)
  for (int e = tid; e < w * w; e += blockDim.x) {
    const int i = e % w, j = e / w;
    if (j <= i) base[(int64_t)j * F.ld + i] = (T)S[i][j];
    Inv[F.inv0 + (int64_t)j * w + i] = (T)(i > j ? S[j][i] : (i == j ? xd[i] : 0.0));
  }
  // 2. the rows below
  switch ((w + 15) / 16) {
    case 1: small_rows<T, 16>(F, base, S, xd, tid); break;
    case 2: small_rows<T, 32>(F, base, S, xd, tid); break;
    case 3: small_rows<T, 48>(F, base, S, xd, tid); break;
    default: small_rows<T, 64>(F, base, S, xd, tid); break;
  }
  __syncthreads();  // every row below final before the products read them
  // 3. U = -L21 L21^T by lower 64 x 64 tiles. Thread (tr, tc) = (t % 16, t / 16) sums rows
  // i0 + tr + 16 a and columns j0 + tc + 16 b (a, b < 4) over the w columns in order, from
  // LDS: A(k, r) and B(k, r) return L21(i0 + r, k) and L21(j0 + r, k) (rows past mu: any
  // in-bounds value; those sums are not stored).
  const int tr = tid & 15, tc = tid >> 4;
  const T* l21 = base + w;  // L21(i, k) = l21[k ld + i]
  auto tile = [&](int i0, int j0, auto&& A, auto&& B) {
    double acc[4][4] = {};
    for (int k = 0; k < w; ++k) {
      double av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = A(k, tr + 16 * a);
#pragma unroll
      for (int bb = 0; bb < 4; ++bb) bv[bb] = B(k, tc + 16 * bb);
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int bb = 0; bb < 4; ++bb) acc[a][bb] = fma(av[a], bv[bb], acc[a][bb]);
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int bb = 0; bb < 4; ++bb) {
        const int i = i0 + tr + 16 * a, j = j0 + tc + 16 * bb;
        if (i < mu && j <= i) U[F.upos + (int64_t)j * mu + i] = (T)(-acc[a][bb]);
      }
  };
  constexpr int STAGE = (int)(sizeof(S) / sizeof(T));  // entries of T the block's LDS holds
  if (mu * w <= STAGE) {
    // the whole of L21 staged at once (as T, column-major with ld mu): one round of loads
    T* Ls = reinterpret_cast<T*>(&S[0][0]);
    const int n = mu * w;
    for (int e0 = 0; e0 < n; e0 += 256 * 8) {
      T v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int e = min(e0 + tid + 256 * q, n - 1), k = e / mu;
        v[q] = l21[(int64_t)k * F.ld + (e - k * mu)];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q)
        if (e0 + tid + 256 * q < n) Ls[e0 + tid + 256 * q] = v[q];
    }
    __syncthreads();
    for (int j0 = 0; j0 < mu; j0 += SMALL_W)
      for (int i0 = j0; i0 < mu; i0 += SMALL_W)
        tile(
            i0, j0, [&](int k, int r) { return (double)Ls[k * mu + min(i0 + r, mu - 1)]; },
            [&](int k, int r) { return (double)Ls[k * mu + min(j0 + r, mu - 1)]; });
    return;
  }
  // Otherwise by block columns: the column block's rows j0.. are staged once (Bs) and serve as
  // the diagonal tile's rows too; each tile below stages its rows (As), as doubles [k][r].
  double* As = &S[0][0];
  double* Bs = As + SMALL_W * SMALL_W;
  // rows r0 + t % 64 of columns t / 64 + 4 q (q < 16, those below w) into D[k][r], rows past mu
  // as 0 (read through the range check); the loads issued before any is used
  auto stage = [&](double* D, int r0) {
    const int r = tid & (SMALL_W - 1), k0 = tid >> 6;
    T v[SMALL_W / 4];
#pragma unroll
    for (int q = 0; q < SMALL_W / 4; ++q) {
      const int k = k0 + 4 * q;
      v[q] = buf_load<T>(rs, (k < w && r0 + r < mu) ? (uint32_t)((w + r0 + r) * sizeof(T)) : 0x80000000u, k * ldb);
    }
#pragma unroll
    for (int q = 0; q < SMALL_W / 4; ++q)
      if (k0 + 4 * q < w) D[(k0 + 4 * q) * SMALL_W + r] = (double)v[q];
  };
  for (int j0 = 0; j0 < mu; j0 += SMALL_W) {
    stage(Bs, j0);
    for (int i0 = j0; i0 < mu; i0 += SMALL_W) {
      if (i0 > j0) stage(As, i0);
      __syncthreads();
      const double* Ap = i0 > j0 ? As : Bs;
      tile(
          i0, j0, [&](int k, int r) { return Ap[k * SMALL_W + r]; }, [&](int k, int r) { return Bs[k * SMALL_W + r]; });
      __syncthreads();  // the tile's reads done before the next staging
    }
  }
}

}  // namespace fc
