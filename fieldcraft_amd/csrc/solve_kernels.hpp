// Supernodal triangular solves over the fronts (the reference's solve_group,
// cholesky.rs:904-1037, reorganised for levels of the front tree) and the FP64 vector steps of
// refinement.
//
// Forward, leaves first: a front gathers its children's contributions (each child hands up the
// products of its rows below, u_c = v_c,below - L21_c y_c), solves L11 y = v_own and hands up
// u = v_below - L21 y. Backward, root first: x_own = L11^-T (y - L21^T x_below), the rows below
// read from the solution. Contributions travel up the tree instead of being scattered into
// ancestors, so no two workgroups write one entry and no atomics are used.
//
// A front is solved by blocks of SNB = 128 columns, its factorization's panels, whose diagonal
// blocks' inverses the factorization keeps (column-major, ld = the block's width, zeros above
// the diagonal), so a block's triangle is a small matrix-vector product. One launch per block
// each way; the workgroup that finishes the rows (forward) or columns (backward) of the next
// block also forms that block's solution, so the block solutions never wait for a launch of
// their own. Forward reads v from x and writes y to a second vector; backward reads z from it
// and writes x. The right-hand sides are columns (postorder numbering, leading dimension ldx),
// R of them held at a time (1, or 16 for blocks of load cases); every entry of L is read once
// per R right-hand sides, with independent loads kept in flight.
#pragma once

#include <type_traits>

#include "device.hpp"

namespace fc {

constexpr int SNB = 128;      // columns of a solve block (the factorization's panel width)
constexpr int SROWS = 256;    // rows of a forward workgroup (one right-hand side)
constexpr int SROWS_M = 128;  // ... of several: each row's right-hand sides split between two threads
constexpr int SCOLS = 64;     // columns of a below-the-front workgroup (backward)
constexpr int BCOLS = 128;    // columns of a backward block-step workgroup
constexpr int RC = 16;        // right-hand sides of the multi-column kernels
constexpr int RSTEP = 8;      // right-hand sides of a backward block-step workgroup: a step's launch
                              // holds one copy of its items per group of them, run side by side
constexpr int MAX_RHS = 64;   // right-hand sides one solve pass handles (the host splits larger blocks)
// A block solution not yet formed (the chained one-right-hand-side solve below): a signalling
// NaN, which no arithmetic result is.
constexpr uint64_t PENDING = 0x7FF0DEAD0000BEEFull;

// Where a thread's loads would otherwise wait on branches, they are made in batches from clamped
// (always in-bounds) addresses and the out-of-range values are selected away afterwards, so a
// batch is issued back to back and costs one round of memory latency.

// y[b] = sum_k Inv(b, k) v[k] for b < pw (Inv column-major, ld = pw, zero above the diagonal):
// thread b of the first 128, loads coalesced along b and independent along k. V: LDS, [k][r].
// Several right-hand sides: thread (b, hh) of the 256 takes half hh of them.
template <typename T, int R>
__device__ __forceinline__ void inv_apply(const T* __restrict__ inv, int pw, const double (*V)[R + 1], int nr,
                                          double* __restrict__ out, int64_t ldx) {
  constexpr int RH = R > 1 ? R / 2 : 1;  // right-hand sides a thread sums
  const int b = R > 1 ? (int)(threadIdx.x & (SNB - 1)) : (int)threadIdx.x, r0 = R > 1 ? (int)(threadIdx.x >> 7) * RH : 0;
  if (b >= pw) return;
  double s[RH];
#pragma unroll
  for (int r = 0; r < RH; ++r) s[r] = 0.0;
  if (R == 1 && pw == SNB) {  // a full block, one right-hand side: the whole row in flight at once
    constexpr int KB = sizeof(T) == 4 ? SNB : SNB / 2;
    const __amdgpu_buffer_rsrc_t rs = raw_rsrc(inv);
#pragma unroll
    for (int k0 = 0; k0 < SNB; k0 += KB) {
      T a[KB];
#pragma unroll
      for (int q = 0; q < KB; ++q) a[q] = buf_load<T>(rs, b * sizeof(T), (k0 + q) * SNB * sizeof(T));
#pragma unroll
      for (int q = 0; q < KB; ++q) s[0] = fma((double)a[q], V[k0 + q][0], s[0]);
    }
  } else {  // batches of 32 rows of Inv, each batch's loads in flight together
    const __amdgpu_buffer_rsrc_t rs = raw_rsrc(inv);
    for (int k0 = 0; k0 < pw; k0 += 32) {
      T a[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) a[q] = buf_load<T>(rs, b * sizeof(T), (uint32_t)min(k0 + q, pw - 1) * pw * sizeof(T));
#pragma unroll
      for (int q = 0; q < 32; ++q)
        if (k0 + q < pw)
#pragma unroll
          for (int r = 0; r < RH; ++r) s[r] = fma((double)a[q], V[k0 + q][r0 + r], s[r]);
    }
  }
  for (int r = 0; r < RH && r0 + r < nr; ++r) out[(int64_t)(r0 + r) * ldx + b] = s[r];
}

// x[b] = sum_k Inv(k, b) z[k] (the transpose): thread (b, h) sums the half h of column b of Inv
// (contiguous), then the halves are added in a fixed order. Z: LDS, [k][r]; H: LDS scratch.
// Several right-hand sides are summed eight at a time, with eight loads of Inv in flight (each
// one's sum in the same order as alone): sixteen at once, or sixteen loads, spill registers.
template <typename T, int R>
__device__ __forceinline__ void inv_apply_t(const T* __restrict__ inv, int pw, const double (*Z)[R + 1], int nr,
                                            double* __restrict__ out, int64_t ldx, double (*H)[SNB][R + 1]) {
  const int b = threadIdx.x & (SNB - 1), h = threadIdx.x >> 7, bb = min(b, pw - 1);
  const __amdgpu_buffer_rsrc_t rs = raw_rsrc(inv);
  const uint32_t co = bb * pw * sizeof(T);  // column bb of Inv
  constexpr int RG = R < 8 ? R : 8;         // right-hand sides summed at a time
  static_assert(R % RG == 0, "whole groups of right-hand sides");
  constexpr int KB = R == 1 ? 64 : 8;       // loads per batch (one batch for a single right-hand side)
#pragma unroll 1
  for (int r0 = 0; r0 < R; r0 += RG) {
    double s[RG];
#pragma unroll
    for (int r = 0; r < RG; ++r) s[r] = 0.0;
#pragma unroll 1
    for (int k0 = 0; k0 < 64; k0 += KB) {
      double a[KB];
#pragma unroll
      for (int q = 0; q < KB; ++q) a[q] = (double)buf_load<T>(rs, co, min(64 * h + k0 + q, pw - 1) * sizeof(T));
#pragma unroll
      for (int q = 0; q < KB; ++q) {
        const int k = 64 * h + k0 + q;
        const bool in = k < pw;
        // Z past the block's width is not written by every caller: select it away (a 0 weight on
        // a NaN left in LDS by an earlier kernel would still give a NaN)
#pragma unroll
        for (int r = 0; r < RG; ++r) s[r] = fma(in ? a[q] : 0.0, in ? Z[min(k, SNB - 1)][r0 + r] : 0.0, s[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < RG; ++r) H[h][b][r0 + r] = s[r];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < pw * nr; e += blockDim.x) {
    const int c = e % pw, r = e / pw;
    out[(int64_t)r * ldx + c] = H[0][c][r] + H[1][c][r];
  }
}

// Forward, per front (item f): zero the front's contribution block, add its children's into it
// and into v, then y of the front's first block = Inv_0 v_0.
template <typename T, int R>
__global__ __launch_bounds__(256, 2) void k_fwd_gather(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                    const int32_t* __restrict__ kids, const int32_t* __restrict__ rel,
                                                    const T* __restrict__ Inv, double* v, double* __restrict__ y,
                                                    int64_t ldx, int nrhs, double* u, const int64_t* __restrict__ uoff,
                                                    int mark) {
  __shared__ double Vb[SNB][R + 1];
  const int f = items[blockIdx.x].f;
  const FrontDev F = fronts[f];
  const int w = F.w, mu = F.m - F.w, pw = min(SNB, w);
  if (mark)  // one right-hand side, chained: y of the later blocks still to come (k_fwd_front)
    for (int e = pw + threadIdx.x; e < w; e += blockDim.x)
      reinterpret_cast<uint64_t*>(y)[F.col0 + e] = PENDING;
  double* uf = u + uoff[f] * nrhs;
  for (int e = threadIdx.x; e < mu * nrhs; e += blockDim.x) uf[e] = 0.0;
  __syncthreads();
  for (int q = F.kid0; q < F.kid1; ++q) {
    const int c = kids[q];
    const FrontDev C = fronts[c];
    const int cmu = C.m - C.w;
    const double* uc = u + uoff[c] * nrhs;
    const int32_t* rc = rel + C.rel_off;
    // a child's rows land on distinct rows of the front (rel rises): each thread's row is read
    // for every column before it is written
    for (int i = threadIdx.x; i < cmu; i += blockDim.x) {
      const int p = rc[i];
      double* t = p < w ? v + F.col0 + p : uf + p - w;
      const int64_t ld = p < w ? ldx : mu;
      for (int r0 = 0; r0 < nrhs; r0 += R) {
        double add[R], old[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const int rr = min(r0 + r, nrhs - 1);
          add[r] = uc[(int64_t)rr * cmu + i];
          old[r] = t[(int64_t)rr * ld];
        }
        for (int r = 0; r < R && r0 + r < nrhs; ++r) t[(int64_t)(r0 + r) * ld] = old[r] + add[r];
      }
    }
    __syncthreads();
  }
  for (int q0 = 0; q0 < nrhs; q0 += R) {
    const int nr = min(R, nrhs - q0);
    if (threadIdx.x < pw) {
      double vv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) vv[r] = v[(int64_t)(q0 + min(r, nr - 1)) * ldx + F.col0 + threadIdx.x];
#pragma unroll
      for (int r = 0; r < R; ++r) Vb[threadIdx.x][r] = vv[r];
    }
    __syncthreads();
    inv_apply<T, R>(Inv + F.inv0, pw, Vb, nr, y + (int64_t)q0 * ldx + F.col0, ldx);
    __syncthreads();
  }
}

// Forward step of block [c0, c1) (y_b already in y): rows [r0, r0 + SROWS) at or below c1 lose
// L(r, block) y_b, in v for the front's own columns, in its contribution u below them. The item
// with c = 1 (r0 = c1 and a next block [c1, c2) in the front) then forms y of the next block
// from the rows it has just finished. Item: f, a = c0, b = r0, c.
template <typename T, int R>
__global__ __launch_bounds__(256, 2) void k_fwd_step(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                  const T* __restrict__ L, const T* __restrict__ Inv, double* v,
                                                  double* __restrict__ y, int64_t ldx, int nrhs, double* __restrict__ u,
                                                  const int64_t* __restrict__ uoff, int nitems) {
  __shared__ double Yb[SNB][R + 1];
  __shared__ double Vn[SNB][R + 1];
  // workgroup b takes item b % nitems for the right-hand-side groups b / nitems + k gridDim / nitems
  const int grp = blockIdx.x / nitems, ngrp = gridDim.x / nitems;
  const Item it = items[blockIdx.x - grp * nitems];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(SNB, F.w - c0), c1 = c0 + pw, r0 = it.b;
  const int pn = min(SNB, F.w - c1);  // width of the next block (designated item)
  const T* lf = L + F.lpos + (int64_t)c0 * F.ld;
  const int mu = F.m - F.w;
  double* uf = u + uoff[it.f] * nrhs;
  // several right-hand sides: SROWS_M rows, thread (row, hh) taking half hh of them
  constexpr int RH = R > 1 ? R / 2 : 1, RPI = R > 1 ? SROWS_M : SROWS;
  const int rh0 = R > 1 ? (int)(threadIdx.x / RPI) * RH : 0;
  const int i = r0 + (int)(threadIdx.x % RPI);
  const bool row = i < F.m;
  const __amdgpu_buffer_rsrc_t rs = raw_rsrc(lf);
  const uint32_t lo = min(i, F.m - 1) * sizeof(T);  // rows past the front load its last row (not stored)
  double* p = i < F.w ? v + F.col0 + i : uf + i - F.w;  // the row's entry of v (or of u below the front)
  const int64_t ldp = i < F.w ? ldx : mu;
  for (int q0 = grp * R; q0 < nrhs; q0 += R * ngrp) {
    const int nr = min(R, nrhs - q0);
    // Loads that do not wait on y go first: the row's old values and, for a full block and one
    // right-hand side, its first (FP32: only) batch of L.
    double old[RH];
    if (row)
#pragma unroll
      for (int r = 0; r < RH; ++r) old[r] = p[(int64_t)(q0 + min(rh0 + r, nr - 1)) * ldp];
    // (KP columns of L per batch: the first batch goes out before the barrier, each later one
    // before the previous is multiplied)
    constexpr int KP = sizeof(T) == 4 ? 64 : 32;
    const bool full1 = R == 1 && pw == SNB;
    T lb[2][KP];
    if (full1)
#pragma unroll
      for (int q = 0; q < KP; ++q) lb[0][q] = buf_load<T>(rs, lo, (uint32_t)q * F.ld * sizeof(T));
    if (threadIdx.x < pw) {  // y of the block: every column's load issued before the stores
      double yv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) yv[r] = y[(int64_t)(q0 + min(r, nr - 1)) * ldx + F.col0 + c0 + threadIdx.x];
#pragma unroll
      for (int r = 0; r < R; ++r) Yb[threadIdx.x][r] = yv[r];
    }
    __syncthreads();
    double s[RH];
#pragma unroll
    for (int r = 0; r < RH; ++r) s[r] = 0.0;
    if (full1) {  // a full block, one right-hand side
#pragma unroll
      for (int j0 = 0; j0 < SNB; j0 += KP) {
        const int cur = (j0 / KP) & 1;
        if (j0 + KP < SNB)
#pragma unroll
          for (int q = 0; q < KP; ++q) lb[cur ^ 1][q] = buf_load<T>(rs, lo, (uint32_t)(j0 + KP + q) * F.ld * sizeof(T));
#pragma unroll
        for (int q = 0; q < KP; ++q) s[0] = fma((double)lb[cur][q], Yb[j0 + q][0], s[0]);
      }
    } else {  // batches of 32 columns, each batch's loads in flight together
      for (int j0 = 0; j0 < pw; j0 += 32) {
        T l[32];
#pragma unroll
        for (int q = 0; q < 32; ++q) l[q] = buf_load<T>(rs, lo, (uint32_t)min(j0 + q, pw - 1) * F.ld * sizeof(T));
#pragma unroll
        for (int q = 0; q < 32; ++q)
          if (j0 + q < pw)
#pragma unroll
            for (int r = 0; r < RH; ++r) s[r] = fma((double)l[q], Yb[j0 + q][rh0 + r], s[r]);
      }
    }
    if (row)
      for (int r = 0; r < RH && rh0 + r < nr; ++r) {
        const double nv = old[r] - s[r];
        p[(int64_t)(q0 + rh0 + r) * ldp] = nv;
        if (it.c && i >= c1 && i < c1 + pn) Vn[i - c1][rh0 + r] = nv;
      }
    if (it.c) {
      __syncthreads();
      inv_apply<T, R>(Inv + F.inv0 + (int64_t)(c1 / SNB) * SNB * SNB, pn, Vn, nr,
                      y + (int64_t)q0 * ldx + F.col0 + c1, ldx);
    }
    __syncthreads();
  }
}

// Backward, per front (item f): its rows below, u_f[i] = x[below[i]], gathered from the
// solution (their columns belong to ancestors, solved before), so the transposed products read
// them contiguously.
__global__ __launch_bounds__(256) void k_bwd_gather(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                    const int32_t* __restrict__ rows, const double* __restrict__ x,
                                                    int64_t ldx, int nrhs, double* __restrict__ u,
                                                    const int64_t* __restrict__ uoff) {
  const int f = items[blockIdx.x].f;
  const FrontDev F = fronts[f];
  const int mu = F.m - F.w;
  const int32_t* below = rows + F.rows_off + F.w;
  double* uf = u + uoff[f] * nrhs;
  for (int i = threadIdx.x; i < mu; i += blockDim.x) {
    const int b = below[i];
    for (int r0 = 0; r0 < nrhs; r0 += 16) {  // 16 columns' loads in flight, then the stores
      double xv[16];
#pragma unroll
      for (int r = 0; r < 16; ++r) xv[r] = x[(int64_t)min(r0 + r, nrhs - 1) * ldx + b];
      for (int r = 0; r < 16 && r0 + r < nrhs; ++r) uf[(int64_t)(r0 + r) * mu + i] = xv[r];
    }
  }
}

// Backward: z_j -= L21(:, j)^T x_below for the columns [a, a + SCOLS) of a front (z in y,
// x_below gathered in u). The rows go in tiles of 64, each tile of L (64 rows x SCOLS columns)
// and of x_below (64 rows x R) staged through LDS with coalesced loads, the next tile's loads
// issued before the current one is multiplied. Wave w takes the columns a + 16 w + [0, 16);
// lane (g, r, h) = (lane / 16, lane % 16 % R, lane % 16 / R) sums, for right-hand side r and the
// R columns 16 w + h R + [0, R), the rows of residue g mod 4 in increasing order; the four
// residue sums are then added as (S0 + S1) + (S2 + S3). Every (column, right-hand side) sum is
// formed in that order for any R, so a right-hand side's result does not depend on how many
// are solved together. Item f, a = j0.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_below(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                   const T* __restrict__ L, const double* __restrict__ u,
                                                   const int64_t* __restrict__ uoff, double* __restrict__ z,
                                                   int64_t ldx, int nrhs) {
  static_assert(SCOLS == 64 && 16 % R == 0, "four waves of 16 columns; R divides 16");
  constexpr int TR = 64;                // rows of a tile
  constexpr int XE = (TR * R + 255) / 256;  // x_below entries a thread stages per tile
  __shared__ T Ls[TR][SCOLS + 1];
  __shared__ double Xs[TR][R + 1];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int mu = F.m - F.w;
  const int j0 = it.a, nc = min(SCOLS, F.w - j0);
  const T* lf = L + F.lpos + F.w;  // L21: row i of column j at lf[j ld + i]
  const double* uf = u + uoff[it.f] * nrhs;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int g = lane >> 4, r = (lane & 15) % R, h = (lane & 15) / R;
  const int cb = 16 * wave + h * R;  // the lane's first column (of the item's SCOLS)
  const int lr = tid & (TR - 1), lc = tid >> 6;  // staging: row lr, columns lc + 4 q
  for (int q0 = 0; q0 < nrhs; q0 += R) {
    const int nr = min(R, nrhs - q0);
    T lv[SCOLS / 4];
    double xv[XE];
    auto load = [&](int t0) {  // clamped in bounds; rows past mu and columns past nc staged as 0
      const int i = min(t0 + lr, mu - 1);
#pragma unroll
      for (int q = 0; q < SCOLS / 4; ++q) lv[q] = lf[(int64_t)(j0 + min(lc + 4 * q, nc - 1)) * F.ld + i];
#pragma unroll
      for (int q = 0; q < XE; ++q) {
        const int e = min(tid + 256 * q, TR * R - 1), k = e / R, rr = e % R;
        xv[q] = uf[(int64_t)(q0 + min(rr, nr - 1)) * mu + min(t0 + k, mu - 1)];
      }
    };
    auto stage = [&](int t0) {
      const bool rin = t0 + lr < mu;
#pragma unroll
      for (int q = 0; q < SCOLS / 4; ++q) {
        Ls[lr][lc + 4 * q] = (rin && lc + 4 * q < nc) ? lv[q] : T(0);
      }
#pragma unroll
      for (int q = 0; q < XE; ++q) {
        const int e = tid + 256 * q, k = e / R, rr = e % R;
        if (e < TR * R) Xs[k][rr] = (t0 + k < mu && rr < nr) ? xv[q] : 0.0;
      }
    };
    double acc[R];
#pragma unroll
    for (int c = 0; c < R; ++c) acc[c] = 0.0;
    if (mu > 0) load(0);
    for (int t0 = 0; t0 < mu; t0 += TR) {
      __syncthreads();  // the previous tile's reads are done
      stage(t0);
      if (t0 + TR < mu) load(t0 + TR);
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < TR / 4; ++kk) {
        const int k = 4 * kk + g;
        const double x = Xs[k][r];
#pragma unroll
        for (int c = 0; c < R; ++c) {
          acc[c] = fma((double)Ls[k][cb + c], x, acc[c]);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < R; ++c) {
      double t = acc[c];
      t += __shfl_xor(t, 16);  // (S0 + S1), (S2 + S3)
      t += __shfl_xor(t, 32);  // (S0 + S1) + (S2 + S3)
      if (g == 0 && cb + c < nc && r < nr) z[(int64_t)(q0 + r) * ldx + F.col0 + j0 + cb + c] -= t;
    }
    __syncthreads();  // LDS reused by the next group of right-hand sides
  }
}

// Backward, per front (item f): x of the front's last block = Inv_last^T z_last.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_first(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                   const T* __restrict__ Inv, const double* __restrict__ z,
                                                   double* __restrict__ x, int64_t ldx, int nrhs, int mark) {
  __shared__ double Zb[SNB][R + 1];
  __shared__ double H[2][SNB][R + 1];
  const FrontDev F = fronts[items[blockIdx.x].f];
  const int cl = (F.w - 1) / SNB * SNB, pw = F.w - cl;
  if (mark)  // one right-hand side, chained: x of the earlier blocks still to come (k_bwd_front)
    for (int e = threadIdx.x; e < cl; e += blockDim.x) reinterpret_cast<uint64_t*>(x)[F.col0 + e] = PENDING;
  for (int q0 = 0; q0 < nrhs; q0 += R) {
    const int nr = min(R, nrhs - q0);
    if (threadIdx.x < pw) {
      double zv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) zv[r] = z[(int64_t)(q0 + min(r, nr - 1)) * ldx + F.col0 + cl + threadIdx.x];
#pragma unroll
      for (int r = 0; r < R; ++r) Zb[threadIdx.x][r] = zv[r];
    }
    __syncthreads();
    inv_apply_t<T, R>(Inv + F.inv0 + (int64_t)(cl / SNB) * SNB * SNB, pw, Zb, nr, x + (int64_t)q0 * ldx + F.col0 + cl,
                      ldx, H);
    __syncthreads();
  }
}

// Backward step of block [c0, c1) (x_b already in x): the columns [j0, j0 + BCOLS) before c0
// lose L(block, j)^T x_b in z. The tile L(c0.., j0..) is staged through LDS 64 columns at a
// time with every load independent (coalesced along the block's rows), then four threads share
// a column, each summing a quarter of the rows for every right-hand side, and the quarters are
// added in a fixed order (so each right-hand side's result does not depend on how many are
// solved together). The item with
// c = 1 (j0 = c0 - BCOLS: the previous block's columns) then forms x of that block from the z it
// has just finished. Item: f, a = c0, b = j0, c.
template <typename T, int R>
__global__ __launch_bounds__(256) void k_bwd_step(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                  const T* __restrict__ L, const T* __restrict__ Inv, double* z,
                                                  double* __restrict__ x, int64_t ldx, int nrhs, int nitems) {
  constexpr int HC = 64;  // columns staged at a time
  __shared__ T Lt[HC][SNB + 1];
  __shared__ double Xb[SNB][R + 1];
  __shared__ double Zn[SNB][R + 1];
  __shared__ double part[4][HC][R + 1];  // also the halves of inv_apply_t (2 x SNB = 4 x HC)
  static_assert(2 * SNB == 4 * HC, "the halves fit the partial sums");
  // workgroup b takes item b % nitems for the right-hand-side groups b / nitems + k gridDim / nitems
  const int grp = blockIdx.x / nitems, ngrp = gridDim.x / nitems;
  const Item it = items[blockIdx.x - grp * nitems];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(SNB, F.w - c0), j0 = it.b;
  const T* lf = L + F.lpos + c0;
  const int tid = threadIdx.x, jq = tid & (HC - 1), g = tid >> 6;
  const int nj = min(BCOLS, c0 - j0);
  for (int q0 = grp * R; q0 < nrhs; q0 += R * ngrp) {
    const int nr = min(R, nrhs - q0);
    if (tid < pw) {
      double xv[R];
#pragma unroll
      for (int r = 0; r < R; ++r) xv[r] = x[(int64_t)(q0 + min(r, nr - 1)) * ldx + F.col0 + c0 + tid];
#pragma unroll
      for (int r = 0; r < R; ++r) Xb[tid][r] = xv[r];
    }
    // thread (row k = tid % 128, every other column): 32 independent loads per half, the next
    // half's issued before the current one is multiplied
    const int lk = tid & (SNB - 1), jb = tid >> 7;
    T v[HC / 2];
    const __amdgpu_buffer_rsrc_t rs = raw_rsrc(lf + (int64_t)j0 * F.ld);
    const uint32_t lo = min(lk, pw - 1) * sizeof(T);
    auto load_half = [&](int h0) {
      const int hc = min(HC, nj - h0);
#pragma unroll
      for (int q = 0; q < HC / 2; ++q) v[q] = buf_load<T>(rs, lo, (uint32_t)(h0 + min(jb + 2 * q, hc - 1)) * F.ld * sizeof(T));
    };
    load_half(0);
    for (int h0 = 0; h0 < nj; h0 += HC) {
      const int hc = min(HC, nj - h0);
#pragma unroll
      for (int q = 0; q < HC / 2; ++q) Lt[jb + 2 * q][lk] = (lk < pw && jb + 2 * q < hc) ? v[q] : T(0);
      if (h0 + HC < nj) load_half(h0 + HC);
      __syncthreads();
      {  // partial sums of column jq over rows [32 g, 32 g + 32), every right-hand side
        double t[R];
#pragma unroll
        for (int r = 0; r < R; ++r) t[r] = 0.0;
        if (jq < hc) {
          const int k1 = min(pw, 32 * g + 32);
          for (int k = 32 * g; k < k1; ++k) {
            const double l = (double)Lt[jq][k];
#pragma unroll
            for (int r = 0; r < R; ++r) t[r] = fma(l, Xb[k][r], t[r]);
          }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) part[g][jq][r] = t[r];
      }
      __syncthreads();
      // the four partial sums in a fixed order, the same for any number of right-hand sides;
      // a thread's entries of z are all read before it writes any
      {
        constexpr int PE = (HC * R + 255) / 256;  // entries per thread
        double old[PE];
#pragma unroll
        for (int q = 0; q < PE; ++q) {
          const int e = min(tid + 256 * q, hc * nr - 1), jj = e % hc, r = e / hc;
          old[q] = z[(int64_t)(q0 + r) * ldx + F.col0 + j0 + h0 + jj];
        }
#pragma unroll
        for (int q = 0; q < PE; ++q) {
          const int e = tid + 256 * q;
          if (e < hc * nr) {
            const int jj = e % hc, r = e / hc;
            const double d = (part[0][jj][r] + part[1][jj][r]) + (part[2][jj][r] + part[3][jj][r]);
            const double nz = old[q] - d;
            z[(int64_t)(q0 + r) * ldx + F.col0 + j0 + h0 + jj] = nz;
            if (it.c) Zn[h0 + jj][r] = nz;
          }
        }
      }
      __syncthreads();
    }
    if (it.c) inv_apply_t<T, R>(Inv + F.inv0 + (int64_t)(j0 / SNB) * SNB * SNB, SNB, Zn, nr,
                                x + (int64_t)q0 * ldx + F.col0 + j0, ldx,
                                reinterpret_cast<double (*)[SNB][R + 1]>(&part[0][0][0]));
    __syncthreads();
  }
}

// One right-hand side, backward: the same step with every global load a workgroup needs issued
// at its entry (a block step then waits on memory once), and the same sums in the same order as
// k_bwd_step, so a single column's result equals that column's result in a block.

// Backward step, item as k_bwd_step (128 columns [j0, j0 + 128) before the block): the tile is
// loaded whole at entry (thread = one row of the block, every other column: coalesced), stored
// to LDS, and thread (jq, g) sums rows 32 g + [0, 32) of columns j0 + jq and j0 + 64 + jq; the
// designated item's threads (b, h) hold half h of column b of the previous block's inverse.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_step1(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                   const T* __restrict__ L, const T* __restrict__ Inv, double* z,
                                                   double* __restrict__ x) {
  __shared__ T Lt[SNB][SNB + 1];  // Lt[j][k] = L(c0 + k, j0 + j)
  __shared__ double Xb[SNB];
  __shared__ double Zn[SNB];
  __shared__ double part[2][4][64];
  __shared__ double H[2][SNB];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(SNB, F.w - c0), j0 = it.b;
  const int tid = threadIdx.x, jq = tid & 63, g = tid >> 6;
  const int lk = tid & (SNB - 1), jb = tid >> 7;
  const __amdgpu_buffer_rsrc_t rs = raw_rsrc(L + F.lpos + (int64_t)j0 * F.ld + c0);
  const uint32_t lo = min(lk, pw - 1) * sizeof(T);
  T v[SNB / 2];
#pragma unroll
  for (int q = 0; q < SNB / 2; ++q) v[q] = buf_load<T>(rs, lo, (uint32_t)(jb + 2 * q) * F.ld * sizeof(T));
  double xv = 0.0, zv = 0.0;
  if (tid < SNB) {
    xv = x[F.col0 + c0 + min(tid, pw - 1)];
    zv = z[F.col0 + j0 + tid];
  }
  const int b = tid & (SNB - 1), hh = tid >> 7;
  T a[64];
  if (it.c) {
    const T* col = Inv + F.inv0 + (int64_t)(j0 / SNB) * SNB * SNB + (int64_t)b * SNB + 64 * hh;
#pragma unroll
    for (int q = 0; q < 64; ++q) a[q] = col[q];
  }
  if (tid < SNB) Xb[tid] = tid < pw ? xv : 0.0;
#pragma unroll
  for (int q = 0; q < SNB / 2; ++q) Lt[jb + 2 * q][lk] = v[q];
  __syncthreads();
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    double t = 0.0;
#pragma unroll
    for (int r = 0; r < 32; ++r)
      if (32 * g + r < pw) t = fma((double)Lt[64 * h + jq][32 * g + r], Xb[32 * g + r], t);
    part[h][g][jq] = t;
  }
  __syncthreads();
  if (tid < SNB) {
    const int h = tid >> 6, jj = tid & 63;
    const double d = (part[h][0][jj] + part[h][1][jj]) + (part[h][2][jj] + part[h][3][jj]);
    const double nz = zv - d;
    z[F.col0 + j0 + tid] = nz;
    Zn[tid] = nz;
  }
  if (it.c) {
    __syncthreads();
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 64; ++q) t = fma((double)a[q], Zn[64 * hh + q], t);
    H[hh][b] = t;
    __syncthreads();
    if (tid < SNB) x[F.col0 + j0 + tid] = H[0][tid] + H[1][tid];
  }
}

// ------------------------------------------------------------------------------------------
// One right-hand side: a level's chain of blocks in one launch each way.
//
// The blocks of a front depend on each other in sequence (y of block b needs every earlier
// block's product with its rows; x of block b every later block's), which the launches above
// follow with one launch per block. Here every workgroup of a level's launch stays resident and
// takes its rows (forward) or columns (backward) through the blocks in order. A block's solution
// formed by another workgroup is waited for value by value: the launch before (k_fwd_gather /
// k_bwd_first with `mark`) fills the block solutions still to come with PENDING, a signalling NaN
// that no arithmetic produces, the forming workgroup writes each value with a device-coherent
// atomic store, and a waiting thread polls its own value with device-coherent atomic loads until
// it is not PENDING. No fence is needed: each value is its own signal. A workgroup waits only on
// workgroups of its own front placed earlier in the launch (lower ids), so the waits resolve in
// dispatch order; every wait is bounded, and one that runs past `limit` polls sets *err and gives
// up (the host reports the solve as failed), so every wave exits. The loads that do not depend on
// a block solution (the rows' L values, the next inverse) are issued before the wait. Per row
// (forward) and column (backward) the arithmetic and its order are those of k_fwd_step and
// k_bwd_step1 / k_bwd_first: the solution is the same bit for bit.

constexpr int FROWS = 2 * SNB;  // rows of a forward workgroup (one a thread)

__device__ __forceinline__ void publish(double* p, double v) {
  __hip_atomic_store(reinterpret_cast<uint64_t*>(p), __builtin_bit_cast(uint64_t, v), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

// The value at p once published; bounded (limit < 0: give up at once, a test's way in).
__device__ __forceinline__ double await_value(double* p, int64_t limit, int32_t* err) {
  uint64_t* q = reinterpret_cast<uint64_t*>(p);
  uint64_t b = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (limit < 0) {
    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return __builtin_bit_cast(double, b);
  }
  int64_t n = 0;
  while (b == PENDING) {
    ++n;
    if (n > limit || ((n & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
      __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
    b = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return __builtin_bit_cast(double, b);
}

// Forward, item (f, a = r0): rows [r0, r0 + FROWS) of front f through every block above them.
// y of block 0 comes from k_fwd_gather (an earlier launch); y of a block whose rows are the
// workgroup's is formed here, kept in LDS for the next block and published. A row's L values of
// the next block are loaded while the current block is multiplied (PF of them; all of a block's
// in FP32). Dynamic LDS: SNB^2 T for the inverse of a block to form, on launches that form any
// (0 otherwise, so the launches of levels without chains keep their occupancy).
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_front(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                   const T* __restrict__ L, const T* __restrict__ Inv, double* v,
                                                   double* y, double* u, const int64_t* __restrict__ uoff,
                                                   int32_t* err, int64_t limit) {
  constexpr int IQ = SNB * SNB / 256;          // inverse entries a thread stages
  constexpr int PF = sizeof(T) == 4 ? 64 : 32;  // L values of the next block loaded ahead
  __shared__ double Yb[SNB][2];
  __shared__ double Vn[SNB][2];
  extern __shared__ __align__(16) unsigned char fwd_dyn[];
  T* Is = reinterpret_cast<T*>(fwd_dyn);  // the next block's inverse, column-major as in Inv
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int r0 = it.a, w = F.w, m = F.m;
  const int tid = threadIdx.x, i = r0 + tid;
  const bool row = i < m;
  double* p = i < w ? v + F.col0 + i : u + uoff[it.f] + i - w;  // the row's entry of v (or of u below)
  double val = row ? *p : 0.0;
  const uint32_t lo = min(i, m - 1) * sizeof(T);  // rows past the front load its last row (not stored)
  const int nb = (w + SNB - 1) / SNB;
  // the row's L values of block b (first PF of them), clamped to the block's width
  auto fetch = [&](int b, T (&dst)[PF]) {
    const int c0 = b * SNB, pw = min(SNB, w - c0);
    const __amdgpu_buffer_rsrc_t rs = raw_rsrc(L + F.lpos + (int64_t)c0 * F.ld);
#pragma unroll
    for (int q = 0; q < PF; ++q) dst[q] = buf_load<T>(rs, lo, (uint32_t)min(q, pw - 1) * F.ld * sizeof(T));
  };
  T cur[PF], nxt[PF];
  fetch(0, cur);
  for (int b = 0; b < nb; ++b) {
    const int c0 = b * SNB, pw = min(SNB, w - c0), c1 = c0 + pw;
    if (c1 >= m || c1 > r0 + FROWS - 1) break;  // no row of the workgroup lies below this block (or a later one)
    const int pn = min(SNB, w - c1);
    const bool form = c1 < w && c1 >= r0;  // the next block's rows are the workgroup's: its y is formed here
    // ahead of any wait: the next block's L values, and the inverse of the block to form
    const bool more = b + 1 < nb && c1 + pn < m && c1 + pn <= r0 + FROWS - 1;
    if (more) fetch(b + 1, nxt);
    T ia[IQ];
    if (form) {
      const T* inv = Inv + F.inv0 + (int64_t)(c1 / SNB) * SNB * SNB;
#pragma unroll
      for (int q = 0; q < IQ; ++q) ia[q] = inv[min(tid + 256 * q, pn * pn - 1)];
    }
    if (c0 < r0 || b == 0) {  // y of the block formed elsewhere: k_fwd_gather (b = 0) or an earlier workgroup
      if (tid < pw) Yb[tid][0] = b == 0 ? y[F.col0 + c0 + tid] : await_value(y + F.col0 + c0 + tid, limit, err);
      __syncthreads();
    }
    // the row's sum over the block, j ascending (k_fwd_step's order)
    double s = 0.0;
    if (pw == SNB) {
      // the rest of the row's values, in batches of RB issued one ahead of the batch multiplied
      constexpr int RB = sizeof(T) == 4 ? 32 : 16, NR = (SNB - PF) / RB;
      static_assert(NR * RB == SNB - PF, "the batches cover the block");
      const __amdgpu_buffer_rsrc_t rs = raw_rsrc(L + F.lpos + (int64_t)c0 * F.ld);
      T rb[2][RB];
#pragma unroll
      for (int q = 0; q < RB; ++q) rb[0][q] = buf_load<T>(rs, lo, (uint32_t)(PF + q) * F.ld * sizeof(T));
#pragma unroll
      for (int q = 0; q < PF; ++q) s = fma((double)cur[q], Yb[q][0], s);
#pragma unroll
      for (int r = 0; r < NR; ++r) {
        if (r + 1 < NR)
#pragma unroll
          for (int q = 0; q < RB; ++q) rb[(r + 1) & 1][q] = buf_load<T>(rs, lo, (uint32_t)(PF + (r + 1) * RB + q) * F.ld * sizeof(T));
#pragma unroll
        for (int q = 0; q < RB; ++q) s = fma((double)rb[r & 1][q], Yb[PF + r * RB + q][0], s);
      }
    } else {
#pragma unroll
      for (int q = 0; q < PF; ++q)
        if (q < pw) s = fma((double)cur[q], Yb[q][0], s);
      if (PF < SNB) {
        const __amdgpu_buffer_rsrc_t rs = raw_rsrc(L + F.lpos + (int64_t)c0 * F.ld);
        for (int j0 = PF; j0 < pw; j0 += 32) {
          T l[32];
#pragma unroll
          for (int q = 0; q < 32; ++q) l[q] = buf_load<T>(rs, lo, (uint32_t)min(j0 + q, pw - 1) * F.ld * sizeof(T));
#pragma unroll
          for (int q = 0; q < 32; ++q)
            if (j0 + q < pw) s = fma((double)l[q], Yb[j0 + q][0], s);
        }
      }
    }
    if (row && i >= c1) val -= s;
    if (form) {
      if (row && i >= c1 && i < c1 + pn) Vn[i - c1][0] = val;
#pragma unroll
      for (int q = 0; q < IQ; ++q)
        if (tid + 256 * q < pn * pn) Is[tid + 256 * q] = ia[q];
      __syncthreads();  // (every read of Yb above is done too)
      if (tid < pn) {
        double yn = 0.0;  // inv_apply's sum for one right-hand side: k ascending, one fma each
        if (pn == SNB) {  // batches of LDS reads issued ahead of their products
#pragma unroll
          for (int k0 = 0; k0 < SNB; k0 += 32) {
            T a[32];
            double vv[32];
#pragma unroll
            for (int q = 0; q < 32; ++q) {
              a[q] = Is[(k0 + q) * SNB + tid];
              vv[q] = Vn[k0 + q][0];
            }
#pragma unroll
            for (int q = 0; q < 32; ++q) yn = fma((double)a[q], vv[q], yn);
          }
        } else {
          for (int k = 0; k < pn; ++k) yn = fma((double)Is[k * pn + tid], Vn[k][0], yn);
        }
        Yb[tid][0] = yn;
        publish(y + F.col0 + c1 + tid, yn);
      }
    }
    __syncthreads();  // Yb whole for the next block, and every read of it done
    if (more)
#pragma unroll
      for (int q = 0; q < PF; ++q) cur[q] = nxt[q];
  }
  if (row) *p = val;
}

// Backward, item (f, a = jb): the columns of block jb (a full block, not the front's last)
// through every later block, last first; then x of block jb, published. x of the front's last
// block comes from k_bwd_first (an earlier launch). A front's items are in decreasing jb, so the
// workgroup forming a block's x precedes every workgroup waiting for it.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_front(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                   const T* __restrict__ L, const T* __restrict__ Inv, double* z,
                                                   double* x, int32_t* err, int64_t limit) {
  __shared__ T Lt[SNB][SNB + 1];  // Lt[j][k] = L(c0 + k, j0 + j)
  __shared__ double Xb[SNB];
  __shared__ double Zn[SNB];
  __shared__ double part[2][4][64];
  __shared__ double H[2][SNB];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int jb = it.a, j0 = jb * SNB;
  const int nb = (F.w + SNB - 1) / SNB;
  const int tid = threadIdx.x, jq = tid & 63, g = tid >> 6;
  const int lk = tid & (SNB - 1), jh = tid >> 7;
  // half jh of column lk of the block's inverse, for x of the block at the end
  T ca[64];
  {
    const T* col = Inv + F.inv0 + (int64_t)jb * SNB * SNB + (int64_t)lk * SNB + 64 * jh;
#pragma unroll
    for (int q = 0; q < 64; ++q) ca[q] = col[q];
  }
  double zv = tid < SNB ? z[F.col0 + j0 + tid] : 0.0;
  for (int b = nb - 1; b > jb; --b) {
    const int c0 = b * SNB, pw = min(SNB, F.w - c0);
    const __amdgpu_buffer_rsrc_t rs = raw_rsrc(L + F.lpos + (int64_t)j0 * F.ld + c0);
    const uint32_t lo = min(lk, pw - 1) * sizeof(T);
    {  // the tile into LDS before any wait (it does not depend on x)
      T vt[SNB / 2];
#pragma unroll
      for (int q = 0; q < SNB / 2; ++q) vt[q] = buf_load<T>(rs, lo, (uint32_t)(jh + 2 * q) * F.ld * sizeof(T));
#pragma unroll
      for (int q = 0; q < SNB / 2; ++q) Lt[jh + 2 * q][lk] = vt[q];
    }
    if (tid < SNB) {
      const double xv = b == nb - 1 ? x[F.col0 + c0 + min(tid, pw - 1)]
                                    : await_value(x + F.col0 + c0 + min(tid, pw - 1), limit, err);
      Xb[tid] = tid < pw ? xv : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double t = 0.0;
#pragma unroll
      for (int r = 0; r < 32; ++r)
        if (32 * g + r < pw) t = fma((double)Lt[64 * h + jq][32 * g + r], Xb[32 * g + r], t);
      part[h][g][jq] = t;
    }
    __syncthreads();
    if (tid < SNB) {
      const int h = tid >> 6, jj = tid & 63;
      const double d = (part[h][0][jj] + part[h][1][jj]) + (part[h][2][jj] + part[h][3][jj]);
      zv = zv - d;
      Zn[tid] = zv;
    }
    __syncthreads();
  }
  // x of block jb = Inv_jb^T z_jb (k_bwd_step1's designated item)
  double t = 0.0;
#pragma unroll
  for (int q = 0; q < 64; ++q) t = fma((double)ca[q], Zn[64 * jh + q], t);
  H[jh][lk] = t;
  __syncthreads();
  if (tid < SNB) {
    z[F.col0 + j0 + tid] = zv;
    publish(x + F.col0 + j0 + tid, H[0][tid] + H[1][tid]);
  }
}

// x_perm[k, r] = b[perm[k], r] and back.
__global__ void k_permute(const double* __restrict__ src, double* __restrict__ dst, const int64_t* __restrict__ perm,
                          int64_t n, int nrhs, int64_t lds, int64_t ldd, int inverse) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n; k += (int64_t)gridDim.x * blockDim.x)
    for (int r = 0; r < nrhs; ++r) {
      if (inverse) dst[r * ldd + perm[k]] = src[r * lds + k];
      else dst[r * ldd + k] = src[r * lds + perm[k]];
    }
}

// ------------------------------------------------------------------------------------------
// Vector steps of refinement, FP64, on blocks of k <= MAX_RHS columns (column c at c * ld).
// Reductions are two-stage: each workgroup of a launch reduces one row block of one column
// (row blocks never straddle a segment of independent equations), and the host combines the
// partials in a fixed order, so results are the same on every run.

// Y = A X by rows (Csr::mul, sparse.rs:42-50): eight lanes per row, partial sums combined in a
// fixed order; columns in chunks of eight so a row's entries are read once per chunk.
__global__ __launch_bounds__(256) void k_spmm(int64_t n, const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                              const double* __restrict__ val, const double* __restrict__ x,
                                              double* __restrict__ y, int k, int64_t ld) {
  const int sub = threadIdx.x & 7;
  for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 3; r < n; r += ((int64_t)gridDim.x * blockDim.x) >> 3) {
    const int64_t t0 = ptr[r], t1 = ptr[r + 1];
    for (int c0 = 0; c0 < k; c0 += 8) {
      double s[8] = {};
      // The lane's entries t, t + 8, t + 16, t + 24 at a time: their values and columns, then
      // their x in every column of the chunk, all in flight together rather than each entry's
      // loads waited on in turn. Entries past the row read its last entry and add nothing; the
      // sums keep entry order. The chunk's column count is a constant of the sweep, so no
      // column takes a test of its own.
      auto sweep = [&](auto nc) {
        for (int64_t t = t0 + sub; t < t1; t += 32) {
          double v[4];
          int64_t j[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int64_t tt = min(t + 8 * u, t1 - 1);
            v[u] = val[tt];
            j[u] = idx[tt];
          }
#pragma unroll
          for (int q = 0; q < 8; ++q)
            if (q < decltype(nc)::value) {
              double xv[4];
#pragma unroll
              for (int u = 0; u < 4; ++u) xv[u] = x[(int64_t)(c0 + q) * ld + j[u]];
#pragma unroll
              for (int u = 0; u < 4; ++u) s[q] = t + 8 * u < t1 ? s[q] + v[u] * xv[u] : s[q];
            }
        }
      };
      switch (min(k - c0, 8)) {
        case 1: sweep(std::integral_constant<int, 1>{}); break;
        case 2: sweep(std::integral_constant<int, 2>{}); break;
        case 3: sweep(std::integral_constant<int, 3>{}); break;
        case 4: sweep(std::integral_constant<int, 4>{}); break;
        case 5: sweep(std::integral_constant<int, 5>{}); break;
        case 6: sweep(std::integral_constant<int, 6>{}); break;
        case 7: sweep(std::integral_constant<int, 7>{}); break;
        default: sweep(std::integral_constant<int, 8>{}); break;
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        s[q] += __shfl_xor(s[q], 1);
        s[q] += __shfl_xor(s[q], 2);
        s[q] += __shfl_xor(s[q], 4);
      }
      if (sub == 0)
#pragma unroll
        for (int q = 0; q < 8; ++q)
          if (c0 + q < k) y[(int64_t)(c0 + q) * ld + r] = s[q];
    }
  }
}

struct Coefs {
  double v[MAX_RHS];
};

// r = b - ax over n entries
__global__ void k_residual(int64_t n, const double* __restrict__ b, const double* __restrict__ ax,
                           double* __restrict__ r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - ax[i];
}

// Y[:, c] += a_c X[:, c]
__global__ void k_axpy_cols(int64_t n, int k, int64_t ld, Coefs a, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i / n);
    const int64_t e = (int64_t)c * ld + i % n;
    y[e] += a.v[c] * x[e];
  }
}

// Y[:, c] = X[:, c] + b_c Y[:, c]
__global__ void k_xpby_cols(int64_t n, int k, int64_t ld, const double* __restrict__ x, Coefs b, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n * k; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i / n);
    const int64_t e = (int64_t)c * ld + i % n;
    y[e] = x[e] + b.v[c] * y[e];
  }
}

// partial[c * nblk + blk] = sum a*b (mode 0) or max |a| (mode 1) over rows [rb[blk], rb[blk + 1])
// of column c; grid (nblk, k)
__global__ __launch_bounds__(256) void k_reduce(const int64_t* __restrict__ rb, int nblk, const double* __restrict__ a,
                                                const double* __restrict__ b, int64_t ld, int mode,
                                                double* __restrict__ partial) {
  __shared__ double red[4];
  const int blk = blockIdx.x, c = blockIdx.y;
  const double* ac = a + (int64_t)c * ld;
  const double* bc = b + (int64_t)c * ld;
  double s = 0.0;
  for (int64_t i = rb[blk] + threadIdx.x; i < rb[blk + 1]; i += blockDim.x)
    s = mode == 0 ? s + ac[i] * bc[i] : fmax(s, fabs(ac[i]));
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(s, o);
    s = mode == 0 ? s + t : fmax(s, t);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < 4; ++w) t = mode == 0 ? t + red[w] : fmax(t, red[w]);
    partial[(int64_t)c * nblk + blk] = t;
  }
}

}  // namespace fc
