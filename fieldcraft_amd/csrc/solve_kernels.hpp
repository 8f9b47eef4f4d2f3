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
// A front of at most SB columns is solved by one workgroup. A wider front is solved by blocks
// of SB columns: the block's triangle by one workgroup, then the rows below the block (forward)
// or the columns before it (backward) by many workgroups. The right-hand sides are the columns
// of x (postorder numbering, leading dimension ldx); every entry of L is read once per solve
// whatever their number.
#pragma once

#include "device.hpp"

namespace fc {

constexpr int SB = 64;        // columns of a solve block
constexpr int SROWS = 256;    // rows of a forward GEMV workgroup
constexpr int SCOLS = 32;     // columns of a backward GEMV workgroup
constexpr int MAX_RHS = 64;   // right-hand sides one launch handles (the host splits larger blocks)

// Item: f. Zero the front's contribution block and add its children's into it and into x.
__global__ __launch_bounds__(256) void k_solve_gather(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                      const int32_t* __restrict__ kids, const int32_t* __restrict__ rel,
                                                      double* __restrict__ x, int64_t ldx, int nrhs, double* u,
                                                      const int64_t* __restrict__ uoff) {
  const int f = items[blockIdx.x].f;
  const FrontDev F = fronts[f];
  const int w = F.w, mu = F.m - F.w;
  double* uf = u + uoff[f] * nrhs;
  for (int e = threadIdx.x; e < mu * nrhs; e += blockDim.x) uf[e] = 0.0;
  __syncthreads();
  for (int q = F.kid0; q < F.kid1; ++q) {
    const int c = kids[q];
    const FrontDev C = fronts[c];
    const int cmu = C.m - C.w;
    const double* uc = u + uoff[c] * nrhs;
    const int32_t* rc = rel + C.rel_off;
    for (int e = threadIdx.x; e < cmu * nrhs; e += blockDim.x) {
      const int i = e % cmu, r = e / cmu;
      const int p = rc[i];
      if (p < w) x[r * ldx + F.col0 + p] += uc[e];
      else uf[r * mu + p - w] += uc[e];
    }
    __syncthreads();
  }
}

// The triangle of block [c0, c1) of a front's L11 into LDS (lower part).
template <typename T>
__device__ __forceinline__ void load_block(const T* __restrict__ lf, int ld, int c0, int bw, double (*D)[SB + 1]) {
  for (int e = threadIdx.x; e < SB * SB; e += blockDim.x) {
    const int i = e % SB, j = e / SB;
    D[i][j] = (i < bw && j < bw && i >= j) ? (double)lf[(int64_t)(c0 + j) * ld + c0 + i] : 0.0;
  }
}

// y = L_block^-1 v (forward) on x rows [c0, c0 + bw) of every right-hand side, in place.
__device__ __forceinline__ void block_forward(double (*D)[SB + 1], double* xo, int64_t ldx, int nrhs, int bw,
                                              double (*Y)[MAX_RHS]) {
  for (int e = threadIdx.x; e < bw * nrhs; e += blockDim.x) Y[e % bw][e / bw] = xo[(e / bw) * ldx + e % bw];
  __syncthreads();
  for (int j = 0; j < bw; ++j) {
    if (threadIdx.x < nrhs) Y[j][threadIdx.x] /= D[j][j];
    __syncthreads();
    for (int e = threadIdx.x; e < (bw - j - 1) * nrhs; e += blockDim.x) {
      const int i = j + 1 + e % (bw - j - 1), r = e / (bw - j - 1);
      Y[i][r] -= D[i][j] * Y[j][r];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < bw * nrhs; e += blockDim.x) xo[(e / bw) * ldx + e % bw] = Y[e % bw][e / bw];
  __syncthreads();
}

// x = L_block^-T z (backward) on x rows [c0, c0 + bw), in place.
__device__ __forceinline__ void block_backward(double (*D)[SB + 1], double* xo, int64_t ldx, int nrhs, int bw,
                                               double (*Y)[MAX_RHS]) {
  for (int e = threadIdx.x; e < bw * nrhs; e += blockDim.x) Y[e % bw][e / bw] = xo[(e / bw) * ldx + e % bw];
  __syncthreads();
  for (int j = bw - 1; j >= 0; --j) {
    if (threadIdx.x < nrhs) Y[j][threadIdx.x] /= D[j][j];
    __syncthreads();
    for (int e = threadIdx.x; e < j * nrhs; e += blockDim.x) {
      const int i = e % j, r = e / j;
      Y[i][r] -= D[j][i] * Y[j][r];
    }
    __syncthreads();
  }
  for (int e = threadIdx.x; e < bw * nrhs; e += blockDim.x) xo[(e / bw) * ldx + e % bw] = Y[e % bw][e / bw];
  __syncthreads();
}

// Forward triangle of a block: item f, a = c0. A front of at most SB columns also hands up
// its contribution (b = 1): u -= L21 y.
template <typename T>
__global__ __launch_bounds__(256) void k_forward_block(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                       const T* __restrict__ L, double* __restrict__ x, int64_t ldx,
                                                       int nrhs, double* __restrict__ u, const int64_t* __restrict__ uoff) {
  __shared__ double D[SB][SB + 1];
  __shared__ double Y[SB][MAX_RHS];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, bw = min(SB, F.w - c0);
  const T* lf = L + F.lpos;
  load_block(lf, F.ld, c0, bw, D);
  __syncthreads();
  double* xo = x + F.col0 + c0;
  block_forward(D, xo, ldx, nrhs, bw, Y);
  if (it.b) {  // the whole front: u = v_below - L21 y
    const int mu = F.m - F.w;
    double* uf = u + uoff[it.f] * nrhs;
    for (int i = threadIdx.x; i < mu; i += blockDim.x) {
      double s[8];
      for (int r0 = 0; r0 < nrhs; r0 += 8) {
        const int nr = min(8, nrhs - r0);
        for (int r = 0; r < 8; ++r) s[r] = 0.0;
        for (int j = 0; j < bw; ++j) {
          const double l = (double)lf[(int64_t)j * F.ld + F.w + i];
          for (int r = 0; r < nr; ++r) s[r] += l * Y[j][r0 + r];
        }
        for (int r = 0; r < nr; ++r) uf[(r0 + r) * mu + i] -= s[r];
      }
    }
  }
}

// Forward update of the rows below a solved block: item f, a = c0, b = r0. Rows r0..r0+SROWS
// (>= c0 + SB) lose L(r, block) y_block: rows of the front's columns in x, rows below in u.
template <typename T>
__global__ __launch_bounds__(256) void k_forward_gemv(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                      const T* __restrict__ L, double* __restrict__ x, int64_t ldx,
                                                      int nrhs, double* __restrict__ u, const int64_t* __restrict__ uoff) {
  __shared__ double Y[SB][MAX_RHS];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, bw = min(SB, F.w - c0), r0 = it.b;
  const T* lf = L + F.lpos;
  for (int e = threadIdx.x; e < bw * nrhs; e += blockDim.x) Y[e % bw][e / bw] = x[(e / bw) * ldx + F.col0 + c0 + e % bw];
  __syncthreads();
  const int mu = F.m - F.w;
  double* uf = u + uoff[it.f] * nrhs;
  const int i = r0 + threadIdx.x;
  if (i < F.m && threadIdx.x < SROWS) {
    double s[8];
    for (int q0 = 0; q0 < nrhs; q0 += 8) {
      const int nr = min(8, nrhs - q0);
      for (int r = 0; r < 8; ++r) s[r] = 0.0;
      for (int j = 0; j < bw; ++j) {
        const double l = (double)lf[(int64_t)(c0 + j) * F.ld + i];
        for (int r = 0; r < nr; ++r) s[r] += l * Y[j][q0 + r];
      }
      if (i < F.w)
        for (int r = 0; r < nr; ++r) x[(q0 + r) * ldx + F.col0 + i] -= s[r];
      else
        for (int r = 0; r < nr; ++r) uf[(q0 + r) * mu + i - F.w] -= s[r];
    }
  }
}

// Backward: z_j -= L21(:, j)^T x_below for the columns [a, a + SCOLS) of a front; a wave per
// column at a time. Item f, a = j0.
template <typename T>
__global__ __launch_bounds__(256) void k_backward_below(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                        const int32_t* __restrict__ rows, const T* __restrict__ L,
                                                        double* __restrict__ x, int64_t ldx, int nrhs) {
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int mu = F.m - F.w;
  const T* lf = L + F.lpos;
  const int32_t* below = rows + F.rows_off + F.w;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j1 = min(F.w, it.a + SCOLS);
  for (int j = it.a + wave; j < j1; j += 4) {
    const T* col = lf + (int64_t)j * F.ld + F.w;
    for (int r = 0; r < nrhs; ++r) {
      const double* xr = x + r * ldx;
      double s = 0.0;
      for (int i = lane; i < mu; i += 64) s += (double)col[i] * xr[below[i]];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) x[r * ldx + F.col0 + j] -= s;
    }
  }
}

// Backward triangle of a block: item f, a = c0.
template <typename T>
__global__ __launch_bounds__(256) void k_backward_block(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                        const T* __restrict__ L, double* __restrict__ x, int64_t ldx,
                                                        int nrhs) {
  __shared__ double D[SB][SB + 1];
  __shared__ double Y[SB][MAX_RHS];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, bw = min(SB, F.w - c0);
  load_block(L + F.lpos, F.ld, c0, bw, D);
  __syncthreads();
  block_backward(D, x + F.col0 + c0, ldx, nrhs, bw, Y);
}

// Backward update of the columns before a solved block: z_j -= L(block, j)^T x_block for
// columns [b, b + SCOLS) below c0 = a. Item f, a = c0, b = j0.
template <typename T>
__global__ __launch_bounds__(256) void k_backward_gemv(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                       const T* __restrict__ L, double* __restrict__ x, int64_t ldx,
                                                       int nrhs) {
  __shared__ double Y[SB][MAX_RHS];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, bw = min(SB, F.w - c0);
  const T* lf = L + F.lpos;
  for (int e = threadIdx.x; e < bw * nrhs; e += blockDim.x) Y[e % bw][e / bw] = x[(e / bw) * ldx + F.col0 + c0 + e % bw];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j1 = min(c0, it.b + SCOLS);
  for (int j = it.b + wave; j < j1; j += 4) {
    const double l = lane < bw ? (double)lf[(int64_t)j * F.ld + c0 + lane] : 0.0;
    for (int r = 0; r < nrhs; ++r) {
      double s = l * Y[lane < bw ? lane : 0][r];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) x[r * ldx + F.col0 + j] -= s;
    }
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

// y = A x by rows (Csr::mul, sparse.rs:42-50): eight lanes per row, each row's partial sums
// combined in a fixed order.
__global__ __launch_bounds__(256) void k_spmv(int64_t n, const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                              const double* __restrict__ val, const double* __restrict__ x,
                                              double* __restrict__ y) {
  const int sub = threadIdx.x & 7;
  for (int64_t r = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 3; r < n; r += ((int64_t)gridDim.x * blockDim.x) >> 3) {
    double s = 0.0;
    for (int64_t t = ptr[r] + sub; t < ptr[r + 1]; t += 8) s += val[t] * x[idx[t]];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    s += __shfl_xor(s, 4);
    if (sub == 0) y[r] = s;
  }
}

// ------------------------------------------------------------------------------------------
// Vector steps of refinement, FP64. Reductions are two-stage with a fixed grid (partials of
// RED_BLOCKS workgroups, summed on the host in order), so they are the same on every run.
constexpr int RED_BLOCKS = 512;

__global__ void k_residual(int64_t n, const double* __restrict__ b, const double* __restrict__ ax,
                           double* __restrict__ r) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    r[i] = b[i] - ax[i];
}

__global__ void k_axpy(int64_t n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] += alpha * x[i];
}

__global__ void k_xpby(int64_t n, const double* __restrict__ x, double beta, double* __restrict__ y) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = x[i] + beta * y[i];
}

// partial[block] = sum a*b over the block's stride (mode 0) or max |a| (mode 1)
__global__ __launch_bounds__(256) void k_reduce(int64_t n, const double* __restrict__ a, const double* __restrict__ b,
                                                int mode, double* __restrict__ partial) {
  __shared__ double red[4];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    s = mode == 0 ? s + a[i] * b[i] : fmax(s, fabs(a[i]));
  for (int o = 32; o > 0; o >>= 1) {
    const double t = __shfl_xor(s, o);
    s = mode == 0 ? s + t : fmax(s, t);
  }
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = red[0];
    for (int w = 1; w < 4; ++w) t = mode == 0 ? t + red[w] : fmax(t, red[w]);
    partial[blockIdx.x] = t;
  }
}

}  // namespace fc
