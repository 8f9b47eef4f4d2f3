// Tall-skinny kernels of the shift-invert block Lanczos on the device (gfx950, wave64): the
// basis Q and its M products MQ are n x m column-major in HBM; a block W is n x b (b <= 8).
// Both kernels stream the basis once per call and are HBM-bound.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fc {

constexpr int TS_ROWS = 4096;   // rows of one partial sum of ts_dots (and of the argmax chunks)
constexpr int TS_COLS = 8;      // basis columns one ts_dots workgroup takes
constexpr int TS_MAXB = 8;      // block columns

// partial[(chunk * m + i) * b + j] = sum over the chunk's rows r of A[r, i] B[r, j], for the
// TS_COLS columns i of this workgroup's tile: each thread walks rows r = chunk start + tid +
// 256 k with the b entries of B's row in registers, then the workgroup reduces each of its
// TS_COLS x b sums in a fixed order (wave shuffles, then the four waves in order), so every
// partial is the same bits on every run.
__global__ __launch_bounds__(256) void k_ts_dots(int64_t n, const double* __restrict__ A, int64_t lda, int m,
                                                 const double* __restrict__ B, int64_t ldb, int b,
                                                 double* __restrict__ partial) {
  const int tile = blockIdx.y, chunk = blockIdx.x;
  const int i0 = tile * TS_COLS;
  const int cols = min(TS_COLS, m - i0);
  const int64_t r0 = (int64_t)chunk * TS_ROWS, r1 = min<int64_t>(n, r0 + TS_ROWS);
  double acc[TS_COLS][TS_MAXB];
#pragma unroll
  for (int i = 0; i < TS_COLS; ++i)
#pragma unroll
    for (int j = 0; j < TS_MAXB; ++j) acc[i][j] = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
    double bv[TS_MAXB];
#pragma unroll
    for (int j = 0; j < TS_MAXB; ++j) bv[j] = j < b ? B[(int64_t)j * ldb + r] : 0.0;
#pragma unroll
    for (int i = 0; i < TS_COLS; ++i) {
      if (i < cols) {
        const double a = A[(int64_t)(i0 + i) * lda + r];
#pragma unroll
        for (int j = 0; j < TS_MAXB; ++j) acc[i][j] += a * bv[j];
      }
    }
  }
  __shared__ double red[4][TS_COLS * TS_MAXB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < TS_COLS; ++i)
#pragma unroll
    for (int j = 0; j < TS_MAXB; ++j) {
      double v = acc[i][j];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[wave][i * TS_MAXB + j] = v;
    }
  __syncthreads();
  for (int q = threadIdx.x; q < cols * b; q += 256) {
    const int i = q / b, j = q % b;
    const double v = red[0][i * TS_MAXB + j] + red[1][i * TS_MAXB + j] + red[2][i * TS_MAXB + j] +
                     red[3][i * TS_MAXB + j];
    partial[((int64_t)chunk * m + i0 + i) * b + j] = v;
  }
}

// C[i * b + j] = sum over chunks, in chunk order, of the partials.
__global__ void k_ts_dots_sum(const double* __restrict__ partial, int chunks, int m, int b, double* __restrict__ C) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m * b) return;
  double v = 0.0;
  for (int c = 0; c < chunks; ++c) v += partial[(int64_t)c * m * b + q];
  C[q] = v;
}

// B[r, j] = beta B[r, j] + alpha sum_i A[r, i] C[i * ldc + j] for j < b: one thread per row,
// the coefficients in LDS (m x b <= TS_MAXC entries), each row's sums in column order.
constexpr int TS_MAXC = 4096;
__global__ __launch_bounds__(256) void k_ts_update(int64_t n, const double* __restrict__ A, int64_t lda, int m,
                                                   const double* __restrict__ C, int ldc, double* __restrict__ B,
                                                   int64_t ldb, int b, double alpha, double beta) {
  __shared__ double c[TS_MAXC];
  for (int q = threadIdx.x; q < m * b; q += 256) c[q] = C[(q / b) * ldc + q % b];
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r >= n) return;
  double acc[TS_MAXB];
#pragma unroll
  for (int j = 0; j < TS_MAXB; ++j) acc[j] = 0.0;
  for (int i = 0; i < m; ++i) {
    const double a = A[(int64_t)i * lda + r];
#pragma unroll
    for (int j = 0; j < TS_MAXB; ++j)
      if (j < b) acc[j] += a * c[i * b + j];
  }
#pragma unroll
  for (int j = 0; j < TS_MAXB; ++j)
    if (j < b) {
      double* p = B + (int64_t)j * ldb + r;
      *p = (beta == 0.0 ? 0.0 : beta * *p) + alpha * acc[j];  // beta = 0: B is not read
    }
}

// x[r] *= s (one column)
__global__ void k_scale(int64_t n, double* __restrict__ x, double s) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) x[r] *= s;
}

// y[r] = x[r] (one column)
__global__ void k_copy_col(int64_t n, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (r < n) y[r] = x[r];
}

// per chunk of TS_ROWS rows: the first index of the largest |x| (ties: the lower index)
__global__ void k_argmax_abs(int64_t n, const double* __restrict__ x, int64_t* __restrict__ at) {
  __shared__ double bv[256];
  __shared__ int64_t bi[256];
  const int64_t r0 = (int64_t)blockIdx.x * TS_ROWS, r1 = min<int64_t>(n, r0 + TS_ROWS);
  double best = -1.0;
  int64_t bidx = r0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += 256) {
    const double v = fabs(x[r]);
    if (v > best) {
      best = v;
      bidx = r;
    }
  }
  bv[threadIdx.x] = best;
  bi[threadIdx.x] = bidx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double o = bv[threadIdx.x + s];
      const int64_t oi = bi[threadIdx.x + s];
      if (o > bv[threadIdx.x] || (o == bv[threadIdx.x] && oi < bi[threadIdx.x])) {
        bv[threadIdx.x] = o;
        bi[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) at[blockIdx.x] = bi[0];
}

}  // namespace fc
