// Supernodal triangular solves over the fronts (the reference's solve_group,
// cholesky.rs:904-1037, reorganised for levels of the front tree) and the FP64 residual SpMV.
//
// Forward, leaves first: a front gathers its children's contributions (each child hands up
// the products of its rows below, u_c = v_c,below - L21_c y_c), solves L11 y = v_own and hands
// up u = v_below - L21 y. Backward, root first: x_own = L11^-T (y - L21^T x_below), the rows
// below read from the solution. Contributions travel up the tree instead of being scattered
// into ancestors, so no two workgroups write one entry and no atomics are used. The right-hand
// sides are columns of x (postorder numbering) with leading dimension ldx.
#pragma once

#include "device.hpp"

namespace fc {

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0;
  for (int w = 0; w < BS / 64; ++w) s += red[w];
  __syncthreads();
  return s;
}

// Item: f. u holds each front's contribution (M - W) per right-hand side at uoff[f].
template <typename T>
__global__ __launch_bounds__(256) void k_forward(const int32_t* __restrict__ fl, const FrontDev* __restrict__ fronts,
                                                 const int32_t* __restrict__ kids, const int32_t* __restrict__ rel,
                                                 const T* __restrict__ L, double* __restrict__ x, int64_t ldx, int nrhs,
                                                 double* __restrict__ u, const int64_t* __restrict__ uoff) {
  __shared__ double red[4];
  __shared__ double yshare;
  const int f = fl[blockIdx.x];
  const FrontDev F = fronts[f];
  const int w = F.w, mu = F.m - F.w;
  const T* lf = L + F.lpos;
  for (int r = 0; r < nrhs; ++r) {
    double* xo = x + r * ldx + F.col0;
    double* uf = u + uoff[f] * nrhs + (int64_t)r * mu;
    for (int i = threadIdx.x; i < mu; i += blockDim.x) uf[i] = 0.0;
    __syncthreads();
    for (int q = F.kid0; q < F.kid1; ++q) {
      const int c = kids[q];
      const FrontDev C = fronts[c];
      const int cmu = C.m - C.w;
      const double* uc = u + uoff[c] * nrhs + (int64_t)r * cmu;
      const int32_t* rc = rel + C.rel_off;
      for (int i = threadIdx.x; i < cmu; i += blockDim.x) {
        const int p = rc[i];
        if (p < w) xo[p] += uc[i];
        else uf[p - w] += uc[i];
      }
      __syncthreads();
    }
    // L11 y = v_own, column by column
    for (int j = 0; j < w; ++j) {
      if (threadIdx.x == 0) {
        const double y = xo[j] / (double)lf[(int64_t)j * F.ld + j];
        xo[j] = y;
        yshare = y;
      }
      __syncthreads();
      const double y = yshare;
      const T* col = lf + (int64_t)j * F.ld;
      for (int i = j + 1 + threadIdx.x; i < w; i += blockDim.x) xo[i] -= (double)col[i] * y;
      __syncthreads();
    }
    // u = v_below - L21 y
    for (int i = threadIdx.x; i < mu; i += blockDim.x) {
      double s = 0.0;
      for (int j = 0; j < w; ++j) s += (double)lf[(int64_t)j * F.ld + w + i] * xo[j];
      uf[i] -= s;
    }
    __syncthreads();
  }
}

// Item: f. rows holds each front's rows (postorder columns) at rows_off.
template <typename T>
__global__ __launch_bounds__(256) void k_backward(const int32_t* __restrict__ fl, const FrontDev* __restrict__ fronts,
                                                  const int32_t* __restrict__ rows, const T* __restrict__ L,
                                                  double* __restrict__ x, int64_t ldx, int nrhs) {
  __shared__ double red[4];
  const int f = fl[blockIdx.x];
  const FrontDev F = fronts[f];
  const int w = F.w, mu = F.m - F.w;
  const T* lf = L + F.lpos;
  const int32_t* below = rows + F.rows_off + w;
  for (int r = 0; r < nrhs; ++r) {
    double* xr = x + r * ldx;
    double* xo = xr + F.col0;
    // z_j = y_j - sum_i L21(i, j) x_below(i): a wave per column
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j = wave; j < w; j += blockDim.x / 64) {
      const T* col = lf + (int64_t)j * F.ld + w;
      double s = 0.0;
      for (int i = lane; i < mu; i += 64) s += (double)col[i] * xr[below[i]];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) xo[j] -= s;
    }
    __syncthreads();
    // L11^T x = z, last column first
    for (int j = w - 1; j >= 0; --j) {
      const T* col = lf + (int64_t)j * F.ld;
      double s = 0.0;
      for (int i = j + 1 + threadIdx.x; i < w; i += blockDim.x) s += (double)col[i] * xo[i];
      s = block_sum<256>(s, red);
      if (threadIdx.x == 0) xo[j] = (xo[j] - s) / (double)col[j];
      __syncthreads();
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

// y = A x by rows, each row summed in column order (Csr::mul, sparse.rs:42-50). A wave per
// row block of 64 rows... one thread per row keeps the reference's order of the sum.
__global__ void k_spmv(int64_t n, const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                       const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < n; r += (int64_t)gridDim.x * blockDim.x) {
    double s = -0.0;
    for (int64_t t = ptr[r]; t < ptr[r + 1]; ++t) s += val[t] * x[idx[t]];
    y[r] = s;
  }
}

}  // namespace fc

namespace fc {

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
