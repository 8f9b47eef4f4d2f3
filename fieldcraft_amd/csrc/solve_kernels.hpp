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
// A front is solved by blocks of NB columns, its factorization's panels, whose diagonal blocks'
// inverses the factorization keeps. One launch per block: every workgroup forms the block's
// solution with the inverse (a 128 x 128 product, cheap to repeat), the first one stores it,
// and each updates its own share of the rows below (forward) or of the columns before
// (backward). The forward pass reads v from x and writes y to a second vector; the backward
// pass reads z from that vector and writes x, so no workgroup reads what another writes. The
// right-hand sides are columns (postorder numbering, leading dimension ldx); every entry of L
// is read once per RC right-hand sides.
#pragma once

#include "device.hpp"

namespace fc {

constexpr int SNB = 128;      // columns of a solve block (the factorization's panel width)
constexpr int SROWS = 256;    // rows of a forward workgroup
constexpr int SCOLS = 32;     // columns of a backward workgroup
constexpr int RC = 16;        // right-hand sides a workgroup keeps in LDS at a time
constexpr int MAX_RHS = 64;   // right-hand sides one launch handles (the host splits larger blocks)

// Item: f. Zero the front's contribution block and add its children's into it and into v.
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

// Forward step of block [c0, c1): y_b = Linv v_b (stored by the item with c = 1), then rows
// [r0, r0 + SROWS) at or below c1 lose L(r, block) y_b: rows of the front's columns in v (x),
// rows below in the front's contribution u. Item: f, a = c0, b = r0, c = store flag.
template <typename T>
__global__ __launch_bounds__(256) void k_fwd_step(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                  const T* __restrict__ L, const T* __restrict__ Inv, double* v,
                                                  double* __restrict__ y, int64_t ldx, int nrhs, double* __restrict__ u,
                                                  const int64_t* __restrict__ uoff) {
  __shared__ double Vb[SNB][RC + 1];
  __shared__ double Yb[SNB][RC + 1];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(SNB, F.w - c0), c1 = c0 + pw, r0 = it.b;
  const T* inv = Inv + F.inv0 + (int64_t)(c0 / SNB) * SNB * SNB;  // ld = pw
  const T* lf = L + F.lpos;
  const int mu = F.m - F.w;
  double* uf = u + uoff[it.f] * nrhs;
  const int i = r0 + threadIdx.x;
  const bool row = threadIdx.x < SROWS && i >= c1 && i < F.m;
  for (int q0 = 0; q0 < nrhs; q0 += RC) {
    const int nr = min(RC, nrhs - q0);
    for (int e = threadIdx.x; e < pw * nr; e += blockDim.x) {
      const int k = e % pw, r = e / pw;
      Vb[k][r] = v[(int64_t)(q0 + r) * ldx + F.col0 + c0 + k];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < pw * nr; e += blockDim.x) {
      const int b = e % pw, r = e / pw;
      double s = 0.0;
      for (int k = 0; k <= b; ++k) s += (double)inv[(int64_t)k * pw + b] * Vb[k][r];
      Yb[b][r] = s;
      if (it.c) y[(int64_t)(q0 + r) * ldx + F.col0 + c0 + b] = s;
    }
    __syncthreads();
    if (row) {
      double s[RC];
#pragma unroll
      for (int r = 0; r < RC; ++r) s[r] = 0.0;
      for (int j = 0; j < pw; ++j) {
        const double l = (double)lf[(int64_t)(c0 + j) * F.ld + i];
#pragma unroll
        for (int r = 0; r < RC; ++r) s[r] += l * Yb[j][r];
      }
      if (i < F.w) {
        for (int r = 0; r < nr; ++r) v[(int64_t)(q0 + r) * ldx + F.col0 + i] -= s[r];
      } else {
        for (int r = 0; r < nr; ++r) uf[(int64_t)(q0 + r) * mu + i - F.w] -= s[r];
      }
    }
    __syncthreads();
  }
}

// Backward: z_j -= L21(:, j)^T x_below for the columns [a, a + SCOLS) of a front, z in y and
// x_below in x; a wave per column at a time. Item f, a = j0.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_below(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                   const int32_t* __restrict__ rows, const T* __restrict__ L,
                                                   const double* __restrict__ x, double* __restrict__ z, int64_t ldx,
                                                   int nrhs) {
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
      const double* xr = x + (int64_t)r * ldx;
      double s = 0.0;
      for (int i = lane; i < mu; i += 64) s += (double)col[i] * xr[below[i]];
      for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
      if (lane == 0) z[(int64_t)r * ldx + F.col0 + j] -= s;
    }
  }
}

// Backward step of block [c0, c1): x_b = Linv^T z_b (stored by the item with c = 1), then the
// columns [j0, j0 + SCOLS) before c0 lose L(block, j)^T x_b in z. Item: f, a = c0, b = j0,
// c = store flag.
template <typename T>
__global__ __launch_bounds__(256) void k_bwd_step(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                  const T* __restrict__ L, const T* __restrict__ Inv, double* z,
                                                  double* __restrict__ x, int64_t ldx, int nrhs) {
  __shared__ double Zb[SNB][RC + 1];
  __shared__ double Xb[SNB][RC + 1];
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(SNB, F.w - c0), j0 = it.b;
  const T* inv = Inv + F.inv0 + (int64_t)(c0 / SNB) * SNB * SNB;  // ld = pw
  const T* lf = L + F.lpos;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int j1 = min(c0, j0 + SCOLS);
  for (int q0 = 0; q0 < nrhs; q0 += RC) {
    const int nr = min(RC, nrhs - q0);
    for (int e = threadIdx.x; e < pw * nr; e += blockDim.x) {
      const int k = e % pw, r = e / pw;
      Zb[k][r] = z[(int64_t)(q0 + r) * ldx + F.col0 + c0 + k];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < pw * nr; e += blockDim.x) {
      const int b = e % pw, r = e / pw;
      double s = 0.0;
      const T* col = inv + (int64_t)b * pw;
      for (int k = b; k < pw; ++k) s += (double)col[k] * Zb[k][r];
      Xb[b][r] = s;
      if (it.c) x[(int64_t)(q0 + r) * ldx + F.col0 + c0 + b] = s;
    }
    __syncthreads();
    for (int j = j0 + wave; j < j1; j += 4) {
      const T* col = lf + (int64_t)j * F.ld + c0;
      const bool h0 = lane < pw, h1 = lane + 64 < pw;  // Xb holds pw rows only
      const double l0 = h0 ? (double)col[lane] : 0.0;
      const double l1 = h1 ? (double)col[lane + 64] : 0.0;
      for (int r = 0; r < nr; ++r) {
        double s = (h0 ? l0 * Xb[lane][r] : 0.0) + (h1 ? l1 * Xb[lane + 64][r] : 0.0);
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane == 0) z[(int64_t)(q0 + r) * ldx + F.col0 + j] -= s;
      }
    }
    __syncthreads();
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
