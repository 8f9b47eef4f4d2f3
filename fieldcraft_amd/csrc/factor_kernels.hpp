// Kernels of the multifrontal numeric phase (gfx950). All are work-list kernels: workgroup b
// runs item items[item0 + b] of its launch; one launch covers one level of the front tree.
//
// Per level: the children's update matrices are added into the fronts' L parts (extend-add);
// each front is factored by panels of 128 columns, left-looking inside the front like the
// reference's panel loop (cholesky.rs:1138-1186, panels of 64 there): the panel minus its
// products with the front's earlier columns (MFMA GEMM), its diagonal block factored with the
// pivot test of cholesky.rs:1161, the rows below solved against it; then the front's update
// matrix U = -L21 L21^T (MFMA SYRK) and the children's update matrices added into it. The
// dense kernels are in gemm_kernels.hpp; these are the entry, scaling and extend-add steps.
#pragma once

#include "device.hpp"

namespace fc {

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
// Item: f, a, b, and at ranges[c .. c + d) the children's column runs that land in [a, b),
// window of EA_COLS parent columns after window, children in increasing order within a
// window (children with none are not listed; the runs are found when the launches are planned).

template <typename T>
__global__ __launch_bounds__(256) void k_extend_add(const Item* __restrict__ items, const EaRange* __restrict__ ranges,
                                                    const FrontDev* __restrict__ fronts,
                                                    const int32_t* __restrict__ rel, T* __restrict__ L, const T* U,
                                                    T* Uw) {
  // U (children, read) and Uw (this front, written) are disjoint blocks of one arena
  const Item it = items[blockIdx.x];
  const FrontDev P = fronts[it.f];
  const int pw = P.w;
  const int64_t pmu = P.m - P.w;
  for (int q = it.c; q < it.c + it.d; ++q) {
    const EaRange rg = ranges[q];
    const FrontDev C = fronts[rg.child];
    const int cmu = C.m - C.w;
    const int32_t* r = rel + C.rel_off;
    const int jc0 = rg.jc0, jc1 = rg.jc1;  // child columns whose parent column lies in [a, b)
    const T* cu = U + C.upos;
    // The child's entries that land in [j0, j1): columns jc0 + t (t < nc <= EA_COLS, rel is
    // injective), rows jc0 + t .. cmu - 1 of each, taken as one flat range so that small
    // children still fill the workgroup. Entry e of the range is column t with
    // off[t] <= e < off[t + 1], row jc0 + t + e - off[t].
    const int nc = jc1 - jc0;
    int off[EA_COLS + 1];
    T* dcol[EA_COLS];
    off[0] = 0;
#pragma unroll
    for (int t = 0; t < EA_COLS; ++t) {
      const int jc = min(jc0 + min(t, max(nc - 1, 0)), cmu - 1);  // in bounds
      off[t + 1] = off[t] + (t < nc ? cmu - (jc0 + t) : 0);
      const int jp = nc > 0 ? r[jc] : 0;
      dcol[t] = jp < pw ? L + P.lpos + (int64_t)jp * P.ld : Uw + P.upos + (int64_t)(jp - pw) * pmu - pw;
    }
    const int total = off[EA_COLS];
    // EA_BATCH entries a thread at a time: the row map and the child's entries, then the
    // targets, then the sums (distinct targets within one child). Entries past the range read
    // its last entry (in bounds) and store nothing.
    for (int e0 = threadIdx.x; e0 < total; e0 += EA_BATCH * blockDim.x) {
      int ri[EA_BATCH], tc[EA_BATCH];
      T sv[EA_BATCH], dv[EA_BATCH];
#pragma unroll
      for (int k = 0; k < EA_BATCH; ++k) {
        const int e = min(e0 + k * (int)blockDim.x, total - 1);
        int t = 0;
#pragma unroll
        for (int q2 = 1; q2 < EA_COLS; ++q2) t += e >= off[q2];
        const int ic = jc0 + t + e - off[t];
        tc[k] = t;
        ri[k] = r[ic];
        sv[k] = cu[(int64_t)(jc0 + t) * cmu + ic];
      }
#pragma unroll
      for (int k = 0; k < EA_BATCH; ++k) {
        T* d = dcol[0];
#pragma unroll
        for (int q2 = 1; q2 < EA_COLS; ++q2) d = tc[k] == q2 ? dcol[q2] : d;
        dv[k] = d[ri[k]];
      }
#pragma unroll
      for (int k = 0; k < EA_BATCH; ++k)
        if (e0 + k * (int)blockDim.x < total) {
          T* d = dcol[0];
#pragma unroll
          for (int q2 = 1; q2 < EA_COLS; ++q2) d = tc[k] == q2 ? dcol[q2] : d;
          d[ri[k]] = dv[k] + sv[k];
        }
    }
    __syncthreads();
  }
}

}  // namespace fc
