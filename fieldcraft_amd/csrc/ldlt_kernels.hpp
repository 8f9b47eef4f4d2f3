// LDL^T with bounded Bunch-Kaufman (rook) pivoting inside supernodes (crates/solver/src/ldlt.rs)
// on the device: one workgroup per supernode, the supernodes of one level of the supernodal tree
// per launch, and the pivoted solves (cholesky.rs:904-1037). Each entry's sums run in the
// reference's order, one thread per sum, so the factor and the solutions are the reference's bit
// for bit (each function here turns floating-point contraction off).
//
// Storage: the reference's lower trapezoids by columns (supernode s at voff[s], m_s rows,
// column c holding rows c..m_s: offset(m, c) = c m - c (c - 1) / 2), pivot kinds per column
// (0 one, 1 first, 2 second of a 2 x 2), each supernode's pivot order (symbolic local column at
// each pivot position).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fc {

constexpr double LDL_ALPHA = 0.6403882032022076;  // (1 + sqrt 17) / 8, ldlt.rs:78
constexpr int LDL_ROOK_STEPS = 64;                // ldlt.rs:82
constexpr int LDL_PANEL = 64;                     // cholesky.rs:54
constexpr double LDL_PIVOT_TOLERANCE = 1e-10;     // cholesky.rs:48
enum : int8_t { PIV_ONE = 0, PIV_FIRST = 1, PIV_SECOND = 2 };

// The complex scalar of complex-symmetric LDL^T (ldlt.rs over Complex64, scalar.rs:115-131) with
// num-complex 0.4's arithmetic: products (ac - bd, ad + bc), quotients by the squared norm, sums
// by parts; the modulus hypot(re, im) for the pivot tests. Transposes, never conjugates.
struct zd {
  double re, im;
};
__device__ __forceinline__ zd operator+(zd a, zd b) {
#pragma clang fp contract(off)
  return zd{a.re + b.re, a.im + b.im};
}
__device__ __forceinline__ zd operator-(zd a, zd b) {
#pragma clang fp contract(off)
  return zd{a.re - b.re, a.im - b.im};
}
__device__ __forceinline__ zd operator*(zd a, zd b) {
#pragma clang fp contract(off)
  return zd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ zd operator/(zd a, zd b) {
#pragma clang fp contract(off)
  const double nrm = b.re * b.re + b.im * b.im;
  return zd{(a.re * b.re + a.im * b.im) / nrm, (a.im * b.re - a.re * b.im) / nrm};
}
__device__ __forceinline__ zd& operator+=(zd& a, zd b) { return a = a + b; }
__device__ __forceinline__ zd& operator-=(zd& a, zd b) { return a = a - b; }
__device__ __forceinline__ zd& operator/=(zd& a, zd b) { return a = a / b; }
__device__ __forceinline__ double ldl_mod(double x) { return fabs(x); }
__device__ __forceinline__ double ldl_mod(zd x) { return hypot(x.re, x.im); }

struct LdlSupernode {
  int64_t voff;      // trapezoid
  int64_t rows;      // row structure (postorder columns) in the rows array
  int64_t first;     // first column (postorder)
  int64_t upd0, upd1;  // updates it receives
  int64_t scratch;   // two work columns of m entries
  int32_t m, w;
};

struct LdlUpdate {
  int64_t k;     // descendant supernode
  int64_t rel;   // relative map (mu entries: rows p1.. of k -> rows of the target) in the rel array
  int32_t p1, p2;
};

__device__ __forceinline__ int64_t ldl_offset(int64_t m, int64_t c) { return c * m - c * (c > 0 ? c - 1 : 0) / 2; }
#define LDL_AT(t, m, r, c) ((t)[ldl_offset((m), (c)) + (r) - (c)])

// Entry k of times_d(row) (cholesky.rs:460-484) for the row `row(j)` of a trapezoid t (m rows)
// over its columns k0..: a 1 x 1 pivot scales, a 2 x 2 pivot mixes the pair with D's block.
template <typename T, typename Row>
__device__ __forceinline__ T times_d_at(const Row& row, const T* t, int64_t m, int64_t k0,
                                             const int8_t* piv, int k, int len) {
#pragma clang fp contract(off)
  if (piv[k] == PIV_FIRST) {
    const T a = row(k), b = k + 1 < len ? row(k + 1) : T{};
    return a * LDL_AT(t, m, k0 + k, k0 + k) + b * LDL_AT(t, m, k0 + k + 1, k0 + k);
  }
  if (piv[k] == PIV_SECOND && k > 0 && piv[k - 1] == PIV_FIRST) {
    const T a = row(k - 1), b = row(k);
    return a * LDL_AT(t, m, k0 + k, k0 + k - 1) + b * LDL_AT(t, m, k0 + k, k0 + k);
  }
  return row(k) * LDL_AT(t, m, k0 + k, k0 + k);
}

constexpr int LDL_YC = 16;   // target columns of one update window (and one tile)
constexpr int LDL_KC = 128;  // source columns of Y staged in LDS at a time
constexpr int LDL_PP = 8;    // pairs per thread per batch

// own(rel(i), rel(c)) -= sum_{k < kw} L(p + i, k0 + k) Y(p + c, k), Y = L D (products and
// subtract, cholesky.rs:531-686 with pivots), for c < min(wu, mu), c <= i < mu, restricted to the
// target columns rel(c) in [cb, ce) (at most LDL_YC of them) and the target rows rel(i) in
// [rb, re): each (i, c) summed by one thread in k order from 0.0. The window's Y entries are
// staged in LDS LDL_KC columns at a time; each thread carries LDL_PP sums across the stages.
// All threads of the workgroup call it together.
template <typename T>
__device__ inline void ldl_update(T* own, int64_t m, const T* lk, int64_t mk, int64_t k0, int kw,
                                  const int8_t* pk, int64_t p, int64_t mu, int64_t wu, const int32_t* rel, bool rel_id,
                                  int64_t rel0, int64_t cb, int64_t ce, int64_t rb, int64_t re) {
#pragma clang fp contract(off)
  const int64_t cols = wu < mu ? wu : mu;
  // the columns c whose target lies in [cb, ce) and the rows i whose target lies in [rb, re):
  // rel is increasing, so both are ranges
  auto first_at_least = [&](int64_t v, int64_t hi) {
    if (rel_id) {
      const int64_t t = v - rel0;
      return t < 0 ? (int64_t)0 : (t > hi ? hi : t);
    }
    int64_t lo = 0;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (rel[mid] < v) lo = mid + 1;
      else hi = mid;
    }
    return lo;
  };
  const int64_t ca = first_at_least(cb, cols), cz = first_at_least(ce, cols);
  const int64_t ia = first_at_least(rb, mu), iz = first_at_least(re, mu);
  if (ca >= cz || ia >= iz) return;
  const int ncol = (int)(cz - ca);
  // pairs of column ca + t: rows max(ca + t, ia) .. iz; pre[t] = pairs before it
  __shared__ int64_t pre[LDL_YC + 1];
  __shared__ T ys[LDL_YC * LDL_KC];
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int t = 0; t < ncol; ++t) {
      pre[t] = acc;
      const int64_t lo = ca + t > ia ? ca + t : ia;
      acc += iz > lo ? iz - lo : 0;
    }
    pre[ncol] = acc;
  }
  __syncthreads();
  const int64_t pairs = pre[ncol];
  const int64_t d0 = ldl_offset(mk, k0);  // diagonal entry of column k0
  for (int64_t q0 = 0; q0 < pairs; q0 += (int64_t)blockDim.x * LDL_PP) {
    int64_t li[LDL_PP], tgt[LDL_PP];
    int yrow[LDL_PP];
    T sum[LDL_PP];
#pragma unroll
    for (int t = 0; t < LDL_PP; ++t) {
      const int64_t q = q0 + threadIdx.x + (int64_t)t * blockDim.x;
      sum[t] = T{};
      yrow[t] = -1;
      li[t] = 0;
      tgt[t] = 0;
      if (q < pairs) {
        int c = 0;
        while (c + 1 < ncol && pre[c + 1] <= q) ++c;
        const int64_t lo = ca + c > ia ? ca + c : ia;
        const int64_t i = lo + (q - pre[c]);
        yrow[t] = c;
        li[t] = d0 + p + i - k0;  // L(p + i, k0)
        const int64_t ri = rel_id ? rel0 + i : rel[i], rc = rel_id ? rel0 + ca + c : rel[ca + c];
        tgt[t] = ldl_offset(m, rc) + ri - rc;
      }
    }
    for (int kb = 0; kb < kw; kb += LDL_KC) {
      const int kn = kw - kb < LDL_KC ? kw - kb : LDL_KC;
      __syncthreads();  // the previous stage is consumed
      for (int e = threadIdx.x; e < ncol * kn; e += blockDim.x) {
        const int t = e / kn, kk = e - t * kn, k = kb + kk;
        const int64_t r = p + ca + t;  // row of Y
        const int64_t dg = ldl_offset(mk, k0 + k);  // D(k0 + k, k0 + k)
        const T a = lk[dg + r - (k0 + k)];
        T y;
        const int8_t pv = pk[k];
        if (pv == PIV_FIRST) {  // row(k) d11 + row(k + 1) d21
          const T bb = k + 1 < kw ? LDL_AT(lk, mk, r, k0 + k + 1) : T{};
          y = a * lk[dg] + bb * lk[dg + 1];
        } else if (pv == PIV_SECOND && k > 0 && pk[k - 1] == PIV_FIRST) {  // row(k - 1) d21 + row(k) d22
          y = LDL_AT(lk, mk, r, k0 + k - 1) * LDL_AT(lk, mk, k0 + k, k0 + k - 1) + a * lk[dg];
        } else {
          y = a * lk[dg];
        }
        ys[t * LDL_KC + kk] = y;
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < LDL_PP; ++t) {
        if (yrow[t] < 0) continue;
        const T* yv = ys + yrow[t] * LDL_KC;
        int64_t at = li[t];
        T acc = sum[t];
        int64_t step = mk - k0 - kb - 1;  // entries of column k0 + k + 1 lie mk - k0 - k - 1 on
#pragma unroll 4
        for (int kk = 0; kk < kn; ++kk) {
          acc += lk[at] * yv[kk];
          at += step;
          step -= 1;
        }
        sum[t] = acc;
        li[t] = at;
      }
    }
#pragma unroll
    for (int t = 0; t < LDL_PP; ++t)
      if (yrow[t] >= 0) own[tgt[t]] -= sum[t];
  }
  __syncthreads();
}

// A's values into the trapezoids (the columns of A, cholesky.rs:1054-1068): every upper entry
// has one slot; the trapezoids were zeroed.
template <typename T>
__global__ void k_ldlt_scatter(int64_t nnz, const T* __restrict__ aval, const int64_t* __restrict__ pos,
                               T* __restrict__ val) {
#pragma clang fp contract(off)
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < nnz) val[pos[e]] = aval[e];
}

template <typename T>
__global__ void k_ldlt_zero(int64_t n, T* __restrict__ v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = T{};
}

// The descendants' updates of the supernodes of one level (the assemble of cholesky.rs:1069-1124
// with Y = L D): workgroup t takes the target columns [tiles[t].b, tiles[t].e) of supernode
// tiles[t].j and applies every update to them in order, so each entry receives its updates in
// the reference's order whatever the tiling.
struct LdlTile {
  int64_t j, b, e;  // supernode, target columns [b, e)
  int64_t rb, re;   // target rows [rb, re)
};
template <typename T>
__global__ __launch_bounds__(256) void k_ldlt_assemble(const LdlTile* __restrict__ tiles, const LdlSupernode* __restrict__ S,
                                                       const LdlUpdate* __restrict__ U, const int32_t* __restrict__ rel,
                                                       T* __restrict__ val, const int8_t* __restrict__ piv) {
#pragma clang fp contract(off)
  const LdlTile tl = tiles[blockIdx.x];
  const LdlSupernode J = S[tl.j];
  T* own = val + J.voff;
  for (int64_t u = J.upd0; u < J.upd1; ++u) {
    const LdlUpdate up = U[u];
    const LdlSupernode K = S[up.k];
    ldl_update(own, J.m, val + K.voff, K.m, 0, K.w, piv + K.first, up.p1, K.m - up.p1, up.p2 - up.p1, rel + up.rel, false,
               0, tl.b, tl.e, tl.rb, tl.re);
  }
}

// One supernode per workgroup (supernode list sn[blockIdx.x]), assembled: the panels
// (ldlt.rs:256-358). fail[s] receives the failing pivot's postorder
// column (or stays -1); two[s] the supernode's 2 x 2 pivots.
template <typename T, bool ONE>
__global__ __launch_bounds__(1024) void k_ldlt_supernode(const int64_t* __restrict__ sn, const LdlSupernode* __restrict__ S,
                                                        const LdlUpdate* __restrict__ U, const int32_t* __restrict__ rel,
                                                        const double* __restrict__ scale, T* __restrict__ val,
                                                        int8_t* __restrict__ piv, int32_t* __restrict__ order,
                                                        T* __restrict__ work, int64_t* __restrict__ fail,
                                                        int32_t* __restrict__ two, int64_t* __restrict__ state, int lds_rows) {
#pragma clang fp contract(off)
  const int64_t j = sn[blockIdx.x];
  const LdlSupernode J = S[j];
  const int64_t m = J.m, w = J.w;
  T* own = val + J.voff;
  int8_t* pv = piv + J.first;
  int32_t* od = order + J.first;
  (void)U;
  (void)rel;
  int64_t* st = state + 4 * j;  // ONE: start and end of the last panel, failure, 2 x 2 pivots
  if (!ONE) {
    for (int64_t c = threadIdx.x; c < w; c += blockDim.x) {
      pv[c] = PIV_ONE;
      od[c] = (int32_t)c;
    }
    __syncthreads();
  } else if (st[2] >= 0 || st[1] >= w) {
    __syncthreads();
    if (threadIdx.x == 0) st[0] = st[1] = w;  // nothing left: the trailing update stays idle
    return;
  }
  // two work columns (swapped by pointer, as the reference's), in dynamic LDS when the launch
  // gave room for m rows (lds_rows), else in the work array
  extern __shared__ __align__(16) unsigned char ldl_dyn[];
  T* A = m <= lds_rows ? reinterpret_cast<T*>(ldl_dyn) : work + J.scratch;
  T* B = A + m;
  const double* sc = scale + J.first;
  __shared__ double red_v[16];   // one entry per wave (up to 1024 threads)
  __shared__ int64_t red_i[16], red_n[16];
  __shared__ int64_t sh[6];          // decisions broadcast by thread 0
  __shared__ double shd[2];
  int twos = 0;
  // current(c) into out: column c rows k..m less the panel's factored columns c0..k (ldlt.rs:176-193)
  __shared__ T ydc[LDL_PANEL + 2];
  auto current = [&](int64_t c0, int64_t k, int64_t c, T* out) {
    auto row = [&](int q) { return LDL_AT(own, m, c, c0 + q); };
    for (int64_t q = threadIdx.x; q < k - c0; q += blockDim.x)
      ydc[q] = times_d_at(row, own, m, c0, pv + c0, (int)q, (int)(k - c0));
    __syncthreads();
    const int64_t o0 = ldl_offset(m, c0) - c0;
    for (int64_t r = k + threadIdx.x; r < m; r += blockDim.x) {
      T v = r >= c ? LDL_AT(own, m, r, c) : LDL_AT(own, m, c, r);
      int64_t o = o0 + r;  // (r, q): the next column's entry lies m - q - 1 on
#pragma unroll 8
      for (int64_t q = c0; q < k; ++q) {
        v -= own[o] * ydc[q - c0];
        o += m - q - 1;
      }
      out[r] = v;
    }
    __syncthreads();
  };
  // largest |v[i]| over [lo, hi) except skip (ldlt.rs:149-164): the first index of the largest,
  // or the first NaN's index with NaN
  auto largest = [&](const T* v, int64_t lo, int64_t hi, int64_t skip, int64_t* at, double* value) {
    double bv = 0.0;
    int64_t bi = -1, nan_i = -1;
    for (int64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      if (i == skip) continue;
      const double x = ldl_mod(v[i]);
      if (isnan(x)) {
        if (nan_i < 0) nan_i = i;
      } else if (x > bv) {
        bv = x;
        bi = i;
      }
    }
    // the wave's best (largest value, then the lowest index; zeros carry no index) and first
    // NaN, then thread 0 over the waves in order
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off, 64);
      const int64_t oi = __shfl_xor(bi, off, 64), on = __shfl_xor(nan_i, off, 64);
      if (ov > bv || (ov == bv && ov > 0.0 && oi < bi)) {
        bv = ov;
        bi = oi;
      }
      if (on >= 0 && (nan_i < 0 || on < nan_i)) nan_i = on;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
      red_v[wv] = bv;
      red_i[wv] = bi;
      red_n[wv] = nan_i;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      int64_t best_i = -1, first_nan = -1;
      double best = 0.0;
      for (int t = 0; t < (int)(blockDim.x >> 6); ++t) {
        if (red_v[t] > best || (red_v[t] == best && red_v[t] > 0.0 && red_i[t] < best_i)) {
          best = red_v[t];
          best_i = red_i[t];
        }
        if (red_n[t] >= 0 && (first_nan < 0 || red_n[t] < first_nan)) first_nan = red_n[t];
      }
      sh[4] = best_i < 0 ? lo : best_i;
      shd[0] = best;
      if (first_nan >= 0) {  // the reference returns at the first NaN it meets, whatever came before
        sh[4] = first_nan;
        shd[0] = __builtin_nan("");
      }
    }
    __syncthreads();
    *at = sh[4];
    *value = shd[0];
    __syncthreads();
  };
  auto bswap = [&](int64_t p, int64_t q) {
    if (p == q) return;
    if (p > q) {
      const int64_t t = p;
      p = q;
      q = t;
    }
    // interchange (ldlt.rs:131-145): rows p, q of the columns before p, the diagonals, the strip
    // between them, the rows below q
    const int64_t n1 = p, n2 = q - p - 1, n3 = m - q - 1;
    for (int64_t e = threadIdx.x; e < n1 + 1 + n2 + n3; e += blockDim.x) {
      T *x, *y;
      if (e < n1) {
        x = &LDL_AT(own, m, p, e);
        y = &LDL_AT(own, m, q, e);
      } else if (e == n1) {
        x = &LDL_AT(own, m, p, p);
        y = &LDL_AT(own, m, q, q);
      } else if (e < n1 + 1 + n2) {
        const int64_t c = p + 1 + (e - n1 - 1);
        x = &LDL_AT(own, m, c, p);
        y = &LDL_AT(own, m, q, c);
      } else {
        const int64_t r = q + 1 + (e - n1 - 1 - n2);
        x = &LDL_AT(own, m, r, p);
        y = &LDL_AT(own, m, r, q);
      }
      const T t = *x;
      *x = *y;
      *y = t;
    }
    if (threadIdx.x == 0) {
      const int32_t t = od[p];
      od[p] = od[q];
      od[q] = t;
      T x = A[p];
      A[p] = A[q];
      A[q] = x;
      x = B[p];
      B[p] = B[q];
      B[q] = x;
    }
    __syncthreads();
  };
  int64_t c0 = ONE ? st[1] : 0;
  int64_t failed = -1;
  while (c0 < w && failed < 0) {
    const int64_t end = c0 + LDL_PANEL < w ? c0 + LDL_PANEL : w;
    int64_t k = c0;
    while (k < end) {
      current(c0, k, k, A);
      const double akk = ldl_mod(A[k]);
      int64_t r0;
      double omega_k;
      largest(A, k + 1, w, -1, &r0, &omega_k);
      if (!isfinite(akk) || !isfinite(omega_k)) {
        failed = k;
        break;
      }
      bool is_two = false;
      int64_t one_at = k, ti = k, tr = k;
      if (!(akk >= LDL_ALPHA * omega_k)) {
        int64_t i = k, r = r0;
        double omega_i = omega_k;
        int step = 0;
        for (;;) {
          current(c0, k, r, B);
          int64_t s_at;
          double omega_r;
          largest(B, k, w, r, &s_at, &omega_r);
          const double arr = ldl_mod(B[r]);
          if (!isfinite(arr) || !isfinite(omega_r)) {
            failed = k;
            break;
          }
          step += 1;
          if (arr >= LDL_ALPHA * omega_r) {
            T* t = A;
            A = B;
            B = t;
            one_at = r;
            break;
          }
          if (omega_r <= omega_i || step >= LDL_ROOK_STEPS) {
            is_two = true;
            ti = i;
            tr = r;
            break;
          }
          T* t = A;
          A = B;
          B = t;
          i = r;
          omega_i = omega_r;
          r = s_at;
        }
        if (failed >= 0) break;
      }
      if (!is_two) {
        bswap(k, one_at);
        if (!(ldl_mod(A[k]) > LDL_PIVOT_TOLERANCE * sc[od[k]])) {
          failed = k;
          break;
        }
        // one (ldlt.rs:207-221)
        const T d = A[k];
        const int64_t start = ldl_offset(m, k);
        for (int64_t r = k + threadIdx.x; r < m; r += blockDim.x) own[start + r - k] = r == k ? d : A[r] / d;
        if (threadIdx.x == 0) pv[k] = PIV_ONE;
        __syncthreads();
        k += 1;
      } else {
        bswap(k, ti);
        const int64_t r = tr == k ? ti : tr;
        bswap(k + 1, r);
        const double off = ldl_mod(A[k + 1]);
        const double s0 = sc[od[k]], s1 = sc[od[k + 1]];
        if (!(off > LDL_PIVOT_TOLERANCE * (s0 > s1 ? s0 : s1))) {
          failed = k;
          break;
        }
        // two (ldlt.rs:224-253)
        const T d11 = A[k], d21 = A[k + 1], d22 = B[k + 1];
        const T det = d11 * d22 - d21 * d21;
        const int64_t q0 = ldl_offset(m, k), q1 = ldl_offset(m, k + 1);
        for (int64_t rr = k + 2 + threadIdx.x; rr < m; rr += blockDim.x) {
          const T x = A[rr], y = B[rr];
          own[q0 + rr - k] = (x * d22 - y * d21) / det;
          own[q1 + rr - k - 1] = (y * d11 - x * d21) / det;
        }
        if (threadIdx.x == 0) {
          own[q0] = d11;
          own[q0 + 1] = d21;
          own[q1] = d22;
          pv[k] = PIV_FIRST;
          pv[k + 1] = PIV_SECOND;
        }
        twos += 1;
        __syncthreads();
        k += 2;
      }
    }
    if (failed >= 0) break;
    const int64_t c1 = k;
    if (ONE) {  // the trailing update is k_ldlt_trailing's, over many workgroups
      if (threadIdx.x == 0) {
        st[0] = c0;
        st[1] = c1;
      }
      break;
    }
    for (int64_t cb = c1; cb < w; cb += LDL_YC)
      ldl_update(own, m, own, m, c0, (int)(c1 - c0), pv + c0, c1, m - c1, w - c1, nullptr, true, c1, cb,
                 cb + LDL_YC < w ? cb + LDL_YC : w, cb, m);
    __syncthreads();
    c0 = c1;
  }
  if (threadIdx.x == 0) {
    if (!ONE) {
      fail[j] = failed >= 0 ? J.first + od[failed] : -1;
      two[j] = twos;
    } else {
      st[3] += twos;
      two[j] = (int32_t)st[3];
      if (failed >= 0) {
        st[2] = J.first + od[failed];
        fail[j] = st[2];
        st[0] = st[1] = w;
      }
    }
  }
}

// State of the supernodes factored one panel per launch: pivots and order reset, no panel yet.
__global__ void k_ldlt_state_init(const int64_t* __restrict__ sn, const LdlSupernode* __restrict__ S,
                                  int8_t* __restrict__ piv, int32_t* __restrict__ order, int64_t* __restrict__ fail,
                                  int32_t* __restrict__ two, int64_t* __restrict__ state) {
  const int64_t j = sn[blockIdx.x];
  const LdlSupernode J = S[j];
  for (int64_t c = threadIdx.x; c < J.w; c += blockDim.x) {
    piv[J.first + c] = PIV_ONE;
    order[J.first + c] = (int32_t)c;
  }
  if (threadIdx.x == 0) {
    state[4 * j] = 0;
    state[4 * j + 1] = 0;
    state[4 * j + 2] = -1;
    state[4 * j + 3] = 0;
    fail[j] = -1;
    two[j] = 0;
  }
}

// The trailing update after one panel (ldlt.rs:347-352) of a supernode factored one panel per
// launch: workgroup t takes the target columns [tiles[t].b, tiles[t].e) that lie at or after the
// panel's end.
template <typename T>
__global__ __launch_bounds__(256) void k_ldlt_trailing(const LdlTile* __restrict__ tiles, const LdlSupernode* __restrict__ S,
                                                       T* __restrict__ val, const int8_t* __restrict__ piv,
                                                       const int64_t* __restrict__ state) {
#pragma clang fp contract(off)
  const LdlTile tl = tiles[blockIdx.x];
  const LdlSupernode J = S[tl.j];
  const int64_t c0 = state[4 * tl.j], c1 = state[4 * tl.j + 1];
  if (state[4 * tl.j + 2] >= 0 || c1 >= J.w || c1 <= c0 || tl.e <= c1) return;
  const int64_t m = J.m, w = J.w;
  T* own = val + J.voff;
  ldl_update(own, m, own, m, c0, (int)(c1 - c0), piv + J.first + c0, c1, m - c1, w - c1, nullptr, true, c1,
             tl.b > c1 ? tl.b : c1, tl.e, tl.rb, tl.re);
}

// D's blocks for the inertia: out[2 c] = D(c, c), out[2 c + 1] = D(c + 1, c) at the first column
// of a 2 x 2 pivot (else 0), one workgroup per supernode.
__global__ void k_ldlt_diag(const LdlSupernode* __restrict__ S, const double* __restrict__ val,
                            const int8_t* __restrict__ piv, double* __restrict__ out) {
  const LdlSupernode J = S[blockIdx.x];
  const double* t = val + J.voff;
  for (int64_t c = threadIdx.x; c < J.w; c += blockDim.x) {
    out[2 * (J.first + c)] = LDL_AT(t, J.m, c, c);
    out[2 * (J.first + c) + 1] = piv[J.first + c] == PIV_FIRST ? LDL_AT(t, J.m, c + 1, c) : 0.0;
  }
}

// dst[i] = src[perm[i]] (into postorder) or, inverse, dst[perm[i]] = src[i] (back).
__global__ void k_ldlt_permute(int64_t n, const int64_t* __restrict__ perm, const double* __restrict__ src,
                               double* __restrict__ dst, bool inverse) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (inverse) dst[perm[i]] = src[i];
  else dst[i] = src[perm[i]];
}

// The forward pass of the solve (solve_group, cholesky.rs:904-981, with LDL^T's unit L, block
// diagonal D and pivot orders) level by level on one right-hand side x (postorder, in place).
// First the descendants' contributions to one level's unknowns, read as a gather: workgroup t
// takes the columns [b, e) of supernode j and, update by update in increasing descendant order
// (the order the one-pass sweep applies them), subtracts L_k(p1 + i, q) z_k(q) for each column q
// of descendant k in order; z holds the descendants' forward values before D^-1.
template <typename T>
__global__ __launch_bounds__(256) void k_ldlt_fwd_gather(const LdlTile* __restrict__ tiles, const LdlSupernode* __restrict__ S,
                                                         const LdlUpdate* __restrict__ U, const int32_t* __restrict__ rel,
                                                         const T* __restrict__ val, const T* __restrict__ z,
                                                         T* __restrict__ x) {
#pragma clang fp contract(off)
  const LdlTile tl = tiles[blockIdx.x];
  const LdlSupernode J = S[tl.j];
  for (int64_t u = J.upd0; u < J.upd1; ++u) {
    const LdlUpdate up = U[u];
    const LdlSupernode K = S[up.k];
    const int32_t* rl = rel + up.rel;  // rl[i]: the column of J that row p1 + i of K lands on
    const int64_t cnt = up.p2 - up.p1;
    int64_t ia = 0, iz = cnt, hi = cnt;
    while (ia < hi) { const int64_t mid = (ia + hi) / 2; if (rl[mid] < tl.b) ia = mid + 1; else hi = mid; }
    hi = cnt;
    iz = ia;
    while (iz < hi) { const int64_t mid = (iz + hi) / 2; if (rl[mid] < tl.e) iz = mid + 1; else hi = mid; }
    const T* lk = val + K.voff;
    const T* zk = z + K.first;
    for (int64_t i = ia + threadIdx.x; i < iz; i += blockDim.x) {
      T* xi = x + J.first + rl[i];
      T v = *xi;
      int64_t at = up.p1 + i;  // L_k(p1 + i, 0); column q + 1's entry lies K.m - q - 1 on
      for (int64_t q = 0; q < K.w; ++q) {
        v -= lk[at] * zk[q];
        at += K.m - q - 1;
      }
      *xi = v;
    }
    __syncthreads();
  }
}

// Then each supernode of the level (one workgroup each): its own unknowns in pivot order, the
// unit lower triangle's forward substitution, z = the values before D^-1, D^-1 (1 x 1 and 2 x 2
// blocks), the result left in pivot order in x. y is a work vector of n entries.
template <typename T>
__global__ __launch_bounds__(256) void k_ldlt_fwd_own(const int64_t* __restrict__ sn, const LdlSupernode* __restrict__ S,
                                                      const T* __restrict__ val, const int8_t* __restrict__ piv,
                                                      const int32_t* __restrict__ order, T* __restrict__ x,
                                                      T* __restrict__ ywork, T* __restrict__ z) {
#pragma clang fp contract(off)
  const LdlSupernode K = S[sn[blockIdx.x]];
  const int64_t m = K.m, w = K.w;
  const T* t = val + K.voff;
  const int8_t* pv = piv + K.first;
  const int32_t* od = order + K.first;
  T* own = x + K.first;
  T* y = ywork + K.first;
  T* zk = z + K.first;
  for (int64_t c = threadIdx.x; c < w; c += blockDim.x) y[c] = own[od[c]];
  __syncthreads();
  for (int64_t q = 0; q < w; ++q) {
    const int64_t end = q + (pv[q] == PIV_FIRST ? 2 : 1);
    const T yq = y[q];
    for (int64_t r = end + threadIdx.x; r < w; r += blockDim.x) y[r] -= LDL_AT(t, m, r, q) * yq;
    __syncthreads();
  }
  for (int64_t c = threadIdx.x; c < w; c += blockDim.x) zk[c] = y[c];
  __syncthreads();
  for (int64_t c = threadIdx.x; c < w; c += blockDim.x) {
    if (pv[c] == PIV_FIRST) {
      const T d11 = LDL_AT(t, m, c, c), d21 = LDL_AT(t, m, c + 1, c), d22 = LDL_AT(t, m, c + 1, c + 1);
      const T det = d11 * d22 - d21 * d21;
      const T p = y[c], q2 = y[c + 1];
      y[c] = (d22 * p - d21 * q2) / det;
      y[c + 1] = (d11 * q2 - d21 * p) / det;
    } else if (pv[c] == PIV_ONE) {
      y[c] /= LDL_AT(t, m, c, c);
    }
  }
  __syncthreads();
  for (int64_t c = threadIdx.x; c < w; c += blockDim.x) own[c] = y[c];
}

// The column of the backward pass at pivot position c: it finishes unknown q from the rows
// end.. (cholesky.rs:982-1030 with 2 x 2 pivots).
__device__ __forceinline__ void ldl_bwd_col(const int8_t* pv, int64_t c, int64_t& q, int64_t& end) {
  if (pv[c] == PIV_SECOND && c > 0) {
    q = c - 1;
    end = c + 1;
  } else if (pv[c] == PIV_FIRST) {
    q = c + 1;
    end = c + 2;
  } else {
    q = c;
    end = c + 1;
  }
}

// sum - b[0] - b[1] - ... - b[cnt - 1] in that order, the next 16 operands read from LDS while the
// current 16 are used; b is 16-byte aligned and readable 16 entries past cnt.
__device__ __forceinline__ double ldl_chain(double sum, const double* b, int cnt) {
#pragma clang fp contract(off)
  const int nb = cnt / 16;
  const double2* p2 = reinterpret_cast<const double2*>(b);
  double2 cur[8], nxt[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) cur[u] = p2[u];
  for (int bt = 0; bt < nb; ++bt) {
#pragma unroll
    for (int u = 0; u < 8; ++u) nxt[u] = p2[(bt + 1) * 8 + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      sum -= cur[u].x;
      sum -= cur[u].y;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) cur[u] = nxt[u];
  }
  for (int r = nb * 16; r < cnt; ++r) sum -= b[r];
  return sum;
}

// The complex chain: sum - b[0] - ... - b[cnt - 1] by parts, one element per 16-byte read.
__device__ __forceinline__ zd ldl_chain(zd sum, const zd* b, int cnt) {
#pragma clang fp contract(off)
  const int nb = cnt / 8;
  const double2* p2 = reinterpret_cast<const double2*>(b);
  double2 cur[8], nxt[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) cur[u] = p2[u];
  for (int bt = 0; bt < nb; ++bt) {
#pragma unroll
    for (int u = 0; u < 8; ++u) nxt[u] = p2[(bt + 1) * 8 + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      sum.re -= cur[u].x;
      sum.im -= cur[u].y;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) cur[u] = nxt[u];
  }
  for (int r = nb * 8; r < cnt; ++r) sum = sum - b[r];
  return sum;
}

constexpr int LDL_BW_CAP = 3840;  // products of one column held in LDS (two such buffers)

// The backward pass (cholesky.rs:982-1030) for the supernodes of one level (sn[blockIdx.x]),
// levels from the root down: a supernode's unknowns are its own and its sums read only its
// ancestors' (final) unknowns, so the supernodes of a level are independent. Each sum runs in
// the reference's order on one lane. With m <= cap (<= LDL_BW_CAP), waves 1.. stage the next column's
// products while lane 0 runs the current column's sum: the one product that needs the unknown
// being finished (its row is the current column's q) is left to lane 0, which forms it first.
// Larger supernodes stage each column in chunks before its sum.
template <typename T>
__global__ __launch_bounds__(256) void k_ldlt_solve_bwd(const int64_t* __restrict__ sn, const LdlSupernode* __restrict__ S,
                                                       const int64_t* __restrict__ rows_all, const T* __restrict__ val,
                                                       const int8_t* __restrict__ piv, const int32_t* __restrict__ order,
                                                       T* __restrict__ x, T* __restrict__ ywork, int cap) {
#pragma clang fp contract(off)
  const LdlSupernode K = S[sn[blockIdx.x]];
  const int64_t m = K.m, w = K.w;
  const T* t = val + K.voff;
  const int8_t* pv = piv + K.first;
  const int32_t* od = order + K.first;
  const int64_t* rows = rows_all + K.rows;
  T* own = x + K.first;
  T* y = ywork + K.first;
  if (cap < 16 || cap > LDL_BW_CAP) cap = LDL_BW_CAP;
  __shared__ __align__(16) T P[2][LDL_BW_CAP + 16];
  __shared__ T Lr[2];      // L(late row, q) of the staged column
  __shared__ int64_t late[2];   // position of its late product, or -1
  __shared__ T acc;
  for (int64_t c = threadIdx.x; c < w; c += blockDim.x) y[c] = own[c];
  __syncthreads();
  if (m <= cap && blockDim.x > 64) {
    {
      int64_t q, end;
      ldl_bwd_col(pv, w - 1, q, end);  // end = w: every product reads an ancestor's unknown
      const T* col = t + ldl_offset(m, q) - q;
      for (int64_t r = end + threadIdx.x; r < m; r += blockDim.x) P[0][r - end] = col[r] * (r < w ? y[r] : x[rows[r]]);
      if (threadIdx.x == 0) late[0] = -1;
    }
    __syncthreads();
    for (int64_t c = w - 1; c >= 0; --c) {
      const int cur = (int)((w - 1 - c) & 1), nx = cur ^ 1;
      int64_t q, end;
      ldl_bwd_col(pv, c, q, end);
      if (threadIdx.x == 0) {
        if (late[cur] >= 0) P[cur][late[cur]] = Lr[cur] * y[end + late[cur]];
        y[q] = ldl_chain(y[q], P[cur], (int)(m - end));
      } else if (threadIdx.x >= 64 && c > 0) {
        int64_t qn, en;
        ldl_bwd_col(pv, c - 1, qn, en);
        const T* col = t + ldl_offset(m, qn) - qn;
        for (int64_t r = en + (threadIdx.x - 64); r < m; r += blockDim.x - 64)
          if (r != q) P[nx][r - en] = col[r] * (r < w ? y[r] : x[rows[r]]);
        if (threadIdx.x == 64) {
          if (q >= en) {
            late[nx] = q - en;
            Lr[nx] = col[q];
          } else {
            late[nx] = -1;
          }
        }
      }
      __syncthreads();
    }
  } else {
    for (int64_t c = w - 1; c >= 0; --c) {
      int64_t q, end;
      ldl_bwd_col(pv, c, q, end);
      const T* col = t + ldl_offset(m, q) - q;  // col[r] = L(r, q)
      if (threadIdx.x == 0) acc = y[q];
      for (int64_t r0 = end; r0 < m; r0 += cap) {
        const int64_t r1 = r0 + cap < m ? r0 + cap : m;
        for (int64_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) P[0][r - r0] = col[r] * (r < w ? y[r] : x[rows[r]]);
        __syncthreads();
        if (threadIdx.x == 0) acc = ldl_chain(acc, P[0], (int)(r1 - r0));
        __syncthreads();
      }
      if (threadIdx.x == 0) y[q] = acc;
      __syncthreads();
    }
  }
  for (int64_t c = threadIdx.x; c < w; c += blockDim.x) own[od[c]] = y[c];
}

}  // namespace fc
