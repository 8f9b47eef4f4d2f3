// Dense kernels of a front's partial factorization on the matrix cores (gfx950).
//
// Every GEMM-shaped step of the numeric phase is one item of one kernel, k_gemm:
//   panel update   L[r.., c0..c1) -= L[r.., 0..c0) L[c0..c1, 0..c0)^T      (G_SUB)
//   rows below     L[r.., c0..c1)  = L[r.., c0..c1) Linv^T                   (G_SET, in place)
//   update matrix  U[i.., j..)     = -L[W+i.., 0..W) L[W+j.., 0..W)^T        (G_NEG)
// where Linv is the inverse of the panel's factored diagonal block (k_potrf_blk). A launch
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

enum GemmMode : int32_t { G_SUB = 0, G_NEG = 1, G_SET = 2, G_PART = 3, G_ADD = 4 };
enum GemmSrc : int8_t { S_L = 0, S_U = 1, S_INV = 2, S_PART = 3, S_W = 4 };

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

// SPLIT (FP32, in-place rows-below products of the mixed form): the epilogue also writes the
// result's scaled BF16 copies (see k_gemm_b16), so a finished panel needs no pass of its own.
template <typename T, bool SPLIT = false>
__global__ __launch_bounds__(256, 2) void k_gemm(const GemmItem* __restrict__ items, T* L, T* U, T* Inv, T* Part, T* W,
                                                 const FrontDev* __restrict__ fronts, __bf16* __restrict__ Lh,
                                                 __bf16* __restrict__ Ll, const int32_t* __restrict__ fexp) {
  constexpr int PAD = 4;
  __shared__ T As[2][GK][GB + PAD];
  __shared__ T Bs[2][GK][GB + PAD];
  const GemmItem it = items[blockIdx.x];
  auto base = [&](int8_t src) {
    return src == S_L ? L : (src == S_U ? U : (src == S_INV ? Inv : (src == S_W ? W : Part)));
  };
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
  // By row blocks: G_SUB and G_ADD read the block's 16 old values first, every load in flight at
  // once, through a buffer resource (a tile spans far below 2 GiB of its array; an entry outside
  // the tile reads 0 from the range check and is not stored). An update and its load in one
  // branch made the compiler wait for each load in turn: 64 round trips a tile.
  const bool rmw = it.mode == G_SUB || it.mode == G_ADD;
  const auto rc = raw_rsrc(C);
  const uint32_t ldcb = (uint32_t)it.ldc * sizeof(T);
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    T old[4][4] = {};
    if (rmw) {
#pragma unroll
      for (int bj = 0; bj < 4; ++bj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<T>::row(lane, v);
          const bool in = i < it.m && j < it.n && i - j + it.diag >= 0;
          old[bj][v] = buf_load<T>(rc, in ? (uint32_t)(i * sizeof(T)) : 0x80000000u, (uint32_t)j * ldcb);
        }
    }
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<T>::row(lane, v);
        if (i < it.m && j < it.n && i - j + it.diag >= 0) {
          const T x = acc[bi][bj][v];
          C[(int64_t)j * it.ldc + i] =
              it.mode == G_SUB ? old[bj][v] - x : (it.mode == G_ADD ? old[bj][v] + x : (it.mode == G_NEG ? -x : x));
        }
      }
  }
  if constexpr (SPLIT) {
    static_assert(sizeof(T) == 4, "copies of an FP32 result");
    // a lane's four values of a block are consecutive columns of one row: one 8-byte store a
    // copy. Columns from n up to the next multiple of 32 (the last panel of a front) are zeros
    // (the product's B rows past n load as zeros).
    const FrontDev F = fronts[it.front];
    const float sc = ldexpf(1.0f, -fexp[it.front]);
    const int64_t oc = it.c - F.lpos;
    const int row0 = (int)(oc % F.ld), col0 = (int)(oc / F.ld);
    const int ncp = (col0 + it.n + 31) / 32 * 32 - col0;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj) {
        const int i = wi + 16 * bi + il, j = wj + 16 * bj + 4 * (lane >> 4);
        if (i < it.m && j < ncp) {
          __bf16 h[4], l[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float x = (float)acc[bi][bj][v] * sc;
            h[v] = (__bf16)x;
            l[v] = (__bf16)(x - (float)h[v]);
          }
          const int col = col0 + j;
          const int64_t o = F.bpos + ((int64_t)(col / 32) * F.ld + row0 + i) * 32 + col % 32;
          *(uint2*)(Lh + o) = *(const uint2*)h;
          *(uint2*)(Ll + o) = *(const uint2*)l;
        }
      }
  }
}

// A tile's entries C(i, j) = A(i, j) (i < m <= GB, j < n <= GB): the rows below a panel saved
// before the product that overwrites them (the FP64 TRSM refines its result against them).
// Wave v copies columns v, v + 4, ...; lane l rows l and l + 64 of each, so every wave load and
// store covers 64 consecutive entries of one column. CB columns' loads are issued before any
// store (rows and columns past the tile load its last entry, which is not stored).
template <typename T>
__global__ __launch_bounds__(256) void k_copy_tiles(const GemmItem* __restrict__ items, const T* __restrict__ L,
                                                    T* __restrict__ W) {
  constexpr int CB = 8;
  static_assert(GB <= 128, "two rows a lane cover a tile's rows");
  const GemmItem it = items[blockIdx.x];
  const T* A = L + it.a;
  T* C = W + it.c;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j0 = wv; j0 < it.n; j0 += 4 * CB) {
    T v[CB][2];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        v[c][h] = A[(int64_t)min(j0 + 4 * c, it.n - 1) * it.lda + min(lane + 64 * h, it.m - 1)];
#pragma unroll
    for (int c = 0; c < CB; ++c)
#pragma unroll
      for (int h = 0; h < 2; ++h)
        if (j0 + 4 * c < it.n && lane + 64 * h < it.m) C[(int64_t)(j0 + 4 * c) * it.ldc + lane + 64 * h] = v[c][h];
  }
}

// ------------------------------------------------------------------------------------------
// The Schur products of the mixed form on the 16x16x32 BF16 matrix cores with FP32
// accumulation. Both operands of a product come from one front, whose columns are scaled by
// that front's power of two 2^-fexp[front] (which bounds them by 1: |L_ij| <= sqrt(A_ii) for an
// SPD matrix); the result is scaled back exactly. With TERMS = 3 each scaled value x is carried
// as hi = bf16(x) and lo = bf16(x - hi) and a product is hi*hi + hi*lo + lo*hi, about 16
// significant bits; TERMS = 1 keeps hi*hi.
//
// The BF16 copies are written once per panel, by the kernels that finish it (the diagonal
// block's by k_potrf_blk, the rows below by the in-place k_gemm<float, true>), in a k-chunked
// layout: entry (i, j) at bpos + ((j / 32) ld + i) 32 + j % 32, zero past the front's width up to
// a multiple of 32, so a 128-row x 32-column stage of an operand is one contiguous 8 KB block
// per copy and a product's last stage needs no bounds. The product kernel moves it into LDS with the
// direct global-to-LDS loads (no registers on the way), 1 KB (16 rows) per wave instruction,
// while the previous stage is multiplied. Tiles, modes and split-k pieces are those of k_gemm
// (128 x 128, 4 waves of 64 x 64).
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
constexpr int BK = 32;  // k per stage: one MFMA depth, one chunk of the copies

// LDS address of a shared-memory pointer, and a 16-byte LDS read issued as an instruction the
// compiler does not track: it would otherwise wait for every direct global-to-LDS load in
// flight before any LDS read. The caller waits (lgkm_wait) before using the values.
__device__ __forceinline__ uint32_t lds_addr(const void* p) {
  return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void*)p;
}
template <int OFF>
__device__ __forceinline__ v8bf lds_read16(uint32_t a) {
  v8bf r;
  asm volatile("ds_read_b128 %0, %1 offset:%2" : "=v"(r) : "v"(a), "i"(OFF));
  return r;
}

// A product's tile with both operands staged through LDS by direct loads, two stages deep (64
// KB for three terms: two workgroups share a CU). Stage s lives in buffer s % 2: it is loaded
// while stage s - 1 is multiplied, then the workgroup waits for its own loads and meets at a
// barrier (which also frees buffer (s - 1) % 2 for stage s + 1).
//
// Stage buffer: copy t of A (128 rows x 32 k), then copy t of B. A row's four 16-byte chunks sit
// in the order chunk ^ (row / 2 % 4), the arrangement under which the fragment reads measure
// free of bank conflicts on gfx950; a wave's direct load fills 1 KB = 16 rows, lane l taking
// the chunk of LDS slot l.
// Rows past the tile's m (n) load the last row of the tile (in bounds); their products land in
// entries that are not stored. The copies are zero past the front's width, so the last stage
// needs no k bound.
template <int TERMS>
__global__ __launch_bounds__(256, 2) void k_gemm_b16(const GemmItem* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                     const __bf16* __restrict__ Lh, const __bf16* __restrict__ Ll,
                                                     float* L, float* U, float* Part, const int32_t* __restrict__ fexp) {
  static_assert(TERMS == 1 || TERMS == 3, "one or three BF16 products");
  constexpr int NS = TERMS == 3 ? 2 : 1;       // hi (and lo) copies of each operand
  constexpr int CE = GB * BK;                  // elements of one copy of one operand per stage
  constexpr int SE = 2 * NS * CE;              // elements of a stage buffer
  constexpr int PW = 2 * NS * (GB / 16) / 4;   // 1 KB direct loads per wave per stage
  __shared__ __attribute__((aligned(16))) __bf16 S[2 * SE];
  const GemmItem it = items[blockIdx.x];
  const FrontDev F = fronts[it.front];
  const int e = fexp[it.front];
  // the operands' first rows and columns in the front, and their origins in the copies
  const int64_t oa = it.a - F.lpos, ob = it.b - F.lpos;
  const int64_t chunk = (int64_t)F.ld * BK;  // elements from one k-chunk of a row to the next
  const int64_t a0 = F.bpos + (oa / F.ld) / BK * chunk + (oa % F.ld) * BK;
  const int64_t b0 = F.bpos + (ob / F.ld) / BK * chunk + (ob % F.ld) * BK;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wi = (wave >> 1) * 64, wj = (wave & 1) * 64;
  const int lrow = lane >> 2, lch = (lane & 3) ^ ((lrow >> 1) & 3);  // the row and chunk slot l holds
  auto issue = [&](int s, int buf) {
#pragma unroll
    for (int q = 0; q < PW; ++q) {
      const int g = wave + 4 * q;            // which 1 KB block of the stage: operand, copy, rows
      const int op = g / (NS * GB / 16), t = g / (GB / 16) % NS, rb = g % (GB / 16);
      const int row = rb * 16 + lrow;
      const int64_t src = (op == 0 ? a0 + (int64_t)min(row, it.m - 1) * BK
                                   : b0 + (int64_t)min(row, it.n - 1) * BK) + s * chunk + lch * 8;
      const __bf16* gp = (t == 0 ? Lh : Ll) + src;
      __bf16* lp = S + buf * SE + (op * NS + t) * CE + rb * 16 * BK;
      __builtin_amdgcn_global_load_lds((const void*)gp, (__attribute__((address_space(3))) void*)lp, 16, 0, 0);
    }
  };
  v4f acc[4][4];
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 4; ++bj) acc[bi][bj] = v4f{0, 0, 0, 0};
  const int stages = (it.k + BK - 1) / BK;
  // a lane's fragment: row lane % 16 of a 16-row block, k chunk lane / 16 (byte offset in a copy)
  const int fr = lane & 15;
  const uint32_t foff = (uint32_t)(fr * 4 + ((lane >> 4) ^ ((fr >> 1) & 3))) * 16;
  const uint32_t sbase = lds_addr(S);
  auto mul = [&](int buf) {
    const uint32_t ah = sbase + (uint32_t)(buf * SE + wi * BK) * 2 + foff;
    const uint32_t bh = sbase + (uint32_t)(buf * SE + NS * CE + wj * BK) * 2 + foff;
    v8bf fi[4], fj[4];
    fi[0] = lds_read16<0>(ah); fi[1] = lds_read16<1024>(ah); fi[2] = lds_read16<2048>(ah); fi[3] = lds_read16<3072>(ah);
    fj[0] = lds_read16<0>(bh); fj[1] = lds_read16<1024>(bh); fj[2] = lds_read16<2048>(bh); fj[3] = lds_read16<3072>(bh);
    if (TERMS == 3) {
      const uint32_t al = ah + CE * 2, bl = bh + CE * 2;
      v8bf gi[4], gj[4];
      gi[0] = lds_read16<0>(al); gi[1] = lds_read16<1024>(al); gi[2] = lds_read16<2048>(al); gi[3] = lds_read16<3072>(al);
      gj[0] = lds_read16<0>(bl); gj[1] = lds_read16<1024>(bl); gj[2] = lds_read16<2048>(bl); gj[3] = lds_read16<3072>(bl);
      asm volatile("s_waitcnt lgkmcnt(0)"  // the fragments are in registers
                   : "+v"(fi[0]), "+v"(fi[1]), "+v"(fi[2]), "+v"(fi[3]), "+v"(fj[0]), "+v"(fj[1]), "+v"(fj[2]),
                     "+v"(fj[3]), "+v"(gi[0]), "+v"(gi[1]), "+v"(gi[2]), "+v"(gi[3]), "+v"(gj[0]), "+v"(gj[1]),
                     "+v"(gj[2]), "+v"(gj[3]));
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(gj[bj], fi[bi], acc[bi][bj], 0, 0, 0);
#pragma unroll
      for (int bi = 0; bi < 4; ++bi)
#pragma unroll
        for (int bj = 0; bj < 4; ++bj)
          acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fj[bj], gi[bi], acc[bi][bj], 0, 0, 0);
    } else {
      asm volatile("s_waitcnt lgkmcnt(0)"
                   : "+v"(fi[0]), "+v"(fi[1]), "+v"(fi[2]), "+v"(fi[3]), "+v"(fj[0]), "+v"(fj[1]), "+v"(fj[2]),
                     "+v"(fj[3]));
    }
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < 4; ++bj)
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fj[bj], fi[bi], acc[bi][bj], 0, 0, 0);
  };
  if (stages > 0) issue(0, 0);
  for (int s = 0; s < stages; ++s) {
    // this wave's loads of stage s are complete, its LDS reads of stage s - 1 too; after the
    // barrier every wave's are
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (s + 1 < stages) issue(s + 1, (s + 1) & 1);
    mul(s & 1);
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
  float* C = (it.csrc == S_L ? L : (it.csrc == S_U ? U : Part)) + it.c;
  // by row blocks, G_SUB's 16 old values of a block loaded first (as in k_gemm)
  const auto rc = raw_rsrc(C);
  const uint32_t ldcb = (uint32_t)it.ldc * sizeof(float);
#pragma unroll
  for (int bi = 0; bi < 4; ++bi) {
    float old[4][4] = {};
    if (it.mode == G_SUB) {
#pragma unroll
      for (int bj = 0; bj < 4; ++bj)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<float>::row(lane, v);
          const bool in = i < it.m && j < it.n && i - j + it.diag >= 0;
          old[bj][v] = buf_load<float>(rc, in ? (uint32_t)(i * sizeof(float)) : 0x80000000u, (uint32_t)j * ldcb);
        }
    }
#pragma unroll
    for (int bj = 0; bj < 4; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = wi + 16 * bi + il, j = wj + 16 * bj + Mma<float>::row(lane, v);
        if (i < it.m && j < it.n && i - j + it.diag >= 0) {
          const float x = acc[bi][bj][v] * unscale;
          C[(int64_t)j * it.ldc + i] = it.mode == G_SUB ? old[bj][v] - x : (it.mode == G_NEG ? -x : x);
        }
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
    T* p = C + (int64_t)j * it.ldc + i;
    const T old = it.mode == G_SUB ? *p : T(0);
    // the pieces in order, eight loads in flight at a time (one round trip per piece otherwise)
    T s = 0;
    int q = 0;
    for (; q + 8 <= it.parts; q += 8) {
      T v[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) v[t] = Part[it.part0 + (int64_t)(q + t) * GB * GB + e];
#pragma unroll
      for (int t = 0; t < 8; ++t) s += v[t];
    }
    for (; q < it.parts; ++q) s += Part[it.part0 + (int64_t)q * GB * GB + e];
    *p = it.mode == G_SUB ? old - s : -s;
  }
}

// ------------------------------------------------------------------------------------------
// The diagonal block of a panel: factored column by column with the reference's pivot test
// (cholesky.rs:1146-1171: a pivot not above 1e-10 |A_jj|, or not finite, fails the column),
// then inverted, Linv = L11^-1, for the rows below (a GEMM). The arithmetic is FP64 whatever
// the storage. Item: f, a = c0; Linv goes to the panel's slot in Inv (F.inv0 + (c0 / NB) NB^2).

constexpr double PIVOT_TOL = 1e-10;  // cholesky.rs:48

// ------------------------------------------------------------------------------------------
// The diagonal block of a panel (any width up to NB): it splits into halves of h0 = 64 and
// h1 = pw - 64 columns (h1 = 0 for panels up to 64 wide),
//   [A11    ]   [L11    ] [L11^T L21^T]          [X11    ]
//   [A21 A22] = [L21 L22] [      L22^T],  Linv = [X21 X22],  X21 = -X22 L21 X11,
// Each half is factored and inverted inside one wave (wave_potrf_inv): the column values pass
// between lanes by shuffles, so the 128 sequential column steps need no workgroup barrier and
// wait on no memory. Everything else is a 64 x 64 x 64 product on the FP64 MFMA with its
// operands read from LDS (strip64, a 64 x 16 strip per wave):
// L21 = A21 X11^T, A22 -= L21 L21^T, T = L21 X11, X21 = -X22 T.
//
// LDS holds the block in FP64: S[i][j] = L(i, j) for i >= j and Linv(j, i) for i < j, the
// inverse's diagonal in xd. The pivot test is the reference's. Outputs: L11 over the block's lower
// triangle in L, and Linv (pw x pw, column-major, zeros above the diagonal) in its slot of Inv.

// Wave-local Cholesky and inverse of the h x h block at S[o.., o..] (h <= 64), the wave as an
// 8 x 8 lane grid: lane (ti, tk) = (lane % 8, lane / 8) holds a[r][c] = A(ti + 8 r, tk + 8 c)
// and x[r][c] = X(ti + 8 r, tk + 8 c). Step j scales column j of L and broadcasts it by
// shuffles (the trailing update is then local), and takes the same step of the inverse,
// right-looking on L X = I: row j of X is divided by L(j, j), then X(i, :) -= L(i, j) X(j, :)
// for i > j. Lane 0 applies the pivot test and records the first failing column.
// Steps 8 CB .. 8 CB + 7 (a template, so that every register index is a constant). Only the
// 8 x 8 register blocks the step can change are touched: (r, c) with r >= c >= CB for L and
// r >= CB, c <= CB for X; within them the remaining conditions are per-lane selects, not
// branches.
// A double from one lane (uniform index) to the whole wave, by scalar reads of its halves.
__device__ __forceinline__ double readlane_d(double v, int l) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((int)(unsigned)b, l);
  const unsigned hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// EXACT (FP64 storage): d = sqrt(p) and column / d, correctly rounded like the reference.
// Otherwise (FP32 storage) the step takes rd from the hardware reciprocal square root with two
// Newton steps and multiplies: the values are rounded to FP32 afterwards, and the step's
// dependent chain (the part that sets this kernel's time) is shorter.
template <int CB, bool EXACT, int NBK>
__device__ __forceinline__ void wpi_steps(double (&a)[NBK][NBK], double (&x)[NBK][NBK], int h, int lane, double dl,
                                          int64_t col, unsigned long long* __restrict__ fail, bool& bad) {
  const int ti = lane & 7, tk = lane >> 3;
  for (int cc = 0; cc < 8; ++cc) {
    const int j = 8 * CB + cc;
    if (j >= h) return;
    const double p = readlane_d(a[CB][CB], cc + 8 * cc);  // the pivot, on the step's dependent chain
    const double a0 = readlane_d(dl, j);
    if (lane == 0 && !bad && (!(p > PIVOT_TOL * a0) || !isfinite(p))) {
      atomicMin(fail, (unsigned long long)(col + j));
      bad = true;
    }
    double d, rd;
    if constexpr (EXACT) {
      d = sqrt(p);
      rd = 1.0 / d;
    } else {
      const double y = __builtin_amdgcn_rsq(p), hp = 0.5 * p;
      const double y1 = y * fma(-hp * y, y, 1.5);
      rd = y1 * fma(-hp * y1, y1, 1.5);
      d = p * rd;
    }
    {  // column j of L: rows below scaled, the pivot replaced by d (lanes tk == cc own it)
      const bool own = tk == cc;
#pragma unroll
      for (int r = CB; r < NBK; ++r) {
        const bool below = r > CB || ti > cc, diag = r == CB && ti == cc;
        const double sc = EXACT ? a[r][CB] / d : a[r][CB] * rd;
        a[r][CB] = own ? (below ? sc : (diag ? d : a[r][CB])) : a[r][CB];
      }
    }
    {  // row j of X divided by L(j, j) (lanes ti == cc own it)
      const bool own = ti == cc;
#pragma unroll
      for (int c = 0; c <= CB; ++c) x[CB][c] = own ? x[CB][c] * rd : x[CB][c];
    }
    double li[NBK], lk[NBK], xj[NBK];
#pragma unroll
    for (int r = CB; r < NBK; ++r) li[r] = __shfl(a[r][CB], ti + 8 * cc);  // L(ti + 8r, j)
#pragma unroll
    for (int c = CB; c < NBK; ++c) lk[c] = __shfl(a[c][CB], tk + 8 * cc);  // L(tk + 8c, j)
#pragma unroll
    for (int c = 0; c <= CB; ++c) xj[c] = __shfl(x[CB][c], cc + 8 * tk);  // X(j, tk + 8c)
    const bool kj = tk > cc, ik = ti >= tk, ij = ti > cc;
#pragma unroll
    for (int r = CB; r < NBK; ++r) {
#pragma unroll
      for (int c = CB; c <= r; ++c) {  // L(i, k) -= L(i, j) L(k, j) for i >= k > j
        const bool upd = (c > CB || kj) && (r > c || ik);
        const double t = fma(-li[r], lk[c], a[r][c]);
        a[r][c] = upd ? t : a[r][c];
      }
      const bool below = r > CB || ij;  // X(i, :) -= L(i, j) X(j, :) for i > j
#pragma unroll
      for (int c = 0; c <= CB; ++c) {
        const double t = fma(-li[r], xj[c], x[r][c]);
        x[r][c] = below ? t : x[r][c];
      }
    }
  }
}

// NBK x NBK register blocks a lane, as wave_potrf_inv32 (h <= 8 NBK).
template <bool EXACT, int NBK = 8>
__device__ inline void wave_potrf_inv(double (*S)[NB + 1], double* xd, int o, int h, int lane,
                                      const double* __restrict__ dg, int64_t col,
                                      unsigned long long* __restrict__ fail, bool& bad) {
  const int ti = lane & 7, tk = lane >> 3;
  // the reference diagonal of every column, read once: the pivot test then costs a shuffle
  const double dl = lane < h ? fabs(dg[lane]) : 0.0;
  double a[NBK][NBK], x[NBK][NBK];
#pragma unroll
  for (int r = 0; r < NBK; ++r)
#pragma unroll
    for (int c = 0; c < NBK; ++c) {
      const int i = ti + 8 * r, k = tk + 8 * c;
      const double v = S[o + i][o + k];  // in bounds for every lane (o + 8 NBK <= NB)
      a[r][c] = (i < h && k < h && i >= k) ? v : (i == k ? 1.0 : 0.0);
      x[r][c] = i == k ? 1.0 : 0.0;
    }
  wpi_steps<0, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
  wpi_steps<1, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
  wpi_steps<2, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
  wpi_steps<3, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
  if constexpr (NBK > 4) {
    wpi_steps<4, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
    wpi_steps<5, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
    wpi_steps<6, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
    wpi_steps<7, EXACT, NBK>(a, x, h, lane, dl, col, fail, bad);
  }
  // L below and on the diagonal, X above it (transposed), X's diagonal in xd; every lane
  // stores to in-bounds addresses, the entries outside the h x h block are not written
#pragma unroll
  for (int r = 0; r < NBK; ++r)
#pragma unroll
    for (int c = 0; c < NBK; ++c) {
      const int i = ti + 8 * r, k = tk + 8 * c;
      if (i < h && k < h) {
        if (i >= k) S[o + i][o + k] = a[r][c];
        if (i > k) S[o + k][o + i] = x[r][c];
        if (i == k) xd[o + i] = x[r][c];
      }
    }
}

// The same wave-local Cholesky and inverse in FP32 arithmetic, for the FP32-storage forms
// (their diagonal blocks are FP32 work; the block's values are FP32 already). The steps are
// those of wpi_steps, with the broadcast column zeroed outside the rows and columns a step
// changes (L(i, j) for i <= j, L(k, j) for k <= j): the trailing updates then need no selects,
// an update with a zero weight leaving its entry as it was. Entries above the diagonal of L's
// diagonal register blocks take meaningless (finite) values and are never read or stored.
// A pivot not above PIVOT_FP32 |A_jj| (or not finite) sets fb: the caller then factors the
// whole block again in FP64, where the reference's pivot test decides.
constexpr float PIVOT_FP32 = 1e-3f;

__device__ __forceinline__ float readlane_f(float v, int l) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), l));
}

// NBK x NBK register blocks a lane (the wave covers blocks of up to 8 NBK columns): 8 for the
// 64-column halves, 4 for 32-column quarters, whose steps touch a quarter of the registers.
template <int CB, int NBK>
__device__ __forceinline__ void wpi32_steps(float (&a)[NBK][NBK], float (&x)[NBK][NBK], int h, int lane, float dl,
                                            bool& fb) {
  const int ti = lane & 7, tk = lane >> 3;
  for (int cc = 0; cc < 8; ++cc) {
    const int j = 8 * CB + cc;
    if (j >= h) return;
    const float p = readlane_f(a[CB][CB], cc + 8 * cc);  // the pivot, on the step's dependent chain
    const float a0 = readlane_f(dl, j);
    fb = fb || !(p > PIVOT_FP32 * a0) || !(p < __builtin_huge_valf());
    const float y = __builtin_amdgcn_rsqf(p);
    const float rd = y * fmaf(-0.5f * p * y, y, 1.5f);  // 1 / sqrt(p), one Newton step
    const float fc = tk == cc ? rd : 1.0f, fr = ti == cc ? rd : 1.0f;
#pragma unroll
    for (int r = CB; r < NBK; ++r) a[r][CB] *= fc;  // column j of L (the pivot becomes p rd = L(j, j))
#pragma unroll
    for (int c = 0; c <= CB; ++c) x[CB][c] *= fr;  // row j of X divided by L(j, j)
    float li[NBK], lk[NBK], xj[NBK];
#pragma unroll
    for (int r = CB; r < NBK; ++r) li[r] = __shfl(a[r][CB], ti + 8 * cc);  // L(ti + 8r, j)
#pragma unroll
    for (int c = CB; c < NBK; ++c) lk[c] = __shfl(a[c][CB], tk + 8 * cc);  // L(tk + 8c, j)
#pragma unroll
    for (int c = 0; c <= CB; ++c) xj[c] = __shfl(x[CB][c], cc + 8 * tk);  // X(j, tk + 8c)
    li[CB] = ti > cc ? li[CB] : 0.0f;
    lk[CB] = tk > cc ? lk[CB] : 0.0f;
#pragma unroll
    for (int r = CB; r < NBK; ++r) {
#pragma unroll
      for (int c = CB; c <= r; ++c) a[r][c] = fmaf(-li[r], lk[c], a[r][c]);  // L(i, k) -= L(i, j) L(k, j)
#pragma unroll
      for (int c = 0; c <= CB; ++c) x[r][c] = fmaf(-li[r], xj[c], x[r][c]);  // X(i, :) -= L(i, j) X(j, :)
    }
  }
}

template <int NBK = 8>
__device__ inline void wave_potrf_inv32(double (*S)[NB + 1], double* xd, int o, int h, int lane,
                                        const double* __restrict__ dg, bool& fb) {
  const int ti = lane & 7, tk = lane >> 3;
  const float dl = lane < h ? (float)fabs(dg[lane]) : 0.0f;
  float a[NBK][NBK], x[NBK][NBK];
#pragma unroll
  for (int r = 0; r < NBK; ++r)
#pragma unroll
    for (int c = 0; c < NBK; ++c) {
      const int i = ti + 8 * r, k = tk + 8 * c;
      const float v = (float)S[o + i][o + k];  // in bounds for every lane (o + 8 NBK <= NB)
      a[r][c] = (i < h && k < h && i >= k) ? v : (i == k ? 1.0f : 0.0f);
      x[r][c] = i == k ? 1.0f : 0.0f;
    }
  wpi32_steps<0, NBK>(a, x, h, lane, dl, fb);
  wpi32_steps<1, NBK>(a, x, h, lane, dl, fb);
  wpi32_steps<2, NBK>(a, x, h, lane, dl, fb);
  wpi32_steps<3, NBK>(a, x, h, lane, dl, fb);
  if constexpr (NBK > 4) {
    wpi32_steps<4, NBK>(a, x, h, lane, dl, fb);
    wpi32_steps<5, NBK>(a, x, h, lane, dl, fb);
    wpi32_steps<6, NBK>(a, x, h, lane, dl, fb);
    wpi32_steps<7, NBK>(a, x, h, lane, dl, fb);
  }
#pragma unroll
  for (int r = 0; r < NBK; ++r)
#pragma unroll
    for (int c = 0; c < NBK; ++c) {
      const int i = ti + 8 * r, k = tk + 8 * c;
      if (i < h && k < h) {
        if (i >= k) S[o + i][o + k] = a[r][c];
        if (i > k) S[o + k][o + i] = x[r][c];
        if (i == k) xd[o + i] = x[r][c];
      }
    }
}

// One wave's 16 x 16 tile of a product over k in [0, 32) on the FP64 MFMA:
//   acc[v] = sum_k col(t, k) row(i, k),  i = lane % 16,  t = lane / 16 + 4 v
// where col(t, k) and row(i, k) return in-bounds LDS values for t, i in [0, 16).
template <class FC, class FR>
__device__ __forceinline__ v4d tile16(int lane, FC col, FR row) {
  const int r = lane & 15, q = lane >> 4;
  v4d acc = {};
#pragma unroll
  for (int k0 = 0; k0 < 32; k0 += 4) acc = Mma<double>::step(col(r, k0 + q), row(r, k0 + q), acc);
  return acc;
}

// The h x h block at S[o.., o..] (h <= 64) by the whole workgroup, with the layout the wave
// functions leave: quarters of 32 columns, each factored and inverted inside one wave (4 x 4
// register blocks a lane, whose steps cost under half the 64-column ones'), the products
// between them on the FP64 MFMA, a 16 x 16 tile a wave (tile rows 16 (w % 2), columns
// 16 (w / 2)), operands from LDS, triangular ones as 0/1 weights:
//   L21 = A21 X11^T, A22 -= L21 L21^T, (L22, X22) in one wave, T = L21 X11, X21 = -X22 T.
// T is 32 x 32 of LDS. Every thread calls it; it ends on a barrier.
template <bool F32, bool EXACT>
__device__ inline void half_potrf_inv(double (*S)[NB + 1], double* xd, double (*T)[33], int o, int h, int tid,
                                      const double* __restrict__ dg, bool& fb, int64_t col,
                                      unsigned long long* __restrict__ fail, bool& bad) {
  const int lane = tid & 63, wave = tid >> 6;
  const int h0 = min(32, h), h1 = h - h0, p = o + 32;
  // one quarter's chain in wave 0: FP32 arithmetic (a small pivot sets fb), or FP64 with the
  // reference's pivot test
  auto chain = [&](int oq, int hq, int d) {
    if constexpr (F32) wave_potrf_inv32<4>(S, xd, oq, hq, lane, dg + d, fb);
    else wave_potrf_inv<EXACT, 4>(S, xd, oq, hq, lane, dg + d, col + d, fail, bad);
  };
  if (wave == 0) chain(o, h0, 0);
  __syncthreads();
  if (h1 <= 0) return;  // (uniform) h0 = 32 below
  const int bi = 16 * (wave & 1), bt = 16 * (wave >> 1);  // the wave's tile
  const int ri = lane & 15;
  auto rowq = [&](int i) { return p + min(bi + i, h1 - 1); };  // a row of the second quarter, clamped
  {  // L21(i, t) = sum_{k <= t} A21(i, k) X11(t, k); X11(t, k) = S[o + k][o + t] (t > k), xd[o + t]
    const v4d acc = tile16(
        lane,
        [&](int t, int k) {
          const int c = bt + t;
          return S[o + k][o + c] * (c > k ? 1.0 : 0.0) + xd[o + c] * (c == k ? 1.0 : 0.0);
        },
        [&](int i, int k) { return S[rowq(i)][o + k]; });
    __syncthreads();  // every A21 read before L21 is stored over it
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (bi + ri < h1) S[p + bi + ri][o + bt + (lane >> 4) + 4 * v] = acc[v];
  }
  __syncthreads();
  if (bt <= bi) {  // A22(i, j) -= sum_k L21(i, k) L21(j, k), j <= i (reads the left quarter, writes the right)
    const v4d acc = tile16(
        lane, [&](int t, int k) { return S[p + min(bt + t, h1 - 1)][o + k]; }, [&](int i, int k) { return S[rowq(i)][o + k]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = bi + ri, j = bt + (lane >> 4) + 4 * v;
      if (i < h1 && j <= i) S[p + i][p + j] -= acc[v];
    }
  }
  __syncthreads();
  if (wave == 0) chain(p, h1, 32);
  {  // T(i, c) = sum_{k >= c} L21(i, k) X11(k, c); X11(k, c) = S[o + c][o + k] (k > c), xd[o + c]
    const v4d acc = tile16(
        lane,
        [&](int t, int k) {
          const int c = bt + t;
          return S[o + c][o + k] * (k > c ? 1.0 : 0.0) + xd[o + c] * (k == c ? 1.0 : 0.0);
        },
        [&](int i, int k) { return S[rowq(i)][o + k]; });
#pragma unroll
    for (int v = 0; v < 4; ++v) T[bi + ri][bt + (lane >> 4) + 4 * v] = acc[v];
  }
  __syncthreads();  // T, and the second quarter's factor and inverse, done
  {  // X21(i, c) = -sum_{k <= i} X22(i, k) T(k, c); X22(i, k) = S[p + k][p + i] (i > k), xd[p + i];
     // stored at S[o + c][p + i]. T's rows past h1 hold finite values (clamped rows) times 0
    const v4d acc = tile16(
        lane, [&](int t, int k) { return T[k][bt + t]; },
        [&](int i, int k) {
          const int ii = bi + i;
          return S[p + k][p + ii] * (ii > k ? 1.0 : 0.0) + xd[p + ii] * (ii == k ? 1.0 : 0.0);
        });
#pragma unroll
    for (int v = 0; v < 4; ++v)
      if (bi + ri < h1) S[o + bt + (lane >> 4) + 4 * v][p + bi + ri] = -acc[v];
  }
  __syncthreads();
}

// One wave's 64 x 16 strip of a product over k in [0, 64) on the FP64 MFMA:
//   acc[ib][v] += sum_k col(t, k) row(i, k),  i = 16 ib + lane % 16,  t = (lane / 16) + 4 v
// where col(r, k) and row(i, k) return in-bounds LDS values (col's r is the strip column in
// [0, 16)). Each k-step reads one column value and four row values per lane for four MFMAs.
template <class FC, class FR>
__device__ __forceinline__ void strip64(v4d (&acc)[4], int lane, FC col, FR row) {
  const int r = lane & 15, q = lane >> 4;
#pragma unroll 4
  for (int k0 = 0; k0 < 64; k0 += 4) {
    const int k = k0 + q;
    const double x = col(r, k);
    double y[4];
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) y[ib] = row(16 * ib + r, k);
#pragma unroll
    for (int ib = 0; ib < 4; ++ib) acc[ib] = Mma<double>::step(x, y[ib], acc[ib]);
  }
}

template <typename T>
__global__ __launch_bounds__(256, 1) void k_potrf_blk(const Item* __restrict__ items, const FrontDev* __restrict__ fronts,
                                                      T* __restrict__ L, const double* __restrict__ diag,
                                                      unsigned long long* __restrict__ fail, T* __restrict__ Inv,
                                                      __bf16* __restrict__ Lh, __bf16* __restrict__ Ll,
                                                      const int32_t* __restrict__ fexp) {
  constexpr int H = 64;
  __shared__ double S[NB][NB + 1];
  __shared__ double xd[NB];
  __shared__ int redo;  // the FP32 attempt met a small pivot: factor again in FP64
  __shared__ double Tq[32][33];  // the quarter products' scratch (half_potrf_inv)
  const Item it = items[blockIdx.x];
  const FrontDev F = fronts[it.f];
  const int c0 = it.a, pw = min(NB, F.w - c0);
  const int h0 = min(H, pw), h1 = pw - h0;
  T* base = L + F.lpos + (int64_t)c0 * F.ld + c0;
  const double* dg = diag + F.col0 + c0;
  const int64_t col = (int64_t)F.col0 + c0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rr = tid & (NB - 1), jp = tid >> 7;  // a thread's row, and which of every two columns
  // Mixed form (Lh given): the block's scaled BF16 copies, from the same FP32 values as L (zero
  // above the diagonal and past the block up to the next multiple of 32 columns), 8 columns of a
  // row per 16-byte store. S holds the final L here (the products that follow overwrite L21).
  auto copies = [&]() {
    if (!Lh) return;
    const float sc = ldexpf(1.0f, -fexp[it.f]);
    const int ncp = (c0 + pw + 31) / 32 * 32 - c0;
    for (int e = tid; e < pw * (ncp / 8); e += blockDim.x) {
      const int r = e % pw, g = e / pw;
      v8bf h, l;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int j = 8 * g + q;
        const float x = (j <= r ? (float)(T)S[r][min(j, NB - 1)] : 0.0f) * sc;
        h[q] = (__bf16)x;
        l[q] = (__bf16)(x - (float)h[q]);
      }
      const int col = c0 + 8 * g;
      const int64_t o = F.bpos + ((int64_t)(col / 32) * F.ld + c0 + r) * 32 + col % 32;
      *(v8bf*)(Lh + o) = h;
      *(v8bf*)(Ll + o) = l;
    }
  };
  // the pw x pw square (its upper triangle is zeros in L, in bounds): 64 independent loads,
  // through a buffer resource (the block spans far below 2 GiB of L): an entry outside the square
  // takes the lane offset 2^31, past the resource's 2^31 - 1 bytes, and reads 0 from the buffer
  // unit's range check, with no branch (a branch around each load made the compiler wait for
  // each one in turn: 64 round trips, 28 us of a 73 us block).
  // xd beyond the block's width is read by the products below with a 0 weight: it must hold a
  // finite value, not what an earlier kernel left in LDS (a NaN there times 0 is a NaN)
  const auto rs = raw_rsrc(base);
  const uint32_t ldb = (uint32_t)F.ld * sizeof(T);
  auto load = [&]() {
    if (tid < NB) xd[tid] = 0.0;
    T v[NB / 2];
#pragma unroll
    for (int q = 0; q < NB / 2; ++q) {
      const int j = jp + 2 * q;
      v[q] = buf_load<T>(rs, (rr < pw && j < pw) ? (uint32_t)(rr * sizeof(T)) : 0x80000000u, j * ldb);
    }
#pragma unroll
    for (int q = 0; q < NB / 2; ++q) S[rr][jp + 2 * q] = (double)v[q];
  };
  // each wave: strip columns t0 + [0, 16) of a 64 x 64 product, rows 16 ib + lane % 16
  const int r = lane & 15, q = lane >> 4, t0 = 16 * wave;
  // L11 and X11; for a wide panel then L21 = A21 X11^T, A22 -= L21 L21^T, L22 and X22, each half
  // by the whole workgroup in 32-column quarters. F32: the wave steps in FP32 arithmetic (a
  // small pivot sets redo); otherwise FP64 with the reference's pivot test.
  bool bad = false;
  auto factor = [&](auto f32) {
    constexpr bool F32 = decltype(f32)::value;
    bool fb = false;
    half_potrf_inv<F32, sizeof(T) == 8>(S, xd, Tq, 0, h0, tid, dg, fb, col, fail, bad);
    __syncthreads();
    if (h1 > 0) {
      {  // L21 = A21 X11^T: L21(i, t) = sum_{k <= t} A21(i, k) X11(t, k), X11(t, k) = S[k][t] (t > k)
        v4d acc[4] = {};
        strip64(
            acc, lane,
            [&](int c, int k) {  // 0/1 weights, not selects: the loads stay unconditional
              const int t = t0 + c;
              return S[k][t] * (t > k ? 1.0 : 0.0) + xd[k] * (t == k ? 1.0 : 0.0);
            },
            [&](int i, int k) { return S[h0 + min(i, h1 - 1)][k]; });
        __syncthreads();  // every A21 read before L21 is stored over it
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = 16 * ib + r;
            if (i < h1) S[h0 + i][t0 + q + 4 * v] = acc[ib][v];
          }
      }
      __syncthreads();
      {  // A22 -= L21 L21^T over the lower triangle (reads the left half, writes the right)
        v4d acc[4] = {};
        strip64(
            acc, lane, [&](int c, int k) { return S[h0 + min(t0 + c, h1 - 1)][k]; },
            [&](int i, int k) { return S[h0 + min(i, h1 - 1)][k]; });
#pragma unroll
        for (int ib = 0; ib < 4; ++ib)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = 16 * ib + r, j = t0 + q + 4 * v;
            if (i < h1 && j <= i) S[h0 + i][h0 + j] -= acc[ib][v];
          }
      }
      __syncthreads();
      half_potrf_inv<F32, sizeof(T) == 8>(S, xd, Tq, h0, h1, tid, dg + h0, fb, col + h0, fail, bad);
    }
    if constexpr (F32) {
      if (wave == 0 && lane == 0) redo = fb ? 1 : 0;
    }
    __syncthreads();
  };
  load();
  __syncthreads();
  if constexpr (sizeof(T) == 4) {
    factor(std::true_type{});
    if (redo) {  // uniform: read by every thread after the barrier that ends factor()
      __syncthreads();  // every thread has read redo before S is loaded again
      load();
      __syncthreads();
      factor(std::false_type{});
    }
  } else {
    factor(std::false_type{});
  }
  // L is final: write it out (reads first, then stores), and its copies
  {
    T v[NB / 2];
#pragma unroll
    for (int qq = 0; qq < NB / 2; ++qq) v[qq] = (T)S[rr][jp + 2 * qq];
#pragma unroll
    for (int qq = 0; qq < NB / 2; ++qq) {
      const int j = jp + 2 * qq;
      if (rr < pw && j <= rr) base[(int64_t)j * F.ld + rr] = v[qq];
    }
  }
  copies();
  if (h1 > 0) {
    {  // T = L21 X11: T(i, t) = sum_{k >= t} L21(i, k) X11(k, t), X11(k, t) = S[t][k] (k > t)
      v4d acc[4] = {};
      strip64(
          acc, lane,
          [&](int c, int k) {
            const int t = t0 + c;
            return S[t][k] * (k > t ? 1.0 : 0.0) + xd[k] * (k == t ? 1.0 : 0.0);
          },
          [&](int i, int k) { return S[h0 + min(i, h1 - 1)][k]; });
      __syncthreads();  // L21 written out and read
#pragma unroll
      for (int ib = 0; ib < 4; ++ib)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 16 * ib + r;
          if (i < h1) S[h0 + i][t0 + q + 4 * v] = acc[ib][v];  // T over L21
        }
    }
    __syncthreads();
    {  // X21 = -X22 T, stored at S[t][h0 + ii]: X22(ii, k) = S[h0 + k][h0 + ii] (ii > k), T(k, t) = S[h0 + k][t]
      v4d acc[4] = {};
      strip64(
          acc, lane, [&](int c, int k) { return S[h0 + k][t0 + c]; },
          [&](int ii, int k) {
            return S[h0 + k][h0 + ii] * (ii > k ? 1.0 : 0.0) + xd[h0 + k] * (ii == k ? 1.0 : 0.0);
          });
#pragma unroll
      for (int ib = 0; ib < 4; ++ib)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ii = 16 * ib + r;
          if (ii < h1) S[t0 + q + 4 * v][h0 + ii] = -acc[ib][v];
        }
    }
  }
  __syncthreads();
  {  // the inverse, column-major with ld = pw, zeros above the diagonal: reads, then stores
    T* inv = Inv + F.inv0 + (int64_t)(c0 / NB) * NB * NB;
    T v[NB / 2];
    const double xr = xd[rr];
#pragma unroll
    for (int q = 0; q < NB / 2; ++q) {
      const int j = jp + 2 * q;
      v[q] = (T)(S[j][rr] * (rr > j ? 1.0 : 0.0) + xr * (rr == j ? 1.0 : 0.0));
    }
#pragma unroll
    for (int q = 0; q < NB / 2; ++q) {
      const int j = jp + 2 * q;
      if (rr < pw && j < pw) inv[(int64_t)j * pw + rr] = v[q];
    }
  }
}

}  // namespace fc
