// Device-side data of the numeric phase: front descriptors and work items, shared by the
// factor and solve kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fc {

// One front as the kernels see it. L part: M x W column-major at lpos with leading dimension
// ld (rows 0..W-1 are the front's own columns, rows W..M-1 the rows below). Update matrix:
// (M-W) x (M-W) column-major at upos with leading dimension M-W (lower triangle used).
struct FrontDev {
  int64_t lpos;
  int64_t upos;
  int64_t rel_off;   // relative map of the rows below into the parent front
  int64_t rows_off;  // front rows (postorder column numbers) in the rows array
  int64_t inv0;      // inverses of the diagonal blocks: panel at c0 at inv0 + (c0 / NB) NB^2, ld = its width
  int64_t bpos;      // mixed form: BF16 copies of the L part, (i, j) at bpos + ((j / 32) ld + i) 32 + j % 32
  int32_t ld, w, m, col0;
  int32_t parent, kid0, kid1, level;  // children are kids[kid0..kid1)
};
static_assert(sizeof(FrontDev) == 80, "FrontDev layout");

// Strided runs of loads (and stores) through a buffer resource: the lane's part of the address
// is one VGPR offset and each access's part a scalar, so a run of independent accesses needs no
// 64-bit address per access. Raw resource (byte offsets, stride 0); every access made through
// it is in bounds by construction, the record count only has to cover the lane offsets.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t raw_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0x7fffffff, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_load(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t off);
template <>
__device__ __forceinline__ float buf_load<float>(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t off) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, lane_off, off, 0));
}
template <>
__device__ __forceinline__ double buf_load<double>(__amdgpu_buffer_rsrc_t r, uint32_t lane_off, uint32_t off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, lane_off, off, 0));
}

template <typename T>
__device__ __forceinline__ void buf_store(__amdgpu_buffer_rsrc_t r, T v, uint32_t lane_off, uint32_t off);
template <>
__device__ __forceinline__ void buf_store<float>(__amdgpu_buffer_rsrc_t r, float v, uint32_t lane_off, uint32_t off) {
  __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, lane_off, off, 0);
}
template <>
__device__ __forceinline__ void buf_store<double>(__amdgpu_buffer_rsrc_t r, double v, uint32_t lane_off, uint32_t off) {
  typedef unsigned int u32x2 __attribute__((__vector_size__(2 * sizeof(unsigned int))));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, lane_off, off, 0);
}

// A unit of work of one launch; the meaning of a, b, c, d depends on the kernel.
struct Item {
  int32_t f, a, b, c, d, e, pad0, pad1;
};

enum Kind : int32_t {
  K_EXTEND_L = 0,  // children's update matrices into the L part: a = first column, b = end
  K_EXTEND_U = 1,  // ... into the update matrix (after the Schur product): a, b columns
  K_UPDATE = 2,    // panel update L[i0.., c0..c1) -= L[i0.., 0..c0) L[c0..c1, 0..c0)^T
  K_POTRF = 3,     // diagonal block of a panel
  K_TRSM = 4,      // rows below a panel against its diagonal block
  K_SYRK = 5,      // update matrix U = -L21 L21^T
  K_REDUCE = 6,    // sums of the pieces of a GEMM split along k
  K_SMALL = 7,     // small fronts' whole partial factorizations, one workgroup a front
};

struct Launch {
  int32_t kind;
  int32_t level;
  int64_t item0;  // first item in the item array
  int32_t count;  // items (= workgroups)
  int32_t pad;
  double flops;   // algorithmic flops of the launch (GEMM-like kinds)
  double entries; // algorithmic entries moved (extend-add: child entries added)
  int32_t stream = 0;  // 0: the caller's stream, 1: the factor's side stream (lookahead)
  int32_t slot = 0;    // event slot of a record / wait step
};

constexpr int KINDS = 8;

// A child's run of update-matrix columns landing in an extend-add item's parent columns:
// columns [jc0, jc1) of child front `child`.
struct EaRange {
  int32_t child, jc0, jc1, pad;
};

constexpr int EA_BATCH = 8;   // rows an extend-add thread moves at a time
constexpr int EA_COLS = 8;     // columns of the parent an extend-add window spans
constexpr int EA_ENTRIES = 8192;  // child entries an extend-add workgroup takes (whole windows)

}  // namespace fc
