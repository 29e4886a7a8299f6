// Device helpers shared by the gfx950 kernel sources (ivqm_kernels.hip,
// ivqm_ssim_modes.hip).
#pragma once
#include "ivqm_device.h"

namespace ivqm {
namespace {

// Plane c of a picture. Selects between loaded values instead of indexing
// p[] / stride[] with a run-time c (or selecting between their addresses):
// either keeps a private copy of the table entry in scratch memory.
__device__ __forceinline__ const void* plane_ptr(const PicDev& P, int c) {
  const void* p0 = P.p[0];
  const void* p1 = P.p[1];
  const void* p2 = P.p[2];
  return c == 0 ? p0 : (c == 1 ? p1 : p2);
}
__device__ __forceinline__ int plane_stride(const PicDev& P, int c) {
  const int s0 = P.stride[0], s1 = P.stride[1], s2 = P.stride[2];
  return c == 0 ? s0 : (c == 1 ? s1 : s2);
}

// Sample (x, y) of plane c on the luma grid (chroma at (x >> cs, y >> csy)). Plane offsets are 32-bit with
// 24-bit operands (rows and strides stay below 2^24, planes below 2^32 samples).
// Row y (luma grid) of plane c as a global pointer; with a wave-uniform y the
// pointer lives in scalar registers and a load adds only the lane's column.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* plane_row(const PicDev& P, int c, int y) {
  typedef __attribute__((address_space(1))) const T GT;
  const int s = c ? P.csy : 0;
  return (const GT*)plane_ptr(P, c) + __umul24((uint32_t)(y >> s), (uint32_t)plane_stride(P, c));
}

// Raw buffer resource over plane c of a picture of height H (the exact plane
// size in bytes; data-format word 32-bit). Loads through it take the row as a
// scalar byte offset and the column as a 32-bit vector byte offset, so one
// vector offset serves every row and no address arithmetic is per load.
template <typename T>
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const PicDev& P, int c, int H) {
  const int s = c ? P.csy : 0;
  const uint32_t bytes = (uint32_t)((H + s) >> s) * (uint32_t)plane_stride(P, c) * (uint32_t)sizeof(T);
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(plane_ptr(P, c)), (short)0, (int)bytes, 0x00020000);
}
template <typename T>
__device__ __forceinline__ int ld_buf(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if (sizeof(T) == 2) return (int)__builtin_amdgcn_raw_buffer_load_b16(r, voff, soff, 0);
  return (int)__builtin_amdgcn_raw_buffer_load_b8(r, voff, soff, 0);
}

// Picture planes are device allocations: addressed as global memory they load
// with a scalar base and a 32-bit vector offset (no 64-bit address math).
template <typename T>
__device__ __forceinline__ int ldpix(const PicDev& P, int c, int x, int y) {
  typedef __attribute__((address_space(1))) const T GT;
  const int sx = c ? P.cs : 0, sy = c ? P.csy : 0;
  const GT* p = (const GT*)plane_ptr(P, c);
  return (int)p[__umul24((uint32_t)(y >> sy), (uint32_t)plane_stride(P, c)) + (uint32_t)(x >> sx)];
}

// a * b + c with 24-bit signed operands a, b: the low 32 bits of the product plus c (v_mad_i32_i24).
// Written out because instruction selection otherwise splits the sums of products into separate
// multiplies and three-input adds (five instructions for three terms instead of three).
__device__ __forceinline__ int mad_i24(int a, int b, int c) {
  int d;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "v"(c));
  return d;
}
// The same with a wave-uniform b read from a scalar register (no copy of it into a vector register).
__device__ __forceinline__ int mad_i24_s(int a, int b, int c) {
  int d;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(d) : "v"(a), "s"(b), "v"(c));
  return d;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }  // v_med3_i32

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
// Sum over the 64 lanes with DPP (no LDS traffic), uniform result: prefix
// sums within each 16-lane row (row_shr 1, 2, 4, 8; lanes shifted in from
// outside the row read 0), then row_bcast:15 / row_bcast:31 carry the row
// totals up, and lane 63 holds the wave total.
__device__ __forceinline__ uint32_t wave_sum_u32_dpp(uint32_t v) {
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
  return (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
}
// Three wave sums with their DPP steps interleaved (independent instructions
// between dependent DPP reads instead of wait states).
// FP64 sum over the 64 lanes with DPP moves of both halves (fixed order: row
// prefixes, then row broadcasts); the total is read from lane 63.
__device__ __forceinline__ double wave_sum_f64_dpp(double v) {
#define IVQM_DPPD(ctrl, rmask, bc)                                                                      \
  {                                                                                                    \
    const uint64_t b = __builtin_bit_cast(uint64_t, v);                                                \
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, ctrl, rmask, 0xf, bc); \
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), ctrl, rmask, 0xf, bc); \
    v += __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));                              \
  }
  IVQM_DPPD(0x111, 0xf, true)
  IVQM_DPPD(0x112, 0xf, true)
  IVQM_DPPD(0x114, 0xf, true)
  IVQM_DPPD(0x118, 0xf, true)
  IVQM_DPPD(0x142, 0xa, false)
  IVQM_DPPD(0x143, 0xc, false)
#undef IVQM_DPPD
  const uint64_t b = __builtin_bit_cast(uint64_t, v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), 63);
  return __builtin_bit_cast(double, (uint64_t)lo | ((uint64_t)hi << 32));
}
// uint64 (mod 2^64) sum over the 64 lanes with DPP moves of both halves, same steps as
// wave_sum_u32_dpp; uniform result read from lane 63.
__device__ __forceinline__ uint64_t wave_sum_u64_dpp(uint64_t v) {
#define IVQM_DPPQ(ctrl, rmask, bc)                                                                      \
  {                                                                                                    \
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, ctrl, rmask, 0xf, bc); \
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), ctrl, rmask, 0xf, bc); \
    v += (uint64_t)lo | ((uint64_t)hi << 32);                                                          \
  }
  IVQM_DPPQ(0x111, 0xf, true)
  IVQM_DPPQ(0x112, 0xf, true)
  IVQM_DPPQ(0x114, 0xf, true)
  IVQM_DPPQ(0x118, 0xf, true)
  IVQM_DPPQ(0x142, 0xa, false)
  IVQM_DPPQ(0x143, 0xc, false)
#undef IVQM_DPPQ
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, 63);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), 63);
  return (uint64_t)lo | ((uint64_t)hi << 32);
}
// Index of the last table entry whose first-block index is <= blk.
template <typename E, typename F>
__device__ __forceinline__ int find_entry(const E* __restrict__ t, int n, int blk, F first) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (first(t[mid]) <= blk) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

}  // namespace
}  // namespace ivqm
