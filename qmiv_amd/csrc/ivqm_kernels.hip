// gfx950 (CDNA4) kernels of the QMIV metric hot path.
//
// Every kernel reads a table of frame pairs (or jobs) from device memory and
// flattens all of them into one grid, so one launch per phase serves a whole
// batch of pairs of mixed sizes and sample formats; blocks find their table
// entry by a binary search over the entries' first-block indices (uniform per
// block, so it runs on the scalar unit).
//
//   k_stats          per plane row: sum(P1 - P0) and sum((P0 - P1)^2), one wave per row
//                    (GCD input, PSNR / WS-PSNR row SSDs); zeroes the pair's
//                    search row accumulators as a side job.
//   k_gcd_finalize   per pair: frame totals + global colour difference (FP64
//                    mean, round half away, clip), kept in device memory.
//   k_search         corresponding-pixel-shift search, both directions, LDS
//                    tile with a +-R halo, int32-wrapping row sums, fused
//                    shift-compensated-picture epilogue -- or, for IV-SSIM in
//                    batches, the exact 4x4 sub-block moments of (test
//                    picture, SCP) in place of the SCP pictures.
//   k_ssim           block-averaged 8x8 / stride-4 SSIM: exact integer 4x4
//                    sub-block moments (summed from the pictures, or read
//                    from the search's fused moments), 2x2 combination, FP32
//                    ratios, FP64 per-window-row partial sums.
//   (k_ssim_gen*     every other SSIM configuration: ivqm_ssim_modes.hip)
//   k_ssim_finalize  FP64 frame totals per job and component (ERP-weighted
//                    for IV-SSIM).
//   k_downsample     2x2 rounded-mean pyramid level (MS-SSIM).
//   k_upsample420    4:2:0 / 4:2:2 -> 4:4:4 nearest replication (standalone).
//   k_colour         YCbCr <-> RGB fixed-point conversion to 4:4:4 (standalone).
//
// 4:2:0 and 4:2:2 input is never materialised at 4:4:4 by the metric kernels:
// a chroma sample at luma-grid (x, y) is read at (x >> 1, y >> 1) (4:2:0) or
// (x >> 1, y) (4:2:2), exactly the value 2x2 or horizontal replication
// (libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49, :107-120) stores there.
// The reference's Nearest margin (xMarginOps.h:40-73) is reproduced by
// clamping coordinates.
#include "ivqm_device.h"
#include "ivqm_device_util.h"

#include <limits.h>

#include <algorithm>

#include <type_traits>

namespace ivqm {

namespace {

__device__ __forceinline__ void wave_sum3_u32_dpp(uint32_t v[3]) {
#define IVQM_DPP3(ctrl, rmask, bc)                                                                     \
  {                                                                                                   \
    const uint32_t a0 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], ctrl, rmask, 0xf, bc);   \
    const uint32_t a1 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[1], ctrl, rmask, 0xf, bc);   \
    const uint32_t a2 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[2], ctrl, rmask, 0xf, bc);   \
    v[0] += a0;                                                                                       \
    v[1] += a1;                                                                                       \
    v[2] += a2;                                                                                       \
  }
  IVQM_DPP3(0x111, 0xf, true)
  IVQM_DPP3(0x112, 0xf, true)
  IVQM_DPP3(0x114, 0xf, true)
  IVQM_DPP3(0x118, 0xf, true)
  IVQM_DPP3(0x142, 0xa, false)
  IVQM_DPP3(0x143, 0xc, false)
#undef IVQM_DPP3
  for (int k = 0; k < 3; k++) v[k] = (uint32_t)__builtin_amdgcn_readlane((int)v[k], 63);
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// ---------------------------------------------------------------------------
// k_stats: one wave per plane row (grid-stride over all rows of all pairs),
// four 16-byte loads per picture in flight per lane; wave shuffles, no LDS.
// ---------------------------------------------------------------------------
// Adds the differences r - a of one 16-byte vector of samples (8 x 16-bit or
// 16 x 8-bit) to 32-bit partial sums; the caller widens them once per 4
// vectors (32 or 64 samples): up to 12-bit samples a square is < 2^24 and 64 of
// them fit uint32. 16-bit pairs use v_dot2 on the packed int16 differences
// (|d| < 2^15); bytes use v_dot4: sum (r - a)^2 = r.r + a.a - 2 r.a per word.
typedef short s16x2_t __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ void stats_vec32(const uint4& va, const uint4& vr, int32_t& s, uint32_t& q) {
  const uint32_t wa[4] = {va.x, va.y, va.z, va.w}, wr[4] = {vr.x, vr.y, vr.z, vr.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (sizeof(T) == 2) {
      const s16x2_t d = __builtin_bit_cast(s16x2_t, wr[k]) - __builtin_bit_cast(s16x2_t, wa[k]);
      s = __builtin_amdgcn_sdot2(d, s16x2_t{1, 1}, s, false);
      q = (uint32_t)__builtin_amdgcn_sdot2(d, d, (int)q, false);
    } else {
      const uint32_t rr = __builtin_amdgcn_udot4(wr[k], wr[k], 0u, false);
      const uint32_t aa = __builtin_amdgcn_udot4(wa[k], wa[k], 0u, false);
      const uint32_t ra = __builtin_amdgcn_udot4(wr[k], wa[k], 0u, false);
      q += rr + aa - 2u * ra;
      s += (int32_t)__builtin_amdgcn_udot4(wr[k], 0x01010101u, 0u, false) -
           (int32_t)__builtin_amdgcn_udot4(wa[k], 0x01010101u, 0u, false);
    }
  }
}
// Reference-width form (13- to 16-bit): squares need 64-bit sums.
__device__ __forceinline__ void stats_vec64(const uint4& va, const uint4& vr, int64_t& sd, uint64_t& ssd) {
  const uint32_t wa[4] = {va.x, va.y, va.z, va.w}, wr[4] = {vr.x, vr.y, vr.z, vr.w};
  int32_t s = 0;
  uint64_t q = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int d0 = (int)(wr[k] & 0xffffu) - (int)(wa[k] & 0xffffu);
    const int d1 = (int)(wr[k] >> 16) - (int)(wa[k] >> 16);
    s += d0 + d1;
    q += (uint64_t)((uint32_t)d0 * (uint32_t)d0) + (uint64_t)((uint32_t)d1 * (uint32_t)d1);
  }
  sd += s;
  ssd += q;
}

// Sums of one plane row (lb counts the rows of Y, then Cb, then Cr) by one wave. NARROW: the caller
// guarantees samples below 2^12 (the 32-bit partial form only).
template <typename T, bool NARROW = false>
__device__ __forceinline__ void stats_row(const PairDev& P, int lb, int lane, int64_t* __restrict__ row_sd,
                                          uint64_t* __restrict__ row_ssd) {
  const int H = P.h, W = P.w, Hc = H >> P.p0.csy, Wc = W >> P.p0.cs;
  int c, y, wp;
  if (lb < H) { c = 0; y = lb; wp = W; }
  else if (lb < H + Hc) { c = 1; y = lb - H; wp = Wc; }
  else { c = 2; y = lb - H - Hc; wp = Wc; }
  const T* a = (const T*)plane_ptr(P.p0, c) + (size_t)y * plane_stride(P.p0, c);
  const T* r = (const T*)plane_ptr(P.p1, c) + (size_t)y * plane_stride(P.p1, c);
  int64_t sd = 0;
  uint64_t ssd = 0;
  // 16-byte loads where both rows are aligned, scalar tail
  constexpr int V = 16 / sizeof(T);
  int x_done = 0;
  if (((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(r)) & 15) == 0) {
    const int nv = wp / V;
    // buffer resources over the row's whole vectors: loads past them return 0
    // (zero differences), so no per-load bounds branch or 64-bit address math
    const __amdgpu_buffer_rsrc_t ba =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(a), (short)0, nv * 16, 0x00020000);
    const __amdgpu_buffer_rsrc_t br =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(r), (short)0, nv * 16, 0x00020000);
    // the narrow form keeps a whole 4096-sample 16-bit row (8 vectors per picture and lane) in flight;
    // its 64 squares per lane and pass are < 2^30
    constexpr int NV = NARROW ? 8 : 4;
    for (int i0 = 0; i0 < nv; i0 += 64 * NV) {
      uint4 va[NV], vr[NV];
#pragma unroll
      for (int k = 0; k < NV; k++) {
        const uint32_t off = (uint32_t)(i0 + 64 * k + lane) * 16u;
        va[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(ba, off, 0, 0));
        vr[k] = __builtin_bit_cast(uint4, __builtin_amdgcn_raw_buffer_load_b128(br, off, 0, 0));
      }
      if (sizeof(T) == 1) {
        // bytes: one vector at a time (a scheduling barrier keeps the compiler
        // from hoisting all 80 dot products, which raised the kernel's registers)
        int32_t s = 0;
        uint32_t q = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          stats_vec32<T>(va[k], vr[k], s, q);
          __builtin_amdgcn_sched_barrier(0);
        }
        sd += s;
        ssd += q;
      } else if (NARROW || P.maxv < 4096) {
        int32_t s = 0;
        uint32_t q = 0;
#pragma unroll
        for (int k = 0; k < NV; k++) stats_vec32<T>(va[k], vr[k], s, q);
        sd += s;
        ssd += q;
      } else {
#pragma unroll
        for (int k = 0; k < NV; k++) stats_vec64(va[k], vr[k], sd, ssd);
      }
    }
    x_done = nv * V;
  }
  for (int x = x_done + lane; x < wp; x += 64) {
    const int d = (int)r[x] - (int)a[x];
    sd += d;
    ssd += (uint64_t)((uint32_t)d * (uint32_t)d);
  }
  sd = (int64_t)wave_sum_u64_dpp((uint64_t)sd);  // mod 2^64: two's complement sums
  ssd = wave_sum_u64_dpp(ssd);
  if (lane == 0) {
    row_sd[P.rowsd_off + (int64_t)c * H + y] = sd;
    row_ssd[P.rowsd_off + (int64_t)c * H + y] = ssd;
  }
}

// Mask mode: one 4:4:4 row of component c (subsampled chroma read at (x >> cs, y >> csy)), weighted by
// the mask at (x, y). xDistortionSTD::CalcWeightedSD / CalcWeightedSSD
// (libs/pmbb-core/src/xDistortionSTD.cpp:45-68): sum of d * M with the product in int32 (exact for
// masks up to 14 bits), and sum of (uint64)(int32 square of d) * M. aux gets the samples that
// differ (exact components) and, for luma, the non-zero mask samples (xPixelOps::CountNonZero).
template <typename T, typename TM>
__device__ void stats_row_masked(const PairDev& P, int c, int y, int lane, int64_t* __restrict__ row_sd,
                                 uint64_t* __restrict__ row_ssd, int64_t* __restrict__ aux) {
  const int W = P.w, H = P.h;
  const int cs = c ? P.p0.cs : 0, csy = c ? P.p0.csy : 0;
  const T* a = (const T*)plane_ptr(P.p0, c) + (size_t)(y >> csy) * plane_stride(P.p0, c);
  const T* r = (const T*)plane_ptr(P.p1, c) + (size_t)(y >> csy) * plane_stride(P.p1, c);
  const TM* m = (const TM*)P.msk + (size_t)y * P.msk_stride;
  int64_t sd = 0, nd = 0, nz = 0;
  uint64_t ssd = 0;
#pragma unroll 2
  for (int x = lane; x < W; x += 64) {
    const int d = (int)r[x >> cs] - (int)a[x >> cs];
    const int w = (int)m[x];
    sd += (int64_t)(d * w);
    ssd += (uint64_t)(int64_t)(int32_t)((uint32_t)d * (uint32_t)d) * (uint64_t)w;
    nd += d != 0;
    nz += w != 0;
  }
  sd = (int64_t)wave_sum_u64_dpp((uint64_t)sd);
  ssd = wave_sum_u64_dpp(ssd);
  nd = (int64_t)wave_sum_u64_dpp((uint64_t)nd);
  nz = (int64_t)wave_sum_u64_dpp((uint64_t)nz);
  if (lane == 0) {
    row_sd[P.rowsd_off + (int64_t)c * H + y] = sd;
    row_ssd[P.rowsd_off + (int64_t)c * H + y] = ssd;
    aux[P.aux_off + (int64_t)c * H + y] = nd;
    if (c == 0) aux[P.aux_off + 3 * (int64_t)H + y] = nz;
  }
}

// nrows = sum over pairs of H + 2 * Hc (masked pairs: 3 H); a pair's rows start at stat_blk0.
// Each luma row y also clears the six search row sums of y.
// FORM (stats_form): 0 the narrow 16-bit form alone, 64 VGPRs with a whole 4096-sample row of both
// pictures in flight per wave (8 waves per SIMD); 1 every other pair. Rows of pairs of the other form
// are skipped.
template <int FORM>
__global__ __launch_bounds__(256) void k_stats(const PairDev* __restrict__ pairs, int npairs, int nrows,
                                               int64_t* __restrict__ row_sd, uint64_t* __restrict__ row_ssd,
                                               uint32_t* __restrict__ rows, int64_t* __restrict__ aux,
                                               uint64_t* __restrict__ rows64) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * 4;
  // wave-uniform row index: the pair table entry and row pointers stay scalar
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  for (int rw = blockIdx.x * 4 + wv; rw < nrows; rw += nw) {
    const int pi = find_entry(pairs, npairs, rw, [](const PairDev& p) { return p.stat_blk0; });
    const PairDev& P = pairs[pi];
    const int lb = rw - P.stat_blk0;
    if (stats_form(P) != FORM) continue;  // wave-uniform
    if constexpr (FORM == 0) {
      if (lb < P.h && lane < 6) rows[P.rows_off + (int64_t)lane * P.h + lb] = 0u;
      stats_row<uint16_t, true>(P, lb, lane, row_sd, row_ssd);
      continue;
    }
    if (P.msk) {
      const int c = lb / P.h, y = lb - c * P.h;
      if (c == 0 && lane < 6) rows64[P.rows64_off + (int64_t)lane * P.h + y] = 0u;
      if (P.p0.bps == 2) {
        if (P.msk_bps == 2) stats_row_masked<uint16_t, uint16_t>(P, c, y, lane, row_sd, row_ssd, aux);
        else stats_row_masked<uint16_t, uint8_t>(P, c, y, lane, row_sd, row_ssd, aux);
      } else {
        if (P.msk_bps == 2) stats_row_masked<uint8_t, uint16_t>(P, c, y, lane, row_sd, row_ssd, aux);
        else stats_row_masked<uint8_t, uint8_t>(P, c, y, lane, row_sd, row_ssd, aux);
      }
      continue;
    }
    if (lb < P.h && lane < 6) rows[P.rows_off + (int64_t)lane * P.h + lb] = 0u;
    if (P.p0.bps == 2) stats_row<uint16_t>(P, lb, lane, row_sd, row_ssd);
    else stats_row<uint8_t>(P, lb, lane, row_sd, row_ssd);
  }
}

// ---------------------------------------------------------------------------
// k_gcd_finalize: one block per pair. libs/pmbb-ivqm/src/xGlobClrDiff.cpp:23-56
// with the app's argument order (sum of P1 - P0); chroma sums of 4:2:0 x4, of 4:2:2 x2.
// ---------------------------------------------------------------------------
// Mask mode: xGlobClrDiff::CalcGlobalColorDiffM (xGlobClrDiff.cpp:60-91): mean = SD / (NNM * MaxM);
// the row sums are already 4:4:4 totals; mcount receives NNM and the differing-sample counts.
__global__ __launch_bounds__(256) void k_gcd_finalize(const PairDev* __restrict__ pairs,
                                                      const int64_t* __restrict__ row_sd,
                                                      const uint64_t* __restrict__ row_ssd,
                                                      const int64_t* __restrict__ aux, int32_t* __restrict__ gcd,
                                                      int64_t* __restrict__ sd_out, uint64_t* __restrict__ ssd_out,
                                                      int64_t* __restrict__ mcount) {
  const PairDev P = pairs[blockIdx.x];
  __shared__ int64_t s_sd[4], s_nd[4], s_nz[4];
  __shared__ uint64_t s_ssd[4];
  const int tid = threadIdx.x;
  const int W = P.w, H = P.h, cs = P.p0.cs, csy = P.p0.csy;
  const bool masked = P.msk != nullptr;
  int64_t nnm = 0;
  if (masked) {  // non-zero mask samples
    int64_t nz = 0;
    for (int y = tid; y < H; y += 256) nz += aux[P.aux_off + 3 * (int64_t)H + y];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) nz += __shfl_xor(nz, m, 64);
    if ((tid & 63) == 0) s_nz[tid >> 6] = nz;
    __syncthreads();
    nnm = s_nz[0] + s_nz[1] + s_nz[2] + s_nz[3];
    if (tid == 0) mcount[4 * P.res_idx] = nnm;
  }
  for (int c = 0; c < 3; c++) {
    const int rows = (c && !masked) ? (H >> csy) : H;
    int64_t sd = 0, nd = 0;
    uint64_t ssd = 0;
    for (int y = tid; y < rows; y += 256) {
      sd += row_sd[P.rowsd_off + (int64_t)c * H + y];
      ssd += row_ssd[P.rowsd_off + (int64_t)c * H + y];
      if (masked) nd += aux[P.aux_off + (int64_t)c * H + y];
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      sd += __shfl_xor(sd, m, 64);
      ssd += __shfl_xor(ssd, m, 64);
      nd += __shfl_xor(nd, m, 64);
    }
    if ((tid & 63) == 0) { s_sd[tid >> 6] = sd; s_ssd[tid >> 6] = ssd; s_nd[tid >> 6] = nd; }
    __syncthreads();
    if (tid == 0 && masked) {
      const int64_t t = s_sd[0] + s_sd[1] + s_sd[2] + s_sd[3];
      sd_out[3 * P.res_idx + c] = t;
      ssd_out[3 * P.res_idx + c] = s_ssd[0] + s_ssd[1] + s_ssd[2] + s_ssd[3];
      mcount[4 * P.res_idx + 1 + c] = s_nd[0] + s_nd[1] + s_nd[2] + s_nd[3];
      if (P.gcd_given) {
        gcd[4 * P.res_idx + c] = P.gcd_in[c];
      } else {
        const double avg = (double)t / (double)(nnm * (int64_t)P.msk_max);
        const double rnd = avg + copysign(0.5, avg);
        int g = nnm > 0 ? (int)rnd : 0;  // no weighted sample: reported by the host
        const int md = P.md[c];
        g = g < -md ? -md : (g > md ? md : g);
        gcd[4 * P.res_idx + c] = g;
      }
      if (c == 2) gcd[4 * P.res_idx + 3] = 0;
    }
    if (tid == 0 && !masked) {
      int64_t t = s_sd[0] + s_sd[1] + s_sd[2] + s_sd[3];
      uint64_t q = s_ssd[0] + s_ssd[1] + s_ssd[2] + s_ssd[3];
      if (c) {  // a subsampled chroma sample stands for 2 (4:2:2) or 4 (4:2:0) samples at 4:4:4
        t *= (int64_t)1 << (cs + csy);
        q *= (uint64_t)1 << (cs + csy);
      }
      sd_out[3 * P.res_idx + c] = t;
      ssd_out[3 * P.res_idx + c] = q;
      if (P.gcd_given) {
        gcd[4 * P.res_idx + c] = P.gcd_in[c];
        if (c == 2) gcd[4 * P.res_idx + 3] = 0;
      } else {
        if (P.wrap32) t = (int64_t)(int32_t)(uint32_t)(uint64_t)t;
        const double avg = (double)t / (double)((int64_t)W * (int64_t)H);
        const double rnd = avg + copysign(0.5, avg);
        int g = (int)rnd;
        const int md = P.md[c];
        g = g < -md ? -md : (g > md ? md : g);
        gcd[4 * P.res_idx + c] = g;
        if (c == 2) gcd[4 * P.res_idx + 3] = 0;
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// k_search: corresponding-pixel-shift search.
//
// Reference semantics (x86 default build): libs/pmbb-ivqm/src/
// xCorrespPixelShiftSSE.cpp:18-70 and :75-127 -- for every test pixel
// t = Tst + gcs, the first strict minimum in dy-outer raster order of
// E = sum_c w_c (t_c - r_c)^2 over the (2R+1)^2 reference candidates, int32
// arithmetic that wraps; per-row int32-wrapping sums of the winner's
// per-component squared differences; SCP = clip(r_best - gcs, 0, Max).
//
// Block = 8 waves on one 64 x 16 pixel tile: waves 0-3 run direction 0,
// waves 4-7 direction 1. Each direction's (64 + 2R) x (16 + 2R) reference
// region is staged once in LDS as 12-byte entries, and each direction reads
// its test samples from the other direction's region. A wave covers 4 rows x
// 64 columns; lane l owns the 4 adjacent pixels 4 (l % 16) .. + 3 of row
// l / 16. For each candidate row a lane reads its 4 + 2R entries once and
// scores its 4 pixels' 2R + 1 candidates from them (lanes 48 bytes apart: no
// LDS bank conflicts). A pixel row's squared-difference sums are an in-lane
// add plus a 16-lane DPP scan; SCP rows are written with 8-byte stores.
//
// NARROW variant (host checks the ranges): E = T + E', T = sum w t^2 constant
// per pixel and E' = Q(r) - 2 sum_c w_c t_c r_c with Q(r) = sum_c w_c r_c^2
// precomputed per reference entry. Nothing wraps in this range, so
// argmin E == argmin E'. The key S * E' + idx (S = 2^KB > number of
// candidates, idx the candidate's raster index) makes a plain integer min
// equal to "first strict minimum in raster order". An entry at tile position
// (ly, lx) stores S * Q + ly * D + lx (D = 2R + 1): the candidate (dy, dx) of
// the pixel at tile position (py, px) then scores its key plus the constant
// py * D + px, which leaves the order unchanged and is subtracted from the
// pixel's minimum. Entry {S*Q + off, Y | Cb << 16, Cr}; per candidate three
// 24-bit multiply-adds and a min.
//
// DOT variant (host-checked: S * Max and 2 * w * (Max + |gcd|) fit int16):
// the same key, with {Y*S, Cb*S} stored as an int16 pair: entry
// {S*Q + off, Y*S | Cb*S << 16, Cr}; per candidate one v_dot2 (Y, Cb), one
// 24-bit multiply-add (Cr; its multiplier carries S) and a min.
//
// WIDE variant: the reference-form int32 arithmetic with an explicit
// first-strict-minimum compare in raster order; entry {Y, Cb, Cr}.
// WIDE24 variant (host-checked (Max + |gcd|) * |w| < 2^23): as WIDE, with each
// weighted square (d*d)*w mod 2^32 computed as (d*w)*d by two 24-bit
// multiplies -- the same value mod 2^32.
// ---------------------------------------------------------------------------
typedef short short2_t __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2_t __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
// kSearchWide24D: kSearchWide24 with the default component weights {4, 1, 1}: the LDS holds 2Y, so the
// luma term 4 dY^2 is one 24-bit multiply (2dY)^2 -- the same value mod 2^32 as the reference's int32
// 4 * dY^2 -- and a candidate costs three multiply-adds instead of six multiplies.
// kSearchQuad (host-checked 2 |w| (Max + |gcd|) < 2^23): the reference's wrapping score in expanded form.
// sum_c w_c (t_c - r_c)^2 = T + Q(r) - 2 sum_c w_c t_c r_c holds over the integers, so it holds mod 2^32:
// with Q(r) = sum_c w_c r_c^2 mod 2^32 stored per entry, the multipliers m_c = -2 w_c t_c (24-bit) and
// T = sum_c w_c t_c^2 mod 2^32 per pixel, three 24-bit multiply-adds and an add give exactly the int32
// value the reference compares, wrapped or not. Candidates are reduced per candidate row (min3 folds, no
// index), a strictly smaller row minimum moves the pixel's best row, and after the search the winning
// row is scored once more to find its first column holding the minimum: the first strict minimum in
// raster order, at about half the instructions of the explicit compare-and-select per candidate.
enum SearchMode {
  kSearchWide = 0, kSearchNarrow = 1, kSearchDot = 2, kSearchWide24 = 3, kSearchWide24D = 5, kSearchQuad = 6
};
static_assert(kSearchQuad == kSearchQuadMode, "mode number shared with the host");
template <int R> struct KeyBits {
  static constexpr int n = (2 * R + 1) * (2 * R + 1);
  static constexpr int v = n <= 16 ? 4 : n <= 32 ? 5 : n <= 64 ? 6 : 7;
};

// LW is the region width in entries, LP the LDS row pitch: a multiple of 4
// entries (48 bytes), so a lane's run of entries starting at a multiple of 4
// is 16-byte aligned.
template <int R> struct SearchTile {
  static constexpr int TW = kSearchTileW, TH = 16, LW = TW + 2 * R, LH = TH + 2 * R, LP = (LW + 3) & ~3;
};
static_assert(SearchTile<1>::TH == kSearchTileH, "tile height");

struct SEnt {
  int x, y, z;
};
typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x2_t __attribute__((ext_vector_type(2)));

// NE consecutive entries from a 16-byte aligned LDS address, read as 16-byte
// vectors (lanes 48 bytes apart hit distinct banks) plus an 8-byte tail.
template <int NE>
__device__ __forceinline__ void lds_run(const SEnt* p, SEnt (&e)[NE]) {
  constexpr int NW = 3 * NE;
  static_assert(NW % 4 == 0 || NW % 4 == 2, "run length");
  int w[NW];
  const i32x4_t* q = reinterpret_cast<const i32x4_t*>(p);
#pragma unroll
  for (int i = 0; i < NW / 4; i++) {
    const i32x4_t v = q[i];
    w[4 * i] = v.x;
    w[4 * i + 1] = v.y;
    w[4 * i + 2] = v.z;
    w[4 * i + 3] = v.w;
  }
  if (NW % 4) {
    const i32x2_t v = *reinterpret_cast<const i32x2_t*>(q + NW / 4);
    w[NW - 2] = v.x;
    w[NW - 1] = v.y;
  }
#pragma unroll
  for (int j = 0; j < NE; j++) e[j] = SEnt{w[3 * j], w[3 * j + 1], w[3 * j + 2]};
}

// The raw samples {Y, Cb, Cr} of an LDS entry.
template <int MODE, int KB>
__device__ __forceinline__ void sent_raw(const SEnt& e, int& Y, int& Cb, int& Cr) {
  if (MODE == kSearchDot) {
    Y = (int)(((uint32_t)e.y & 0xffffu) >> KB);
    Cb = (int)((uint32_t)e.y >> (16 + KB));
  } else if (MODE == kSearchNarrow || MODE == kSearchQuad) {
    Y = (int)((uint32_t)e.y & 0xffffu);
    Cb = (int)((uint32_t)e.y >> 16);
  } else {
    Y = MODE == kSearchWide24D ? (e.x >> 1) : e.x;
    Cb = e.y;
  }
  Cr = e.z;
}

// Entries A .. A + N - 1 after a 16-byte aligned LDS address, read as the
// 16-byte chunks that cover them (A * 12 need not be aligned).
template <int A, int N>
__device__ __forceinline__ void lds_entries(const SEnt* p, SEnt (&e)[N]) {
  constexpr int C0 = 12 * A / 16, C1 = (12 * (A + N) + 15) / 16;
  int w[4 * (C1 - C0)];
  const i32x4_t* q = reinterpret_cast<const i32x4_t*>(p);
#pragma unroll
  for (int i = C0; i < C1; i++) {
    const i32x4_t v = q[i];
    w[4 * (i - C0)] = v.x;
    w[4 * (i - C0) + 1] = v.y;
    w[4 * (i - C0) + 2] = v.z;
    w[4 * (i - C0) + 3] = v.w;
  }
#pragma unroll
  for (int j = 0; j < N; j++) {
    const int o = 3 * (A + j) - 4 * C0;
    e[j] = SEnt{w[o], w[o + 1], w[o + 2]};
  }
}

// kSearchQuad entries {Q(r), Y, Cb, Cr}: 16 bytes, every operand a whole register (no unpacking in
// the candidate loop).
struct SEnt4 {
  int x, y, z, w;
};
// Entries A .. A + N - 1 of a run that starts at a group boundary of the padded layout (see lds_col).
template <int A, int N>
__device__ __forceinline__ void lds_entries(const SEnt4* p, SEnt4 (&e)[N]) {
  const i32x4_t* q = reinterpret_cast<const i32x4_t*>(p);
#pragma unroll
  for (int j = 0; j < N; j++) {
    const i32x4_t v = q[(A + j) + ((A + j) >> 2)];
    e[j] = SEnt4{v.x, v.y, v.z, v.w};
  }
}
template <int NE>
__device__ __forceinline__ void lds_run(const SEnt4* p, SEnt4 (&e)[NE]) {
  lds_entries<0, NE>(p, e);
}
template <int MODE, int KB>
__device__ __forceinline__ void sent_raw(const SEnt4& e, int& Y, int& Cb, int& Cr) {
  Y = e.y;
  Cb = e.z;
  Cr = e.w;
}
template <int MODE> struct EntryOf {
  using type = SEnt;
};
template <> struct EntryOf<kSearchQuad> {
  using type = SEnt4;
};
// LDS column of region column x. 16-byte entries get one pad entry after every 4, so the lanes of a
// wave, whose runs start 4 entries apart, read 80 bytes apart: 8 consecutive lanes' 16-byte reads
// cover all 32 banks once (64 bytes apart they would meet in pairs of banks). 12-byte entries are
// 48 bytes apart per lane already.
template <typename E> __device__ __forceinline__ constexpr int lds_col(int x) {
  return std::is_same<E, SEnt4>::value ? x + (x >> 2) : x;
}
template <typename E, int LP> struct LdsPitch {
  static constexpr int v = std::is_same<E, SEnt4>::value ? LP + LP / 4 : LP;
};

// One step of an inclusive scan of three values within each 16-lane row.
template <int CTRL>
__device__ __forceinline__ void dpp3_add(uint32_t v[3]) {
  const uint32_t a0 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], CTRL, 0xf, 0xf, true);
  const uint32_t a1 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[1], CTRL, 0xf, 0xf, true);
  const uint32_t a2 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[2], CTRL, 0xf, 0xf, true);
  v[0] += a0;
  v[1] += a1;
  v[2] += a2;
}
// row_shr 1, 2, 4, 8: lane 15 of every 16-lane row ends with the row's totals.
__device__ __forceinline__ void row16_sum3_u32_dpp(uint32_t v[3]) {
  dpp3_add<0x111>(v);
  dpp3_add<0x112>(v);
  dpp3_add<0x114>(v);
  dpp3_add<0x118>(v);
}

// Cross-row sums with the gfx950 lane-swap instructions (no LDS): v_permlane32_swap
// exchanges lanes 32-63 of its first operand with lanes 0-31 of its second, so
// x' + y' is x's 2-lane sum (lanes l, l ^ 32) in lanes 0-31 and y's in lanes
// 32-63; v_permlane16_swap exchanges the odd 16-lane rows of the first operand
// with the even rows of the second: x' + y' is x's sum over rows (0, 1) in row 0
// and over rows (2, 3) in row 2, y's in rows 1 and 3.
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ uint32_t swap_sum32(uint32_t x, uint32_t y) {
  const auto r = __builtin_amdgcn_permlane32_swap(x, y, false, false);
  return r[0] + r[1];
}
__device__ __forceinline__ uint32_t swap_sum16(uint32_t x, uint32_t y) {
  const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
  return r[0] + r[1];
}

template <typename T, int R, int MODE>
__device__ __forceinline__ void search_tile(const PairDev& Pr, int x0, int y0,
                                            typename EntryOf<MODE>::type (
                                                *lds2)[SearchTile<R>::LH]
                                                      [LdsPitch<typename EntryOf<MODE>::type, SearchTile<R>::LP>::v],
                                            int wY, int wCb,
                                            int wCr, const int32_t* __restrict__ gcd, uint32_t* __restrict__ rows,
                                            uint16_t* __restrict__ scp, uint8_t* __restrict__ best) {
  constexpr int D = 2 * R + 1, NE = 4 + 2 * R;
  constexpr int LH = SearchTile<R>::LH, LP = SearchTile<R>::LP;
  constexpr int KB = KeyBits<R>::v, S = 1 << KB;
  constexpr bool KEYED = MODE == kSearchDot || MODE == kSearchNarrow;
  const int W = Pr.w, H = Pr.h;
  // Waves 0-3 run direction 0 (R2T: Tst = P0 + gcd, Ref = P1), waves 4-7
  // direction 1 (T2R: Tst = P1 - gcd, Ref = P0). Region lds2[d] holds
  // direction d's reference: P1 for d = 0, P0 for d = 1.
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int dir = wv >> 2, wd = wv & 3;  // direction, wave within the direction
  const int tl = (int)threadIdx.x & 255;  // thread within the direction
  const PicDev& Ref = dir ? Pr.p0 : Pr.p1;
  using E = typename EntryOf<MODE>::type;
  constexpr int LPE = LdsPitch<E, LP>::v;
  E (*lds)[LPE] = lds2[dir];
  E (*tlds)[LPE] = lds2[1 - dir];
  const int32_t* g4 = gcd + 4 * Pr.res_idx;
  const int g0 = dir ? -g4[0] : g4[0];
  const int g1 = dir ? -g4[1] : g4[1];
  const int g2 = dir ? -g4[2] : g4[2];

  // Every global load of the region (rows, then halo columns) is issued
  // before any is used: one memory round trip per block. Wave wd stages row
  // pairs wd, wd + 4, ...: the row (and its plane offsets) is wave-uniform and
  // each lane keeps one clamped column. Rows are staged in pairs of
  // consecutive picture rows whose first row is even, (2m, 2m + 1): with 4:2:0
  // chroma both read the same chroma row, loaded once (clamped rows keep that
  // property for even heights). The 2R halo columns are staged by threads < 2R * LH.
  constexpr int PAR = R & 1;               // region row 0 is picture row y0 - R, odd for odd R
  constexpr int NPT = (LH + PAR + 1) / 2;  // row pairs of the region
  constexpr int NPW = (NPT + 3) / 4;       // row pairs per wave
  const uint32_t cx = (uint32_t)clampi(x0 - R + lane, 0, W - 1);
  // byte offsets of the lane's column in luma / chroma rows (4:2:0 and 4:2:2
  // chroma at cx >> 1), and the planes' buffer resources; row byte offsets are scalar
  const uint32_t ol = cx * (uint32_t)sizeof(T), oc = (cx >> Ref.cs) * (uint32_t)sizeof(T);
  const __amdgpu_buffer_rsrc_t rs0 = plane_rsrc<T>(Ref, 0, H), rs1 = plane_rsrc<T>(Ref, 1, H),
                               rs2 = plane_rsrc<T>(Ref, 2, H);
  const uint32_t st0 = (uint32_t)plane_stride(Ref, 0) * (uint32_t)sizeof(T);
  const uint32_t st1 = (uint32_t)plane_stride(Ref, 1) * (uint32_t)sizeof(T);
  const uint32_t st2 = (uint32_t)plane_stride(Ref, 2) * (uint32_t)sizeof(T);
  int ra[NPW][2][3];
#pragma unroll
  for (int k = 0; k < NPW; k++) {
    const int j = wd + 4 * k;
    const int ly0 = 2 * j - PAR, ly1 = ly0 + 1;
    // ra[k][h] is read only under the conditions that load it
    if (j < NPT) {
      const int ya = clampi(y0 - R + ly0, 0, H - 1), yb2 = clampi(y0 - R + ly1, 0, H - 1);
      const uint32_t ua = (uint32_t)ya, ub = (uint32_t)yb2;
      if (ly0 >= 0) ra[k][0][0] = ld_buf<T>(rs0, ol, ua * st0);
      if (ly1 < LH) ra[k][1][0] = ld_buf<T>(rs0, ol, ub * st0);
      if (Ref.csy) {  // 4:2:0: one chroma row serves the row pair
        const uint32_t yc = (ly0 >= 0 ? ua : ub) >> 1;
        const int cb = ld_buf<T>(rs1, oc, yc * st1), cr = ld_buf<T>(rs2, oc, yc * st2);
        ra[k][0][1] = ra[k][1][1] = cb;
        ra[k][0][2] = ra[k][1][2] = cr;
      } else {  // 4:4:4 and 4:2:2: a chroma row per row (column oc)
        if (ly0 >= 0) {
          ra[k][0][1] = ld_buf<T>(rs1, oc, ua * st1);
          ra[k][0][2] = ld_buf<T>(rs2, oc, ua * st2);
        }
        if (ly1 < LH) {
          ra[k][1][1] = ld_buf<T>(rs1, oc, ub * st1);
          ra[k][1][2] = ld_buf<T>(rs2, oc, ub * st2);
        }
      }
    }
  }
  constexpr int NH = 2 * R * LH;         // halo entries (columns 64 .. 64 + 2R - 1)
  constexpr int NHK = (NH + 255) / 256;  // halo entries per thread
  int rh[NHK][3];
#pragma unroll
  for (int k = 0; k < NHK; k++) {
    const int e = tl + 256 * k;
    rh[k][0] = rh[k][1] = rh[k][2] = 0;
    if (e < NH) {
      const int hy = e / (2 * R), hx = e - hy * (2 * R);
      const int x = clampi(x0 - R + 64 + hx, 0, W - 1), y = clampi(y0 - R + hy, 0, H - 1);
      rh[k][0] = ldpix<T>(Ref, 0, x, y);
      rh[k][1] = ldpix<T>(Ref, 1, x, y);
      rh[k][2] = ldpix<T>(Ref, 2, x, y);
    }
  }

  // Keyed entries. Host-checked ranges: w * Max fits 16 bits (DOT), S * Q(r)
  // fits int32 and w * r < 2^24, so the 16- and 24-bit multiplies are exact.
  auto entry = [&](int rY, int rCb, int rCr, int off) -> E {
    if constexpr (MODE == kSearchDot) {
      const u16x2_t pr = {(unsigned short)rY, (unsigned short)rCb};
      const u16x2_t wp = {(unsigned short)wY, (unsigned short)wCb};
      const uint32_t qc = __umul24(__umul24((uint32_t)wCr & 0xffffu, (uint32_t)rCr), (uint32_t)rCr);
      const uint32_t q = __builtin_amdgcn_udot2(pr * wp, pr, qc, false);
      const u16x2_t ps = pr << (unsigned short)KB;
      return E{(int)((q << KB) + (uint32_t)off), __builtin_bit_cast(int, ps), rCr};
    }
    if (MODE == kSearchNarrow) {
      const uint32_t uY = (uint32_t)wY & 0xffffu, uCb = (uint32_t)wCb & 0xffffu, uCr = (uint32_t)wCr & 0xffffu;
      const uint32_t q = __umul24(__umul24(uY, (uint32_t)rY), (uint32_t)rY) +
                         __umul24(__umul24(uCb, (uint32_t)rCb), (uint32_t)rCb) +
                         __umul24(__umul24(uCr, (uint32_t)rCr), (uint32_t)rCr);
      return E{(int)((q << KB) + (uint32_t)off), (int)((uint32_t)rY | ((uint32_t)rCb << 16)), rCr};
    }
    if constexpr (MODE == kSearchQuad) {
      // |w r| < 2^23 (host-checked): 24-bit operands, products mod 2^32
      const uint32_t q = (uint32_t)__mul24(__mul24(wY, rY), rY) + (uint32_t)__mul24(__mul24(wCb, rCb), rCb) +
                         (uint32_t)__mul24(__mul24(wCr, rCr), rCr);
      return E{(int)q, rY, rCb, rCr};
    } else {
      return E{MODE == kSearchWide24D ? 2 * rY : rY, rCb, rCr};
    }
  };
#pragma unroll
  for (int k = 0; k < NPW; k++) {
    const int j = wd + 4 * k;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const int ly = 2 * j - PAR + h;
      if (j < NPT && ly >= 0 && ly < LH) lds[ly][lds_col<E>(lane)] = entry(ra[k][h][0], ra[k][h][1], ra[k][h][2], ly * D + lane);
    }
  }
#pragma unroll
  for (int k = 0; k < NHK; k++) {
    const int e = tl + 256 * k;
    if (e < NH) {
      const int hy = e / (2 * R), hx = e - hy * (2 * R);
      lds[hy][lds_col<E>(64 + hx)] = entry(rh[k][0], rh[k][1], rh[k][2], hy * D + 64 + hx);
    }
  }
  __syncthreads();

  // This lane's pixels: tile row iy, tile columns ix .. ix + 3.
  const int iy = wd * 4 + (lane >> 4), ix = 4 * (lane & 15);
  const int y = y0 + iy, x = x0 + ix;
  // Test samples: the other direction's region at (iy + R, ix + k + R) holds
  // the test picture at the clamped position -- the sample a direct load reads.
  // WIDE: t = s + g per component and pixel. Keyed: the test values of
  // pixels (0, 1) and (2, 3) as int16 pairs per component, tp[c][h], and the
  // multipliers -2 S w t (NARROW, DOT Cr: exact 24-bit multiply-adds, |value|
  // < 2^23 host-checked) and {-2 wY t0, -2 wCb t1} (DOT: one packed 16-bit
  // multiply-add, exact mod 2^16 and the value fits int16, host-checked).
  int t0[4], t1[4], t2[4];
  int t0d[4];  // kSearchWide24D: 2 t0, against the doubled luma entries
  uint32_t tp[3][2];
  int m0[4], m1[4], m2[4];
  E te[4];
  lds_entries<R, 4>(&tlds[iy + R][lds_col<E>(ix)], te);
  if (KEYED) {
    const int KY = -2 * S * wY, KCb = -2 * S * wCb, KCr = -2 * S * wCr;
    uint32_t yc[4];  // raw {Y, Cb} per pixel
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int cr = te[k].z;
      m2[k] = __mul24(cr, KCr) + KCr * g2;
      if (MODE == kSearchDot) {
        const u16x2_t raw = __builtin_bit_cast(u16x2_t, te[k].y) >> (unsigned short)KB;
        const u16x2_t kp = {(unsigned short)(-2 * wY), (unsigned short)(-2 * wCb)};
        const u16x2_t kg = {(unsigned short)(-2 * wY * g0), (unsigned short)(-2 * wCb * g1)};
        m0[k] = __builtin_bit_cast(int, raw * kp + kg);
        m1[k] = 0;
        yc[k] = __builtin_bit_cast(uint32_t, raw);
      } else {
        yc[k] = (uint32_t)te[k].y;
        m0[k] = __mul24((int)(yc[k] & 0xffffu), KY) + KY * g0;
        m1[k] = __mul24((int)(yc[k] >> 16), KCb) + KCb * g1;
      }
    }
    const u16x2_t gy = (unsigned short)g0, gb = (unsigned short)g1, gr = (unsigned short)g2;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const u16x2_t sy = {(unsigned short)yc[2 * h], (unsigned short)yc[2 * h + 1]};
      const u16x2_t sb = {(unsigned short)(yc[2 * h] >> 16), (unsigned short)(yc[2 * h + 1] >> 16)};
      const u16x2_t sr = {(unsigned short)te[2 * h].z, (unsigned short)te[2 * h + 1].z};
      tp[0][h] = __builtin_bit_cast(uint32_t, sy + gy);
      tp[1][h] = __builtin_bit_cast(uint32_t, sb + gb);
      tp[2][h] = __builtin_bit_cast(uint32_t, sr + gr);
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      int s0, s1, s2;
      sent_raw<MODE, KB>(te[k], s0, s1, s2);
      t0[k] = s0 + g0;
      t1[k] = s1 + g1;
      t2[k] = s2 + g2;
      t0d[k] = 2 * t0[k];
    }
  }
  // kSearchQuad: m_c = -2 w_c t_c and T = sum_c w_c t_c^2 mod 2^32 (|w t| < 2^22, host-checked)
  int q0[4], q1[4], q2[4];
  uint32_t qT[4];
  if (MODE == kSearchQuad) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      q0[k] = -2 * wY * t0[k];
      q1[k] = -2 * wCb * t1[k];
      q2[k] = -2 * wCr * t2[k];
      qT[k] = (uint32_t)__mul24(__mul24(wY, t0[k]), t0[k]) + (uint32_t)__mul24(__mul24(wCb, t1[k]), t1[k]) +
              (uint32_t)__mul24(__mul24(wCr, t2[k]), t2[k]);
    }
  }
  // the score of candidate entry c for pixel k (kSearchQuad)
  auto quad_score = [&](int k, const E& c) -> int {
    int cr = 0;
    if constexpr (MODE == kSearchQuad) cr = c.w;
    const int acc = mad_i24(q2[k], cr, mad_i24(q1[k], c.z, mad_i24(q0[k], c.y, c.x)));
    return (int)((uint32_t)acc + qT[k]);
  };

  int bk[4], bi[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    bk[k] = INT_MAX;
    bi[k] = KEYED || MODE == kSearchQuad ? 0 : -1;  // kSearchQuad: the best candidate row; -1: none taken
  }
  if (MODE == kSearchQuad) {
    // one candidate row per iteration (a rolled loop bounds the live registers)
#pragma unroll 1
    for (int dy = 0; dy < D; dy++) {
      E e[NE];
      lds_run<NE>(&lds[iy + dy][lds_col<E>(ix)], e);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        int rm = quad_score(k, e[k]);
#pragma unroll
        for (int dx = 1; dx < D; dx++) rm = min(rm, quad_score(k, e[k + dx]));
        if (rm < bk[k]) {
          bk[k] = rm;
          bi[k] = dy;
        }
      }
    }
  }
  // keyed forms: the candidate rows unrolled; reference forms: a rolled loop (bounds the live registers)
#pragma unroll
  for (int dy = 0; dy < (KEYED ? D : 0); dy++) {
    E e[NE];
    lds_run<NE>(&lds[iy + dy][lds_col<E>(ix)], e);
    if (MODE == kSearchDot) {
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int dx = 0; dx < D; dx++) {
          const E& c = e[k + dx];
          // multiply-add into a fresh register first: v_dot2c accumulates in place
          int acc = __mul24(m2[k], c.z) + c.x;
          acc = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, m0[k]), __builtin_bit_cast(short2_t, c.y), acc,
                                       false);
          bk[k] = min(bk[k], acc);
        }
    } else if (MODE == kSearchNarrow) {
      int eY[NE], eCb[NE];
#pragma unroll
      for (int j = 0; j < NE; j++) {
        eY[j] = (int)((uint32_t)e[j].y & 0xffffu);
        eCb[j] = (int)((uint32_t)e[j].y >> 16);
      }
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int dx = 0; dx < D; dx++) {
          int acc = __mul24(m0[k], eY[k + dx]) + e[k + dx].x;
          acc = __mul24(m1[k], eCb[k + dx]) + acc;
          acc = __mul24(m2[k], e[k + dx].z) + acc;
          bk[k] = min(bk[k], acc);
        }
    }
  }
  if (!KEYED && MODE != kSearchQuad) {
#pragma unroll 1
    for (int dy = 0; dy < D; dy++) {
      E e[NE];
      lds_run<NE>(&lds[iy + dy][lds_col<E>(ix)], e);
#pragma unroll
      for (int k = 0; k < 4; k++)
#pragma unroll
        for (int dx = 0; dx < D; dx++) {
          const E& c = e[k + dx];
          int v;
          if (MODE == kSearchWide24D) {
            // |2 dY| < 2^23: exact 24-bit operands, products mod 2^32 as the reference's int32
            const int d0 = t0d[k] - c.x, d1 = t1[k] - c.y, d2 = t2[k] - c.z;
            uint32_t acc = (uint32_t)__mul24(d0, d0);
            acc += (uint32_t)__mul24(d1, d1);
            acc += (uint32_t)__mul24(d2, d2);
            v = (int)acc;
          } else if (MODE == kSearchWide24) {
            const int d0 = t0[k] - c.x, d1 = t1[k] - c.y, d2 = t2[k] - c.z;
            uint32_t acc = (uint32_t)__mul24(__mul24(d0, wY), d0);
            acc += (uint32_t)__mul24(__mul24(d1, wCb), d1);
            acc += (uint32_t)__mul24(__mul24(d2, wCr), d2);
            v = (int)acc;
          } else {
            const uint32_t d0 = (uint32_t)__mul24(t0[k] - c.x, t0[k] - c.x);
            const uint32_t d1 = (uint32_t)__mul24(t1[k] - c.y, t1[k] - c.y);
            const uint32_t d2 = (uint32_t)__mul24(t2[k] - c.z, t2[k] - c.z);
            v = (int)(d0 * (uint32_t)wY + d1 * (uint32_t)wCb + d2 * (uint32_t)wCr);
          }
          if (v < bk[k]) {
            bk[k] = v;
            bi[k] = dy * D + dx;
          }
        }
    }
  }
  if constexpr (MODE == kSearchQuad) {
    // the test samples again (not kept in registers through the search)
    __builtin_amdgcn_sched_barrier(0);
    asm volatile("" ::: "memory");
    E tq[4];
    lds_entries<R, 4>(&tlds[iy + R][lds_col<E>(ix)], tq);
#pragma unroll
    for (int k = 0; k < 4; k++) {
      t0[k] = tq[k].y + g0;
      t1[k] = tq[k].z + g1;
      t2[k] = tq[k].w + g2;
    }
  }
  int idx[4];
  if (MODE == kSearchQuad) {
    // the first column of the best row that holds the minimum
#pragma unroll
    for (int k = 0; k < 4; k++) {
      E e[D];
      const int by = bi[k];
      switch (k) {
        case 0: lds_entries<0, D>(&lds[iy + by][lds_col<E>(ix)], e); break;
        case 1: lds_entries<1, D>(&lds[iy + by][lds_col<E>(ix)], e); break;
        case 2: lds_entries<2, D>(&lds[iy + by][lds_col<E>(ix)], e); break;
        default: lds_entries<3, D>(&lds[iy + by][lds_col<E>(ix)], e); break;
      }
      int bx = 0;
#pragma unroll
      for (int dx = D - 1; dx >= 0; dx--) {
        const int v = quad_score(k, e[dx]);
        bx = v == bk[k] ? dx : bx;
      }
      idx[k] = by * D + bx;
    }
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) idx[k] = KEYED ? ((bk[k] - (iy * D + ix + k)) & (S - 1)) : max(bi[k], 0);
  }
  // Reference-form scores can all equal INT_MAX (wrapped values): then no candidate is taken and the
  // reference keeps zero distances and a zero winner sample (xCorrespPixelShiftSSE.cpp:50-66, :109-124).
  // Keyed scores stay far below INT_MAX (host-checked ranges).
  bool nowin[4];
#pragma unroll
  for (int k = 0; k < 4; k++) nowin[k] = MODE == kSearchQuad ? bk[k] == INT_MAX : (!KEYED && bi[k] < 0);

  // Epilogue: the winners' samples, squared differences, row sums, SCP, winner map.
  // The winner of pixel k, candidate idx = ddy * D + ddx, is region entry (iy + ddy, ix + k + ddx): for
  // 12-byte entries the byte offset (idx + ddy * (LP - D)) * 12 from the pixel's own tile position.
  auto winner = [&](int k) -> E {
    const uint32_t u = (uint32_t)idx[k];
    const uint32_t ddy = u / (uint32_t)D;
    if constexpr (std::is_same<E, SEnt4>::value) {
      return lds[iy + (int)ddy][lds_col<E>(ix + k + (int)(u - ddy * (uint32_t)D))];
    } else {
      const uint32_t off = __umul24(u, 12u) + __umul24(ddy, (uint32_t)((LP - D) * 12));
      return *reinterpret_cast<const E*>(reinterpret_cast<const char*>(&lds[iy][ix + k]) + off);
    }
  };
  // Pixels outside the picture (tiles on the right and bottom edges) add zero.
  const bool edge = x0 + kSearchTileW > W || y0 + 16 > H;  // block-uniform
  const bool row_ok = y < H;
  uint32_t q[3] = {0u, 0u, 0u};
  uint32_t rp[3][2];  // winner samples of pixels (0, 1) and (2, 3) as u16 pairs, per component
  if (KEYED) {
    // Keyed variants keep |t - r| <= Max + |gcd| < 2^15 (host-checked): the
    // differences are int16 pairs and v_dot2 adds their squares.
    uint32_t yc[4], cr[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const E e = winner(k);
      yc[k] = MODE == kSearchDot ? __builtin_bit_cast(uint32_t, __builtin_bit_cast(u16x2_t, e.y) >> (unsigned short)KB)
                                 : (uint32_t)e.y;
      cr[k] = (uint32_t)e.z;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
      rp[0][h] = (yc[2 * h] & 0xffffu) | (yc[2 * h + 1] << 16);
      rp[1][h] = (yc[2 * h] >> 16) | (yc[2 * h + 1] & 0xffff0000u);
      rp[2][h] = cr[2 * h] | (cr[2 * h + 1] << 16);
    }
    short2_t d[3][2];
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int h = 0; h < 2; h++)
        d[c][h] = __builtin_bit_cast(short2_t, tp[c][h]) - __builtin_bit_cast(short2_t, rp[c][h]);
    if (edge) [[unlikely]] {
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const uint32_t msk =
            (row_ok && x + 2 * h < W ? 0xffffu : 0u) | (row_ok && x + 2 * h + 1 < W ? 0xffff0000u : 0u);
#pragma unroll
        for (int c = 0; c < 3; c++) d[c][h] = __builtin_bit_cast(short2_t, __builtin_bit_cast(uint32_t, d[c][h]) & msk);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
      q[c] = (uint32_t)__builtin_amdgcn_sdot2(d[c][1], d[c][1], __builtin_amdgcn_sdot2(d[c][0], d[c][0], 0, false),
                                              false);
  } else {
    int rr[3][4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      sent_raw<MODE, KB>(winner(k), rr[0][k], rr[1][k], rr[2][k]);
      if (nowin[k]) [[unlikely]]
        rr[0][k] = rr[1][k] = rr[2][k] = 0;
      int e0 = t0[k] - rr[0][k], e1 = t1[k] - rr[1][k], e2 = t2[k] - rr[2][k];
      if (nowin[k]) [[unlikely]]
        e0 = e1 = e2 = 0;
      if (edge) {
        const bool ok = row_ok && x + k < W;
        e0 = ok ? e0 : 0;
        e1 = ok ? e1 : 0;
        e2 = ok ? e2 : 0;
      }
      q[0] += (uint32_t)__mul24(e0, e0);
      q[1] += (uint32_t)__mul24(e1, e1);
      q[2] += (uint32_t)__mul24(e2, e2);
    }
#pragma unroll
    for (int c = 0; c < 3; c++)
#pragma unroll
      for (int h = 0; h < 2; h++) rp[c][h] = ((uint32_t)rr[c][2 * h] & 0xffffu) | ((uint32_t)rr[c][2 * h + 1] << 16);
  }
  row16_sum3_u32_dpp(q);
  if ((lane & 15) == 15 && row_ok) {
    uint32_t* rsp = rows + Pr.rows_off + (size_t)dir * 3 * H + y;
    atomicAdd(rsp, q[0]);
    atomicAdd(rsp + H, q[1]);
    atomicAdd(rsp + 2 * H, q[2]);
  }
  if (Pr.best_off >= 0 && row_ok) {
    typedef __attribute__((address_space(1))) uint8_t GU8;
    GU8* bm = (GU8*)best + Pr.best_off + (size_t)dir * W * H + (size_t)y * W + x;
#pragma unroll
    for (int k = 0; k < 4; k++)
      if (x + k < W) bm[k] = (uint8_t)idx[k];
  }
  // dir 0 writes SCP1 (picture 1), dir 1 writes SCP0 (picture 0). The SCP
  // pitch is a multiple of the 64-column tile width, so a lane's 4 samples
  // stay inside the row; columns >= W are padding.
  // SCP planes are below 4 GiB (host-checked): a uniform plane base plus a
  // 32-bit byte offset per lane.
  const bool want_scp = Pr.scp_off >= 0, want_mom = Pr.mom_off >= 0;  // block-uniform
  if (want_scp || want_mom) {
    const int gc[3] = {g0, g1, g2};
    // SCP samples of pixels (0, 1) and (2, 3) per component: clip(r - g, 0, Max)
    // for r in [0, Max] as a saturating u16 subtract of max(g, 0), a saturating
    // add of max(-g, 0) and a min with Max
    u16x2_t sv[3][2];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const u16x2_t gp = (unsigned short)max(gc[c], 0), gn = (unsigned short)max(-gc[c], 0);
      const u16x2_t mx = (unsigned short)Pr.maxv;
#pragma unroll
      for (int h = 0; h < 2; h++)
        sv[c][h] = __builtin_elementwise_min(
            __builtin_elementwise_add_sat(__builtin_elementwise_sub_sat(__builtin_bit_cast(u16x2_t, rp[c][h]), gp), gn),
            mx);
    }
    if (want_scp && row_ok) {
      typedef __attribute__((address_space(1))) uint8_t GU8;
      typedef __attribute__((address_space(1))) u32x2_t GU32x2;
      const size_t plane_bytes = (size_t)Pr.scp_pitch * H * 2;
      GU8* pic = (GU8*)(scp + Pr.scp_off) + (size_t)(1 - dir) * 3 * plane_bytes;
      const uint32_t eo = ((uint32_t)y * (uint32_t)Pr.scp_pitch + (uint32_t)x) * 2u;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const u32x2_t o = {__builtin_bit_cast(uint32_t, sv[c][0]), __builtin_bit_cast(uint32_t, sv[c][1])};
        // streaming store: the SCP is read once, by the SSIM pass, after the whole search
        __builtin_nontemporal_store(o, (GU32x2*)(pic + c * plane_bytes + eo));
      }
    }
    // Keyed variants only (host-checked): the moments' registers fit beside the
    // reference-form variants' larger state without lowering occupancy.
    if (KEYED && want_mom) {
      // Fused IV-SSIM: the wave's 4 rows x 64 columns are 16 sub-blocks of
      // 4 x 4 pixels; lanes l, l ^ 16, l ^ 32, l ^ 48 hold the 4 rows of one.
      // a = the test picture's sample (t - g), b = the SCP sample. Per lane and
      // component {S_a | S_b << 16, S_aa, S_bb, S_ab} over its 4 pixels (host:
      // at most 12-bit samples, so no half overflows), summed over the 4 rows
      // by lane-swap exchanges that leave component g's totals in row group g.
      u16x2_t ta[3][2];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int h = 0; h < 2; h++) ta[c][h] = __builtin_bit_cast(u16x2_t, tp[c][h]) - (unsigned short)gc[c];
      // one component's {S_a | S_b << 16, S_aa, S_bb, S_ab} over the lane's 4 pixels
      auto moments = [&](int c, uint32_t (&m)[4]) {
        const u16x2_t xa = ta[c][0] + ta[c][1], xb = sv[c][0] + sv[c][1];
        const u16x2_t lo = {xa.x, xb.x}, hi = {xa.y, xb.y};
        m[0] = __builtin_bit_cast(uint32_t, lo + hi);
        m[1] = __builtin_amdgcn_udot2(ta[c][1], ta[c][1], __builtin_amdgcn_udot2(ta[c][0], ta[c][0], 0u, false), false);
        m[2] = __builtin_amdgcn_udot2(sv[c][1], sv[c][1], __builtin_amdgcn_udot2(sv[c][0], sv[c][0], 0u, false), false);
        m[3] = __builtin_amdgcn_udot2(ta[c][1], sv[c][1], __builtin_amdgcn_udot2(ta[c][0], sv[c][0], 0u, false), false);
      };
      uint32_t s0[4], s2[4], p[4];
      moments(0, s0);
      moments(2, s2);
#pragma unroll
      for (int k = 0; k < 4; k++) p[k] = swap_sum32(s0[k], s2[k]);  // lanes 0-31: component 0, 32-63: component 2
      uint32_t s1[4];
      moments(1, s1);
      u32x4_t fin;
#pragma unroll
      for (int k = 0; k < 4; k++) fin[k] = swap_sum16(p[k], swap_sum32(s1[k], s1[k]));  // rows 0, 1, 2: components 0, 1, 2
      const int g = lane >> 4;
      const int sby = (y0 >> 2) + wd, sbx = (x0 >> 2) + (lane & 15);
      if (g < 3 && sby < Pr.mom_rows && sbx < Pr.mom_cols) {
        typedef __attribute__((address_space(1))) u32x4_t GU32x4;
        GU32x4* dst = (GU32x4*)(scp + Pr.mom_off) +
                      (((int64_t)dir * 3 + g) * Pr.mom_rows + sby) * (int64_t)Pr.mom_pitch + sbx;
        // read once, by the SSIM pass, after the whole search
        __builtin_nontemporal_store(fin, dst);
      }
    }
  }
}

// One kernel per form of the search arithmetic, so no form's register allocation is imposed on another:
// FORM 0 the keyed modes (kSearchNarrow, kSearchDot), 1 the reference forms (kSearchWide, kSearchWide24,
// kSearchWide24D), 2 kSearchQuad. Each walks all tiles of the batch and returns at once from tiles of
// pairs of another form (block-uniform).
__host__ __device__ constexpr int search_form(int mode) {
  return mode == kSearchNarrow || mode == kSearchDot ? 0 : (mode == kSearchQuad ? 2 : 1);
}
template <int R, int FORM>
__global__ __launch_bounds__(512) void k_search(const PairDev* __restrict__ pairs, int npairs, int ntiles, int wY,
                                                int wCb, int wCr, const int32_t* __restrict__ gcd,
                                                uint32_t* __restrict__ rows, uint16_t* __restrict__ scp,
                                                uint8_t* __restrict__ best) {
  using TT = SearchTile<R>;
  using E = typename std::conditional<FORM == 2, SEnt4, SEnt>::type;
  __shared__ __attribute__((aligned(16))) E lds[2][TT::LH][LdsPitch<E, TT::LP>::v];
  // XCD-aware order: workgroups are handed to the kXcds XCDs in turn, so XCD g
  // runs tiles g * per, g * per + 1, ...: each XCD walks a contiguous run of
  // tiles, whose halos overlap in its L2.
  const int per = (ntiles + kXcds - 1) / kXcds;
  const int tile = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (tile >= ntiles) return;  // padding of the last XCD's run (uniform per block)
  const int pi = find_entry(pairs, npairs, tile, [](const PairDev& p) { return p.tile0; });
  const PairDev P = pairs[pi];
  if (search_form(P.mode) != FORM) return;
  const int lt = tile - P.tile0;
  const int ntx = (P.w + TT::TW - 1) / TT::TW;
  const int ty = lt / ntx, tx = lt - ty * ntx;
  const int x0 = tx * TT::TW, y0 = ty * TT::TH;
#define IVQM_TILE(T, MODE) search_tile<T, R, MODE>(P, x0, y0, lds, wY, wCb, wCr, gcd, rows, scp, best)
  if constexpr (FORM == 0) {
    if (P.p0.bps == 2) {
      if (P.mode == kSearchDot) IVQM_TILE(uint16_t, kSearchDot);
      else IVQM_TILE(uint16_t, kSearchNarrow);
    } else {
      if (P.mode == kSearchDot) IVQM_TILE(uint8_t, kSearchDot);
      else IVQM_TILE(uint8_t, kSearchNarrow);
    }
  } else if constexpr (FORM == 2) {
    if (P.p0.bps == 2) IVQM_TILE(uint16_t, kSearchQuad);
    else IVQM_TILE(uint8_t, kSearchQuad);
  } else {
    if (P.p0.bps == 2) {
      if (P.mode == kSearchWide24D) IVQM_TILE(uint16_t, kSearchWide24D);
      else if (P.mode == kSearchWide24) IVQM_TILE(uint16_t, kSearchWide24);
      else IVQM_TILE(uint16_t, kSearchWide);
    } else {
      if (P.mode == kSearchWide24D) IVQM_TILE(uint8_t, kSearchWide24D);
      else if (P.mode == kSearchWide24) IVQM_TILE(uint8_t, kSearchWide24);
      else IVQM_TILE(uint8_t, kSearchWide);
    }
  }
#undef IVQM_TILE
}

// ---------------------------------------------------------------------------
// k_ssim: block-averaged SSIM, window 8, stride 4.
// libs/pmbb-ivqm/src/xStructSimSTD.cpp:420-461 (window), xSSIM.cpp:238-254
// (row loop), :381-404 (window counts). Each wave owns 64 sub-block columns
// (63 windows) and a band of NB window rows; it walks the band top to bottom,
// carrying the previous 4x4 sub-block row so every input row is read once per
// band. Window moments are exact integers; the ratio is FP32 on the exact
// numerators n*S_ab - S_a*S_b (formed in FP64, where they are exact, and
// rounded once); window values are summed in FP64.
// ---------------------------------------------------------------------------
// Four samples of a picture row (luma grid x .. x + 3, x a multiple of 4) from
// a wave-uniform row pointer: one vector load; 4:2:0 chroma reads 2 samples
// (one load when the pair is aligned) and replicates them.
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned char u8x4 __attribute__((ext_vector_type(4)));
typedef unsigned char u8x2 __attribute__((ext_vector_type(2)));

template <typename T, int CS>
__device__ __forceinline__ void load4(const __attribute__((address_space(1))) T* row, int x, uint32_t v[4]) {
  // 32-bit byte offsets from the (scalar) row pointer
  typedef __attribute__((address_space(1))) const uint8_t GB;
  if (CS == 0) {
    GB* q = (GB*)row + (uint32_t)x * (uint32_t)sizeof(T);
    if (sizeof(T) == 2) {
      const u16x4 u = *(const __attribute__((address_space(1))) u16x4*)q;
      v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
    } else {
      const u8x4 u = *(const __attribute__((address_space(1))) u8x4*)q;
      v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
    }
  } else {
    GB* q = (GB*)row + ((uint32_t)x >> 1) * (uint32_t)sizeof(T);
    uint32_t a, b;
    if ((reinterpret_cast<uintptr_t>(q) & (2 * sizeof(T) - 1)) == 0) {  // uniform: x >> 1 is even
      if (sizeof(T) == 2) {
        const u16x2 u = *(const __attribute__((address_space(1))) u16x2*)q;
        a = u.x; b = u.y;
      } else {
        const u8x2 u = *(const __attribute__((address_space(1))) u8x2*)q;
        a = u.x; b = u.y;
      }
    } else {
      a = ((const __attribute__((address_space(1))) T*)q)[0];
      b = ((const __attribute__((address_space(1))) T*)q)[1];
    }
    v[0] = a; v[1] = a; v[2] = b; v[3] = b;
  }
}

// Value of lane + 1 (DPP wave shift; no LDS traffic). Lane 63 reads 0: its source lane is out of range and
// bound_ctrl makes such a read 0, so no old value is set up (and an add of the result takes the shift as
// its own DPP operand).
__device__ __forceinline__ uint32_t next_lane32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ uint64_t next_lane64(uint64_t v) {
  return (uint64_t)next_lane32((uint32_t)v) | ((uint64_t)next_lane32((uint32_t)(v >> 32)) << 32);
}

// Sum widths of the window moments (host picks by bit depth): SM 0 up to 12-bit, everything in
// uint32 (64 * 4095^2 < 2^32); SM 1 up to 14-bit, sub-blocks uint32, windows uint64; SM 2 uint64.
// MOM: the sub-block sums come from the search's fused moments (J.mom; SM 0) instead of the pictures --
// the same integers, so the same scores bit for bit.
// Each lane sums its windows' scores over a group of 4 window rows (bands start at multiples of 4, so a group
// lies in one band), weighted by their rows' ERP weights (weight 1 without), and the group's wave total is its
// partial: the grouping does not depend on the band height, so every band height gives the same doubles.
template <typename TA, int CSA, typename TB, int CSB, int SM, int NB, bool MOM = false>
__device__ void ssim_wave(const SsimJobDev& J, int c, int chunk, int band, int lane, float c1n, float c2n,
                          const double* __restrict__ erp_w, double* __restrict__ partial) {
  static_assert(!MOM || SM == 0, "fused moments are 32-bit");
  using SQ = typename std::conditional<SM == 2, uint64_t, uint32_t>::type;  // sub-block squared sums
  using WQ = typename std::conditional<SM == 0, uint32_t, uint64_t>::type;  // window squared sums
  const int wy0 = band * NB;
  if (wy0 >= J.nuy) return;
  const int bx = chunk * 63 + lane;
  const bool col_ok = bx <= J.nux;
  typedef __attribute__((address_space(1))) const TA GTA;
  typedef __attribute__((address_space(1))) const TB GTB;
  GTA* pa = (GTA*)plane_ptr(J.a, c);
  GTB* pb = (GTB*)plane_ptr(J.b, c);
  const int sa = plane_stride(J.a, c), sb = plane_stride(J.b, c);
  uint32_t pA = 0, pB = 0;
  SQ pAA = 0, pBB = 0, pAB = 0;
  static_assert(NB % 4 == 0, "bands hold whole groups of 4 window rows");
  const int ngrp = (J.nuy + 3) >> 2;
  const bool weighted = J.erp && J.erp_off >= 0;
  const double* wrow = erp_w + (weighted ? J.erp_off + J.erp_base : 0);
  double acc = 0.0, acc_cs = 0.0;
#pragma unroll 1
  for (int k = 0; k <= NB; k++) {
    const int by = wy0 + k;
    if (by > J.nuy) break;
    uint32_t A = 0, B = 0;
    SQ AA = 0, BB = 0, AB = 0;
    if (MOM) {
      if (col_ok) {
        typedef __attribute__((address_space(1))) const u32x4_t GM;
        const u32x4_t m = __builtin_nontemporal_load(
            (GM*)J.mom + ((int64_t)c * J.mom_rows + by) * (int64_t)J.mom_pitch + bx);
        A = m.x & 0xffffu;
        B = m.x >> 16;
        AA = (SQ)m.y;
        BB = (SQ)m.z;
        AB = (SQ)m.w;
      }
    } else if (col_ok) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        uint32_t a[4], b[4];
        const int y = by * 4 + r;  // wave-uniform row
        load4<TA, (CSA != 0)>(pa + (size_t)(y >> (CSA == 1 ? 1 : 0)) * sa, bx * 4, a);
        load4<TB, (CSB != 0)>(pb + (size_t)(y >> (CSB == 1 ? 1 : 0)) * sb, bx * 4, b);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          A += a[j];
          B += b[j];
          // __umul24 returns int: widen through uint32_t (16-bit squares reach 2^32 - 2^17)
          AA += (SQ)(uint32_t)__umul24(a[j], a[j]);
          BB += (SQ)(uint32_t)__umul24(b[j], b[j]);
          AB += (SQ)(uint32_t)__umul24(a[j], b[j]);
        }
      }
    }
    if (k > 0) {
      const int wy = by - 1;
      // window = this lane's and the next lane's columns of the previous and this sub-block row
      const uint32_t vA = pA + A, vB = pB + B;
      const WQ vAA = (WQ)pAA + AA, vBB = (WQ)pBB + BB, vAB = (WQ)pAB + AB;
      uint32_t nA = next_lane32(vA), nB = next_lane32(vB);
      WQ nAA, nBB, nAB;
      if (SM == 0) {
        nAA = (WQ)next_lane32((uint32_t)vAA);
        nBB = (WQ)next_lane32((uint32_t)vBB);
        nAB = (WQ)next_lane32((uint32_t)vAB);
      } else {
        nAA = (WQ)next_lane64((uint64_t)vAA);
        nBB = (WQ)next_lane64((uint64_t)vBB);
        nAB = (WQ)next_lane64((uint64_t)vAB);
      }
      // Every lane scores (the sums of lane 63 and of columns past the last window are finite, the
      // denominators positive) and the scores of windows that do not exist are dropped: the window sums
      // take the lane shift as a DPP operand of their adds.
      const bool valid = lane < 63 && bx < J.nux;
      const uint32_t sA = vA + nA, sB = vB + nB;  // < 2^23
      // exact numerators n S_ab - S_a S_b, each rounded once to FP32: in FP64 every term is an integer
      // below 2^53 (S_a S_b < 2^46, n S_xy < 2^45), so the products and the difference are exact
      const double dA = (double)sA, dB = (double)sB;
      const float fcov = (float)fma(64.0, (double)(WQ)(vAB + nAB), -(dA * dB));
      const float fva = (float)fma(64.0, (double)(WQ)(vAA + nAA), -(dA * dA));
      const float fvb = (float)fma(64.0, (double)(WQ)(vBB + nBB), -(dB * dB));
      float num = 2.0f * fcov + c2n;
      float den = fva + fvb + c2n;
      // the twin's score is this job's without the luminance term: the same operations as its own job
      double s_cs = 0.0;
      if (J.twin_part_off >= 0) s_cs = valid ? (double)(num * __builtin_amdgcn_rcpf(den)) : 0.0;
      if (J.calc_l) {  // luminance term: positive terms only, FP32 products suffice
        const float fa = (float)sA, fb = (float)sB;
        num *= 2.0f * fa * fb + c1n;
        den *= fa * fa + fb * fb + c1n;
      }
      const double s = valid ? (double)(num * __builtin_amdgcn_rcpf(den)) : 0.0;
      acc = weighted ? fma(s, wrow[(int64_t)wy * J.stride], acc) : acc + s;
      acc_cs += s_cs;  // the twin (MS-SSIM level 0) is never ERP-weighted
      if ((wy & 3) == 3 || wy == J.nuy - 1) {  // wave-uniform: the group's last row
        const int64_t pi = ((int64_t)c * ngrp + (wy >> 2)) * J.nchunk + chunk;
        acc = wave_sum_f64_dpp(acc);
        if (lane == 0) partial[J.part_off + pi] = acc;
        acc = 0.0;
        if (J.twin_part_off >= 0) {  // block-uniform
          acc_cs = wave_sum_f64_dpp(acc_cs);
          if (lane == 0) partial[J.twin_part_off + pi] = acc_cs;
          acc_cs = 0.0;
        }
      }
    }
    pA = A; pB = B; pAA = AA; pBB = BB; pAB = AB;
  }
}

// CSA / CSB: chroma format of picture a / b for component c (0 4:4:4, 1 4:2:0, 2 4:2:2; luma 0). A job
// pairs two caller pictures of one format, or a caller picture with a 4:4:4 SCP picture (either order).
template <int SM, int NB>
__device__ void ssim_dispatch(const SsimJobDev& J, int c, int chunk, int band, int lane, float c1n, float c2n,
                              const double* erp_w, double* partial) {
  const int csa = c ? chroma_format(J.a) : 0, csb = c ? chroma_format(J.b) : 0;
#define IVQM_SSIM(TA, A, TB, B) ssim_wave<TA, A, TB, B, SM, NB>(J, c, chunk, band, lane, c1n, c2n, erp_w, partial)
  if (J.a.bps == 2 && J.b.bps == 2) {
    if (csa == 0 && csb == 0) IVQM_SSIM(uint16_t, 0, uint16_t, 0);
    else if (csa == 1 && csb == 1) IVQM_SSIM(uint16_t, 1, uint16_t, 1);
    else if (csa == 1 && csb == 0) IVQM_SSIM(uint16_t, 1, uint16_t, 0);
    else if (csa == 0 && csb == 1) IVQM_SSIM(uint16_t, 0, uint16_t, 1);
    else if (csa == 2 && csb == 2) IVQM_SSIM(uint16_t, 2, uint16_t, 2);
    else if (csa == 2 && csb == 0) IVQM_SSIM(uint16_t, 2, uint16_t, 0);
    else IVQM_SSIM(uint16_t, 0, uint16_t, 2);
  } else if (SM < 2) {
    if (J.a.bps == 1 && J.b.bps == 1) {
      if (csa == 0 && csb == 0) IVQM_SSIM(uint8_t, 0, uint8_t, 0);
      else if (csa == 1) IVQM_SSIM(uint8_t, 1, uint8_t, 1);
      else IVQM_SSIM(uint8_t, 2, uint8_t, 2);
    } else if (J.a.bps == 1) {
      if (csa == 0) IVQM_SSIM(uint8_t, 0, uint16_t, 0);
      else if (csa == 1) IVQM_SSIM(uint8_t, 1, uint16_t, 0);
      else IVQM_SSIM(uint8_t, 2, uint16_t, 0);
    }
  }
#undef IVQM_SSIM
}

// One launch per job form, so no form's register allocation is imposed on another: FORM 0 the jobs
// that read pictures with 32-bit sums (up to 12-bit samples), 1 those with wider sums, 2 the jobs that
// read fused moments; blocks of another form exit at once.
template <int NB, int FORM>
__global__ __launch_bounds__(256) void k_ssim(const SsimJobDev* __restrict__ jobs, int njobs, int nblk,
                                              const double* __restrict__ erp_w, double* __restrict__ partial) {
  constexpr bool MOM = FORM == 2;
  // XCD-aware order (as in k_search): XCD g runs blocks g * per, g * per + 1, ...,
  // so neighbouring chunks and band groups, which share cache lines of their
  // rows, run on one XCD close together and meet in its L2.
  const int per = (nblk + kXcds - 1) / kXcds;
  const int b = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (b >= nblk) return;  // padding of the last XCD's run (uniform per block)
  const int ji = find_entry(jobs, njobs, b, [](const SsimJobDev& j) { return j.blk0; });
  const SsimJobDev J = jobs[ji];
  if ((J.mom != nullptr) != MOM) return;  // block-uniform
  if (!MOM && (J.wsq == 0) != (FORM == 0)) return;
  const int lb = b - J.blk0;
  const int per_c = J.nchunk * J.nbg;
  const int c = lb / per_c, rem = lb - c * per_c;
  const int bg = rem / J.nchunk, chunk = rem - bg * J.nchunk;
  // band is wave-uniform: row pointers stay scalar and every load adds only the lane's column
  const int lane = threadIdx.x & 63, band = bg * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if constexpr (FORM == 2)
    ssim_wave<uint16_t, 0, uint16_t, 0, 0, NB, true>(J, c, chunk, band, lane, J.c1n, J.c2n, erp_w, partial);
  else if constexpr (FORM == 0) ssim_dispatch<0, NB>(J, c, chunk, band, lane, J.c1n, J.c2n, erp_w, partial);
  else if (J.wsq == 1) ssim_dispatch<1, NB>(J, c, chunk, band, lane, J.c1n, J.c2n, erp_w, partial);
  else ssim_dispatch<2, NB>(J, c, chunk, band, lane, J.c1n, J.c2n, erp_w, partial);
}

// One block per (job, component): FP64 frame totals of the window-row sums.
// IV-SSIM ERP weighting follows libs/pmbb-ivqm/src/xIVSSIM.cpp:148-165: the
// weight of window row j is taken at its centre row (regular) or at its top
// row plus half the window (block).
__global__ __launch_bounds__(256) void k_ssim_finalize(const SsimJobDev* __restrict__ jobs,
                                                       const double* __restrict__ partial,
                                                       const double* __restrict__ erp_w, double* __restrict__ out) {
  const int j = blockIdx.x / 3, c = blockIdx.x - 3 * (blockIdx.x / 3);
  const SsimJobDev J = jobs[j];
  const int tid = threadIdx.x;
  double acc = 0.0, accw = 0.0;
  const bool weighted = J.erp && J.erp_off >= 0;
  const int nrow = (J.nuy + J.grp - 1) / J.grp;  // partial rows
  for (int r = tid; r < nrow; r += 256) {
    const double* p = partial + J.part_off + ((int64_t)c * nrow + r) * J.nchunk;
    double rs = 0.0;
    for (int k = 0; k < J.nchunk; k++) rs += p[k];
    if (weighted && J.grp == 1) rs *= erp_w[J.erp_off + J.erp_base + (int64_t)r * J.stride];
    acc += rs;
  }
  if (weighted)  // the weight total
    for (int wy = tid; wy < J.nuy; wy += 256) accw += erp_w[J.erp_off + J.erp_base + (int64_t)wy * J.stride];
  __shared__ double s_a[4], s_w[4];
  acc = wave_sum_f64(acc);
  accw = wave_sum_f64(accw);
  if ((tid & 63) == 0) { s_a[tid >> 6] = acc; s_w[tid >> 6] = accw; }
  __syncthreads();
  if (tid == 0) {
    out[(int64_t)J.out_idx * 6 + c * 2 + 0] = (s_a[0] + s_a[1]) + (s_a[2] + s_a[3]);
    out[(int64_t)J.out_idx * 6 + c * 2 + 1] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  }
}

// 2x2 rounded mean: libs/pmbb-core/src/xPixelOpsSTD.cpp:50-67.
template <typename T>
__device__ __forceinline__ void down_px(const PicDev& S, int c, int x, int y, uint16_t* dst, int pitch, int dh) {
  const int v = ldpix<T>(S, c, 2 * x, 2 * y) + ldpix<T>(S, c, 2 * x + 1, 2 * y) + ldpix<T>(S, c, 2 * x, 2 * y + 1) +
                ldpix<T>(S, c, 2 * x + 1, 2 * y + 1) + 2;
  dst[(size_t)c * pitch * dh + (size_t)y * pitch + x] = (uint16_t)(v >> 2);
}

// M consecutive samples from p with the widest loads the address allows (16, 8 or 4 bytes; single
// samples otherwise). The alignment of p is the row's: every lane's offset is a multiple of M samples, so
// the branch is wave-uniform.
template <typename T, int M>
__device__ __forceinline__ void load_samples(const __attribute__((address_space(1))) T* p, uint32_t (&v)[M]) {
  typedef __attribute__((address_space(1))) const uint8_t GB;
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  constexpr int BYTES = M * (int)sizeof(T);
  auto unpack = [&](uint32_t w, int i0) {  // one 32-bit word: 2 (u16) or 4 (u8) samples
    if (sizeof(T) == 2) {
      v[i0] = w & 0xffffu;
      v[i0 + 1] = w >> 16;
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++) v[i0 + k] = (w >> (8 * k)) & 0xffu;
    }
  };
  constexpr int SPW = 4 / (int)sizeof(T);  // samples per 32-bit word
  if (BYTES % 16 == 0 && (a & 15) == 0) {
#pragma unroll
    for (int i = 0; i < BYTES / 16; i++) {
      const u32x4_t w = *(const __attribute__((address_space(1))) u32x4_t*)((GB*)p + 16 * i);
      unpack(w.x, (4 * i) * SPW);
      unpack(w.y, (4 * i + 1) * SPW);
      unpack(w.z, (4 * i + 2) * SPW);
      unpack(w.w, (4 * i + 3) * SPW);
    }
  } else if (BYTES % 8 == 0 && (a & 7) == 0) {
#pragma unroll
    for (int i = 0; i < BYTES / 8; i++) {
      typedef unsigned int u32x2d_t __attribute__((ext_vector_type(2)));
      const u32x2d_t w = *(const __attribute__((address_space(1))) u32x2d_t*)((GB*)p + 8 * i);
      unpack(w.x, (2 * i) * SPW);
      unpack(w.y, (2 * i + 1) * SPW);
    }
  } else if (BYTES % 4 == 0 && (a & 3) == 0) {
#pragma unroll
    for (int i = 0; i < BYTES / 4; i++) unpack(*(const __attribute__((address_space(1))) uint32_t*)((GB*)p + 4 * i), i * SPW);
  } else {
#pragma unroll
    for (int k = 0; k < M; k++) v[k] = p[k];
  }
}

// Eight adjacent outputs x8 .. x8 + 7 of row y (x8 + 7 < dw): the two source rows' 16 samples each with the
// widest loads their alignment allows (wave-uniform: one row per wave), one 16-byte store (pyramid rows are
// 256-byte aligned). A 4:2:0 source's chroma is the level-1 chroma itself (the 2x2 block replicates one
// sample), so it is copied; a 4:2:2 source's 2x2 block holds two samples twice each, one above the other.
template <typename T>
__device__ __forceinline__ void down_x8(const PicDev& S, int c, int x8, int y, uint16_t* dst, int pitch, int dh) {
  typedef __attribute__((address_space(1))) const T GT;
  uint32_t o[8];
  const int cf = c ? chroma_format(S) : 0;
  if (cf == 1) {
    load_samples<T, 8>((GT*)plane_ptr(S, c) + (size_t)y * plane_stride(S, c) + x8, o);
  } else if (cf == 2) {
    GT* r0 = (GT*)plane_ptr(S, c) + (size_t)(2 * y) * plane_stride(S, c) + x8;
    uint32_t a[8], b[8];
    load_samples<T, 8>(r0, a);
    load_samples<T, 8>(r0 + plane_stride(S, c), b);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = (2 * a[k] + 2 * b[k] + 2) >> 2;
  } else {
    GT* r0 = (GT*)plane_ptr(S, c) + (size_t)(2 * y) * plane_stride(S, c) + 2 * x8;
    uint32_t a[16], b[16];
    load_samples<T, 16>(r0, a);
    load_samples<T, 16>(r0 + plane_stride(S, c), b);
#pragma unroll
    for (int k = 0; k < 8; k++) o[k] = (a[2 * k] + a[2 * k + 1] + b[2 * k] + b[2 * k + 1] + 2) >> 2;
  }
  const u32x4_t v = {o[0] | (o[1] << 16), o[2] | (o[3] << 16), o[4] | (o[5] << 16), o[6] | (o[7] << 16)};
  *(__attribute__((address_space(1))) u32x4_t*)(dst + (size_t)c * pitch * dh + (size_t)y * pitch + x8) = v;
}

// Block = 4 rows x 512 outputs (a lane makes 8 adjacent outputs); the last partial group of a row
// takes the one-output form.
__global__ __launch_bounds__(256) void k_downsample(const DownJob* __restrict__ jobs, int njobs) {
  const int b = blockIdx.x;
  const int ji = find_entry(jobs, njobs, b, [](const DownJob& j) { return j.blk0; });
  const DownJob J = jobs[ji];
  const int dw = J.w >> 1, dh = J.h >> 1;
  const int nbx = (dw + kDownBlkW - 1) / kDownBlkW, nby = (dh + kDownBlkH - 1) / kDownBlkH;
  int lb = b - J.blk0;
  const int c = lb / (nbx * nby);
  if (c >= J.ncomp) return;  // the job writes luma only (level 1 with the caller's chroma planes)
  lb -= c * nbx * nby;
  const int by = lb / nbx, bx = lb - by * nbx;
  const int x8 = bx * kDownBlkW + 8 * (threadIdx.x & 63);
  const int y = by * kDownBlkH + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (x8 >= dw || y >= dh) return;
  if (x8 + 7 < dw) {
    if (J.src.bps == 2) down_x8<uint16_t>(J.src, c, x8, y, J.dst, J.pitch, dh);
    else down_x8<uint8_t>(J.src, c, x8, y, J.dst, J.pitch, dh);
  } else {
    for (int x = x8; x < dw; x++) {
      if (J.src.bps == 2) down_px<uint16_t>(J.src, c, x, y, J.dst, J.pitch, dh);
      else down_px<uint8_t>(J.src, c, x, y, J.dst, J.pitch, dh);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_downsample_one(PicDev S, int dw, int dh, uint16_t* __restrict__ dst, int pitch) {
  const int c = blockIdx.z;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= dw || y >= dh) return;
  down_px<T>(S, c, x, y, dst, pitch, dh);
}

// 4:2:0 -> 4:4:4: libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49 (luma copied).
template <typename T>
__global__ __launch_bounds__(256) void k_upsample420(PicDev S, int w, int h, uint16_t* __restrict__ dst, int dst_stride) {
  const int c = blockIdx.z;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  dst[(size_t)c * dst_stride * h + (size_t)y * dst_stride + x] = (uint16_t)ldpix<T>(S, c, x, y);
}

// Colour-space conversion to 4:4:4 planes (standalone, before the metrics): the fixed-point
// kernels the reference dispatches, xColorSpaceSTD::ConvertYCbCr2RGB_I32 / ConvertRGB2YCbCr_I32
// (libs/pmbb-core/src/xColorSpaceSTD.cpp), coefficients xColorSpaceCoeff.h:59-73. The int32
// arithmetic wraps (16-bit samples) as the reference's two's-complement code does.
struct ColourCoef {
  int32_t k[9];  // RGB -> YCbCr: Y_R, Y_G, Y_B, U_R, U_G, V_G, V_B (7 used); YCbCr -> RGB: R_V, G_U, G_V, B_U
};
__device__ __forceinline__ int32_t clipu(int32_t v, int32_t mx) { return min(max(v, 0), mx); }

template <typename T, int DIR>
__global__ __launch_bounds__(256) void k_colour(PicDev S, int w, int h, int bit_depth, ColourCoef K,
                                                uint16_t* __restrict__ dst, int dst_stride) {
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  const int32_t mid = 1 << (bit_depth - 1), mx = (1 << bit_depth) - 1;
  const uint32_t a = (uint32_t)ldpix<T>(S, 0, x, y), b = (uint32_t)ldpix<T>(S, 1, x, y),
                 c = (uint32_t)ldpix<T>(S, 2, x, y);
  int32_t o0, o1, o2;
  if (DIR == 1) {  // RGB -> YCbCr (a, b, c = R, G, B)
    const int32_t ty = (int32_t)((uint32_t)K.k[0] * a + (uint32_t)K.k[1] * b + (uint32_t)K.k[2] * c + 32768u) >> 16;
    const int32_t tu = (int32_t)((uint32_t)K.k[3] * a + (uint32_t)K.k[4] * b + (c << 15) + 32768u) >> 16;
    const int32_t tv = (int32_t)((a << 15) + (uint32_t)K.k[5] * b + (uint32_t)K.k[6] * c + 32768u) >> 16;
    o0 = clipu(ty, mx);
    o1 = clipu(tu + mid, mx);
    o2 = clipu(tv + mid, mx);
  } else {  // YCbCr -> RGB
    const int32_t iu = (int32_t)b - mid, iv = (int32_t)c - mid;
    const uint32_t sy = (a << 16) + 32768u;
    o0 = clipu((int32_t)(sy + (uint32_t)K.k[0] * (uint32_t)iv) >> 16, mx);
    o1 = clipu((int32_t)(sy + (uint32_t)K.k[1] * (uint32_t)iu + (uint32_t)K.k[2] * (uint32_t)iv) >> 16, mx);
    o2 = clipu((int32_t)(sy + (uint32_t)K.k[3] * (uint32_t)iu) >> 16, mx);
  }
  const size_t plane = (size_t)dst_stride * h, o = (size_t)y * dst_stride + x;
  dst[o] = (uint16_t)o0;
  dst[plane + o] = (uint16_t)o1;
  dst[2 * plane + o] = (uint16_t)o2;
}

}  // namespace

// ------------------------------- launchers ---------------------------------

hipError_t launch_colour(const PicDev& src, int w, int h, int bit_depth, int direction, int space, uint16_t* dst,
                         int dst_stride, hipStream_t st) {
  // xColorSpaceCoeffYCbCr::c_RGB2YCbCr_I32 / c_YCbCr2RGB_I32 (xColorSpaceCoeff.h:59-73): BT.601
  // (= SMPTE170M, JPEG), BT.709, SMPTE240M, BT.2020
  static const ColourCoef kFwd[4] = {{{19595, 38470, 7471, -11058, -21710, -27439, -5329}},
                                     {{13933, 46871, 4732, -7509, -25259, -29763, -3005}},
                                     {{13894, 45941, 5701, -7609, -25159, -29150, -3618}},
                                     {{17216, 44434, 3886, -9151, -23617, -30133, -2635}}};
  static const ColourCoef kInv[4] = {{{91881, -22553, -46802, 116130}},
                                     {{103206, -12276, -30679, 121609}},
                                     {{103285, -14852, -31236, 119669}},
                                     {{96639, -10784, -37444, 123299}}};
  if (w <= 0 || h <= 0 || space < 0 || space > 3 || (direction != 0 && direction != 1)) return hipErrorInvalidValue;
  const dim3 grid((w + 63) / 64, (h + 3) / 4);
  const ColourCoef K = direction ? kFwd[space] : kInv[space];
  if (src.bps == 2) {
    if (direction) hipLaunchKernelGGL((k_colour<uint16_t, 1>), grid, dim3(256), 0, st, src, w, h, bit_depth, K, dst, dst_stride);
    else hipLaunchKernelGGL((k_colour<uint16_t, 0>), grid, dim3(256), 0, st, src, w, h, bit_depth, K, dst, dst_stride);
  } else {
    if (direction) hipLaunchKernelGGL((k_colour<uint8_t, 1>), grid, dim3(256), 0, st, src, w, h, bit_depth, K, dst, dst_stride);
    else hipLaunchKernelGGL((k_colour<uint8_t, 0>), grid, dim3(256), 0, st, src, w, h, bit_depth, K, dst, dst_stride);
  }
  return hipGetLastError();
}

hipError_t launch_stats(const PairDev* pairs, int npairs, int nblk, int64_t* row_sd, uint64_t* row_ssd,
                        uint32_t* rows, int64_t* aux, uint64_t* rows64, uint32_t forms, hipStream_t st) {
  if (npairs <= 0 || nblk <= 0) return hipSuccess;
  // nblk = total plane rows; one wave per row, grid-stride. Each form's grid fills the chip once:
  // form 0 (64 VGPRs) at 8 waves per SIMD (256 CUs x 32 waves = 2048 blocks of 4 waves), form 1 (73 VGPRs) at 6
  if (forms & 1u) {
    const int grid = std::min((nblk + 3) / 4, 2048);
    hipLaunchKernelGGL(k_stats<0>, dim3(grid), dim3(256), 0, st, pairs, npairs, nblk, row_sd, row_ssd, rows, aux, rows64);
  }
  if (forms & 2u) {
    const int grid = std::min((nblk + 3) / 4, 1536);
    hipLaunchKernelGGL(k_stats<1>, dim3(grid), dim3(256), 0, st, pairs, npairs, nblk, row_sd, row_ssd, rows, aux, rows64);
  }
  return hipGetLastError();
}

hipError_t launch_gcd_finalize(const PairDev* pairs, int npairs, const int64_t* row_sd, const uint64_t* row_ssd,
                               const int64_t* aux, int32_t* gcd, int64_t* sd, uint64_t* ssd, int64_t* mcount,
                               hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_gcd_finalize, dim3(npairs), dim3(256), 0, st, pairs, row_sd, row_ssd, aux, gcd, sd, ssd,
                     mcount);
  return hipGetLastError();
}

template <int R>
static hipError_t search_r(const PairDev* pairs, int npairs, int ntiles, const int32_t cw[4], const int32_t* gcd,
                           uint32_t* rows, uint16_t* scp, uint8_t* best, uint32_t forms, hipStream_t st) {
  const int per = (ntiles + kXcds - 1) / kXcds;
  if (forms & 1u)
    hipLaunchKernelGGL((k_search<R, 0>), dim3(kXcds * per), dim3(512), 0, st, pairs, npairs, ntiles, cw[0], cw[1],
                       cw[2], gcd, rows, scp, best);
  if (forms & 2u)
    hipLaunchKernelGGL((k_search<R, 1>), dim3(kXcds * per), dim3(512), 0, st, pairs, npairs, ntiles, cw[0], cw[1],
                       cw[2], gcd, rows, scp, best);
  if (forms & 4u)
    hipLaunchKernelGGL((k_search<R, 2>), dim3(kXcds * per), dim3(512), 0, st, pairs, npairs, ntiles, cw[0], cw[1],
                       cw[2], gcd, rows, scp, best);
  return hipGetLastError();
}

hipError_t launch_search(const PairDev* pairs, int npairs, int ntiles, int R, const int32_t cw[4],
                         const int32_t* gcd, uint32_t* rows, uint16_t* scp, uint8_t* best, uint32_t forms,
                         hipStream_t st) {
  if (npairs <= 0 || ntiles <= 0) return hipSuccess;
  switch (R) {
    case 1: return search_r<1>(pairs, npairs, ntiles, cw, gcd, rows, scp, best, forms, st);
    case 2: return search_r<2>(pairs, npairs, ntiles, cw, gcd, rows, scp, best, forms, st);
    case 3: return search_r<3>(pairs, npairs, ntiles, cw, gcd, rows, scp, best, forms, st);
    case 4: return search_r<4>(pairs, npairs, ntiles, cw, gcd, rows, scp, best, forms, st);
    default: return hipErrorInvalidValue;
  }
}

template <int FORM>
static hipError_t ssim_launch(const SsimJobDev* jobs, int njobs, int nblk, int nb, const double* erp_w, double* partial,
                              hipStream_t st) {
  const int grid = kXcds * ((nblk + kXcds - 1) / kXcds);
  switch (nb) {
    case 4: hipLaunchKernelGGL((k_ssim<4, FORM>), dim3(grid), dim3(256), 0, st, jobs, njobs, nblk, erp_w, partial); break;
    case 8: hipLaunchKernelGGL((k_ssim<8, FORM>), dim3(grid), dim3(256), 0, st, jobs, njobs, nblk, erp_w, partial); break;
    case 16:
      hipLaunchKernelGGL((k_ssim<16, FORM>), dim3(grid), dim3(256), 0, st, jobs, njobs, nblk, erp_w, partial);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_ssim(const SsimJobDev* jobs, int njobs, int nblk, int nb, const double* erp_w, double* partial,
                       uint32_t forms, hipStream_t st) {
  if (njobs <= 0 || nblk <= 0) return hipSuccess;
  if (forms & 1u) {
    const hipError_t e = ssim_launch<0>(jobs, njobs, nblk, nb, erp_w, partial, st);
    if (e != hipSuccess) return e;
  }
  if (forms & 4u) {
    const hipError_t e = ssim_launch<1>(jobs, njobs, nblk, nb, erp_w, partial, st);
    if (e != hipSuccess) return e;
  }
  if (forms & 2u) return ssim_launch<2>(jobs, njobs, nblk, nb, erp_w, partial, st);
  return hipSuccess;
}

hipError_t launch_ssim_finalize(const SsimJobDev* jobs, int njobs, const double* partial, const double* erp_w,
                                double* out, hipStream_t st) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ssim_finalize, dim3(njobs * 3), dim3(256), 0, st, jobs, partial, erp_w, out);
  return hipGetLastError();
}

hipError_t launch_downsample(const DownJob* jobs, int njobs, int nblk, hipStream_t st) {
  if (njobs <= 0 || nblk <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_downsample, dim3(nblk), dim3(256), 0, st, jobs, njobs);
  return hipGetLastError();
}

hipError_t launch_downsample_one(const PicDev& src, int w, int h, uint16_t* dst, int dst_pitch, hipStream_t st) {
  const int dw = w >> 1, dh = h >> 1;
  dim3 grid((dw + 63) / 64, (dh + 3) / 4, 3);
  if (src.bps == 2) hipLaunchKernelGGL(k_downsample_one<uint16_t>, grid, dim3(256), 0, st, src, dw, dh, dst, dst_pitch);
  else hipLaunchKernelGGL(k_downsample_one<uint8_t>, grid, dim3(256), 0, st, src, dw, dh, dst, dst_pitch);
  return hipGetLastError();
}

hipError_t launch_upsample420(const PicDev& src, int w, int h, uint16_t* dst, int dst_stride, hipStream_t st) {
  dim3 grid((w + 63) / 64, (h + 3) / 4, 3);
  if (src.bps == 2) hipLaunchKernelGGL(k_upsample420<uint16_t>, grid, dim3(256), 0, st, src, w, h, dst, dst_stride);
  else hipLaunchKernelGGL(k_upsample420<uint8_t>, grid, dim3(256), 0, st, src, w, h, dst, dst_stride);
  return hipGetLastError();
}

}  // namespace ivqm
