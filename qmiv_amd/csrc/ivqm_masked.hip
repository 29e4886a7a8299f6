// gfx950 kernel of the masked corresponding-pixel search (mask mode of the
// QMIV app, apps/QMIV/src/xAppQMIV.cpp:1026-1029).
//
// libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp:151-215
// (CalcDistAsymmetricRowM / FindBestPixelWithinBlockM): a test pixel whose
// mask weight is 0 adds nothing; candidates whose mask weight is 0 are not
// considered (the mask outside the picture is its Nearest-extended margin,
// xAppQMIV.cpp:899, so coordinates are clamped); the winner is the first
// strict minimum of the int32 (wrapping) weighted error in dy-outer raster
// order; the row sum is uint64 of (uint64)(int32 square) x weight per
// component.
//
// Block = 4 waves on one 64 x 16 pixel tile of one direction. The reference
// region (64 + 2R) x (16 + 2R) -- clamped coordinates, the Nearest-extended
// picture and mask -- is staged once in LDS as 8-byte entries {Y | Cb << 16,
// Cr | M << 16}. A lane owns one pixel column; a wave walks 4 rows, scoring
// every candidate from LDS with the mask test folded into a select (a masked
// candidate scores INT_MAX, which the strict < never takes, so it is not
// considered). Candidate (0, 0) carries the pixel's own non-zero weight and is
// always considered. Row sums: a DPP wave reduction per row (no LDS), one
// uint64 atomic per component.
#include "ivqm_device.h"
#include "ivqm_device_util.h"

namespace ivqm {

namespace {

typedef unsigned int u32x2m_t __attribute__((ext_vector_type(2)));

template <int R> struct MaskTile {
  static constexpr int D = 2 * R + 1, LW = kMaskTileW + 2 * R, LH = kMaskTileH + 2 * R;
  static constexpr int KB = D * D <= 16 ? 4 : D * D <= 32 ? 5 : D * D <= 64 ? 6 : 7;  // S = 2^KB > D^2
};
static_assert(kMaskTileH == 16 && kMaskTileW == 64, "4 waves x 4 rows, one column per lane");

template <typename T, typename TM, int R>
__device__ void masked_tile(const PairDev& P, int dir, int x0, int y0, u32x2m_t (*lds)[MaskTile<R>::LW], int4 cw,
                            const int32_t* __restrict__ gcd, uint64_t* __restrict__ rows64) {
  using MT = MaskTile<R>;
  constexpr int D = MT::D, LW = MT::LW, LH = MT::LH;
  const int W = P.w, H = P.h;
  const PicDev& Tp = dir ? P.p1 : P.p0;  // direction 0: P0 + GCD searched in P1; 1: P1 - GCD in P0
  const PicDev& Rp = dir ? P.p0 : P.p1;
  const TM* msk = (const TM*)P.msk;
  const int ms = P.msk_stride;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int x = x0 + lane;
  const int sg = dir ? -1 : 1;
  const int g0 = sg * gcd[4 * P.res_idx], g1 = sg * gcd[4 * P.res_idx + 1], g2 = sg * gcd[4 * P.res_idx + 2];
  // The test samples t = s + sg * GCD of the wave's 4 rows, loaded before the
  // region so both sets of loads are in flight together.
  uint32_t ts[4][3];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int y = y0 + wv * 4 + r;
    ts[r][0] = ts[r][1] = ts[r][2] = 0;
    if (x < W && y < H) {
      ts[r][0] = (uint32_t)(ldpix<T>(Tp, 0, x, y) + g0);
      ts[r][1] = (uint32_t)(ldpix<T>(Tp, 1, x, y) + g1);
      ts[r][2] = (uint32_t)(ldpix<T>(Tp, 2, x, y) + g2);
    }
  }
  for (int e = (int)threadIdx.x; e < LW * LH; e += 256) {
    const int ly = e / LW, lx = e - ly * LW;
    const int rx = clampi(x0 - R + lx, 0, W - 1), ry = clampi(y0 - R + ly, 0, H - 1);
    const uint32_t Y = (uint32_t)ldpix<T>(Rp, 0, rx, ry), Cb = (uint32_t)ldpix<T>(Rp, 1, rx, ry);
    const uint32_t Cr = (uint32_t)ldpix<T>(Rp, 2, rx, ry), M = (uint32_t)msk[(size_t)ry * ms + rx];
    lds[ly][lx] = u32x2m_t{Y | (Cb << 16), Cr | (M << 16)};
  }
  __syncthreads();
  const uint32_t w0 = (uint32_t)cw.x, w1 = (uint32_t)cw.y, w2 = (uint32_t)cw.z;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int iy = wv * 4 + r, y = y0 + iy;  // wave-uniform
    if (y >= H) break;
    uint64_t v[3] = {0, 0, 0};
    const uint32_t m = lds[iy + R][lane + R].y >> 16;
    if (x < W && m != 0) {
      const uint32_t t0 = ts[r][0], t1 = ts[r][1], t2 = ts[r][2];
      int32_t best = INT_MAX;
      int bi = R * D + R;  // (dy, dx) = (0, 0)
#pragma unroll
      for (int dy = 0; dy < D; dy++)
#pragma unroll
        for (int dx = 0; dx < D; dx++) {
          const u32x2m_t c = lds[iy + dy][lane + dx];
          const uint32_t d0 = t0 - (c.x & 0xffffu), d1 = t1 - (c.x >> 16), d2 = t2 - (c.y & 0xffffu);
          int32_t e = (int32_t)(d0 * d0 * w0 + d1 * d1 * w1 + d2 * d2 * w2);
          e = c.y >= 0x10000u ? e : INT_MAX;  // weight (high half) 0: not a candidate
          if (e < best) {
            best = e;
            bi = dy * D + dx;
          }
        }
      const int wdy = bi / D, wdx = bi - wdy * D;
      const u32x2m_t c = lds[iy + wdy][lane + wdx];
      const uint32_t d0 = t0 - (c.x & 0xffffu), d1 = t1 - (c.x >> 16), d2 = t2 - (c.y & 0xffffu);
      v[0] = (uint64_t)(int64_t)(int32_t)(d0 * d0) * (uint64_t)m;
      v[1] = (uint64_t)(int64_t)(int32_t)(d1 * d1) * (uint64_t)m;
      v[2] = (uint64_t)(int64_t)(int32_t)(d2 * d2) * (uint64_t)m;
    }
    const uint64_t s0 = wave_sum_u64_dpp(v[0]), s1 = wave_sum_u64_dpp(v[1]), s2 = wave_sum_u64_dpp(v[2]);
    if (lane < 3) {
      const uint64_t s = lane == 0 ? s0 : (lane == 1 ? s1 : s2);
      atomicAdd((unsigned long long*)&rows64[P.rows64_off + (int64_t)(dir * 3 + lane) * H + y],
                (unsigned long long)s);
    }
  }
}

// Keyed form (host-checked ranges, PairDev::mkey != 0): the unmasked search's
// narrow key S * (Q(r) - 2 sum_c w_c t_c r_c) + offset with the tile-position
// offset ly * D + lx in the entry, so a plain integer min is the first strict
// minimum in raster order (nothing wraps in range, so argmin of the reference's
// E = T + E' is argmin E'). A masked-out candidate's entry holds mkey, which
// exceeds every considered key. Entry {S * Q + off | mkey, Y, Cb, Cr}: three
// 24-bit multiply-adds and a min per candidate.
template <typename T, typename TM, int R>
__device__ void masked_tile_keyed(const PairDev& P, int dir, int x0, int y0, int4 (*lds)[MaskTile<R>::LW], int4 cw,
                                  const int32_t* __restrict__ gcd, uint64_t* __restrict__ rows64) {
  using MT = MaskTile<R>;
  constexpr int D = MT::D, LW = MT::LW, LH = MT::LH, S = 1 << MT::KB;
  const int W = P.w, H = P.h;
  const PicDev& Tp = dir ? P.p1 : P.p0;
  const PicDev& Rp = dir ? P.p0 : P.p1;
  const TM* msk = (const TM*)P.msk;
  const int ms = P.msk_stride;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int x = x0 + lane;
  const int sg = dir ? -1 : 1;
  const int g0 = sg * gcd[4 * P.res_idx], g1 = sg * gcd[4 * P.res_idx + 1], g2 = sg * gcd[4 * P.res_idx + 2];
  int ts[4][3];
  uint32_t mw[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int y = y0 + wv * 4 + r;
    ts[r][0] = ts[r][1] = ts[r][2] = 0;
    mw[r] = 0;
    if (x < W && y < H) {
      ts[r][0] = ldpix<T>(Tp, 0, x, y) + g0;
      ts[r][1] = ldpix<T>(Tp, 1, x, y) + g1;
      ts[r][2] = ldpix<T>(Tp, 2, x, y) + g2;
      mw[r] = (uint32_t)msk[(size_t)y * ms + x];
    }
  }
  const uint32_t uY = (uint32_t)cw.x, uCb = (uint32_t)cw.y, uCr = (uint32_t)cw.z;
  for (int e = (int)threadIdx.x; e < LW * LH; e += 256) {
    const int ly = e / LW, lx = e - ly * LW;
    const int rx = clampi(x0 - R + lx, 0, W - 1), ry = clampi(y0 - R + ly, 0, H - 1);
    const uint32_t Y = (uint32_t)ldpix<T>(Rp, 0, rx, ry), Cb = (uint32_t)ldpix<T>(Rp, 1, rx, ry);
    const uint32_t Cr = (uint32_t)ldpix<T>(Rp, 2, rx, ry), M = (uint32_t)msk[(size_t)ry * ms + rx];
    // Q(r) = sum w r^2 exactly: w * r < 2^24 and S * Q + off < 2^31 (host-checked)
    const uint32_t q = __umul24(__umul24(uY, Y), Y) + __umul24(__umul24(uCb, Cb), Cb) + __umul24(__umul24(uCr, Cr), Cr);
    const int kb = M ? (int)(q * (uint32_t)S + (uint32_t)(ly * D + lx)) : P.mkey;
    lds[ly][lx] = int4{kb, (int)Y, (int)Cb, (int)Cr};
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int iy = wv * 4 + r, y = y0 + iy;  // wave-uniform
    if (y >= H) break;
    uint64_t v[3] = {0, 0, 0};
    const uint32_t m = mw[r];
    if (m != 0) {  // zero for x >= W
      const int t0 = ts[r][0], t1 = ts[r][1], t2 = ts[r][2];
      // multipliers -2 S w t (|value| < 2^23, host-checked)
      const int kY = -2 * S * cw.x * t0, kCb = -2 * S * cw.y * t1, kCr = -2 * S * cw.z * t2;
      int bk = INT_MAX;
      constexpr int UY = R <= 2 ? D : 1;  // larger windows: one candidate row at a time (registers)
#pragma unroll UY
      for (int dy = 0; dy < D; dy++)
#pragma unroll
        for (int dx = 0; dx < D; dx++) {
          const int4 c = lds[iy + dy][lane + dx];
          int acc = __mul24(kY, c.y) + c.x;
          acc = __mul24(kCb, c.z) + acc;
          acc = __mul24(kCr, c.w) + acc;
          bk = min(bk, acc);
        }
      const uint32_t idx = (uint32_t)(bk - (iy * D + lane)) & (uint32_t)(S - 1);
      const uint32_t wdy = idx / (uint32_t)D, wdx = idx - wdy * (uint32_t)D;
      const int4 c = lds[iy + wdy][lane + wdx];
      const int d0 = t0 - c.y, d1 = t1 - c.z, d2 = t2 - c.w;  // |d| < 2^15: squares are positive int32
      v[0] = (uint64_t)(int64_t)(d0 * d0) * (uint64_t)m;
      v[1] = (uint64_t)(int64_t)(d1 * d1) * (uint64_t)m;
      v[2] = (uint64_t)(int64_t)(d2 * d2) * (uint64_t)m;
    }
    const uint64_t s0 = wave_sum_u64_dpp(v[0]), s1 = wave_sum_u64_dpp(v[1]), s2 = wave_sum_u64_dpp(v[2]);
    if (lane < 3) {
      const uint64_t s = lane == 0 ? s0 : (lane == 1 ? s1 : s2);
      atomicAdd((unsigned long long*)&rows64[P.rows64_off + (int64_t)(dir * 3 + lane) * H + y],
                (unsigned long long)s);
    }
  }
}

template <int R>
__global__ __launch_bounds__(256) void k_search_masked(const PairDev* __restrict__ pairs, int npairs, int ntiles,
                                                       int4 cw, const int32_t* __restrict__ gcd,
                                                       uint64_t* __restrict__ rows64) {
  using MT = MaskTile<R>;
  __shared__ int4 lds4[MT::LH][MT::LW];  // keyed entries; the reference form uses its first half as 8-byte entries
  auto lds = reinterpret_cast<u32x2m_t (*)[MT::LW]>(&lds4[0][0]);
  const int b = blockIdx.x;
  if (b >= ntiles) return;
  const int pi = find_entry(pairs, npairs, b, [](const PairDev& p) { return p.mtile0; });
  const PairDev& P = pairs[pi];
  const int nx = (P.w + kMaskTileW - 1) / kMaskTileW, ny = (P.h + kMaskTileH - 1) / kMaskTileH;
  const int lt = b - P.mtile0;
  const int dir = lt / (nx * ny), rem = lt - dir * nx * ny;
  const int ty = rem / nx, tx = rem - ty * nx;
  const int x0 = tx * kMaskTileW, y0 = ty * kMaskTileH;
  if (P.mkey != 0) {
    if (P.p0.bps == 2) {
      if (P.msk_bps == 2) masked_tile_keyed<uint16_t, uint16_t, R>(P, dir, x0, y0, lds4, cw, gcd, rows64);
      else masked_tile_keyed<uint16_t, uint8_t, R>(P, dir, x0, y0, lds4, cw, gcd, rows64);
    } else {
      if (P.msk_bps == 2) masked_tile_keyed<uint8_t, uint16_t, R>(P, dir, x0, y0, lds4, cw, gcd, rows64);
      else masked_tile_keyed<uint8_t, uint8_t, R>(P, dir, x0, y0, lds4, cw, gcd, rows64);
    }
  } else if (P.p0.bps == 2) {
    if (P.msk_bps == 2) masked_tile<uint16_t, uint16_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64);
    else masked_tile<uint16_t, uint8_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64);
  } else {
    if (P.msk_bps == 2) masked_tile<uint8_t, uint16_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64);
    else masked_tile<uint8_t, uint8_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64);
  }
}

}  // namespace

hipError_t launch_search_masked(const PairDev* pairs, int npairs, int ntiles, int R, const int32_t cw[4],
                                const int32_t* gcd, uint64_t* rows64, hipStream_t st) {
  if (npairs <= 0 || ntiles <= 0) return hipSuccess;
  const int4 w = make_int4(cw[0], cw[1], cw[2], cw[3]);
  switch (R) {
    case 1: hipLaunchKernelGGL(k_search_masked<1>, dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64); break;
    case 2: hipLaunchKernelGGL(k_search_masked<2>, dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64); break;
    case 3: hipLaunchKernelGGL(k_search_masked<3>, dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64); break;
    case 4: hipLaunchKernelGGL(k_search_masked<4>, dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace ivqm
