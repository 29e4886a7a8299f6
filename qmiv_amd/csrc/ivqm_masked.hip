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
// One thread per pixel, one wave per 64 pixels of a row, waves of both
// directions in one grid (tiles of 64 x 4 pixels per direction). The
// candidate window is read through the caches (no LDS staging): mask mode is
// a correctness path, not the headline.
#include "ivqm_device.h"
#include "ivqm_device_util.h"

namespace ivqm {

namespace {

template <typename T, typename TM>
__device__ void masked_search_px(const PairDev& P, int dir, int x, int y, int R, int4 cw, const int32_t* gcd,
                                 uint64_t v[3]) {
  v[0] = v[1] = v[2] = 0;
  const int W = P.w, H = P.h;
  if (x >= W) return;
  const TM* msk = (const TM*)P.msk;
  const int ms = P.msk_stride;
  const uint32_t m = msk[(size_t)y * ms + x];
  if (m == 0) return;
  const PicDev& Tp = dir ? P.p1 : P.p0;  // direction 0: P0 + GCD searched in P1; 1: P1 - GCD in P0
  const PicDev& Rp = dir ? P.p0 : P.p1;
  const int sg = dir ? -1 : 1;
  int32_t t[3];
#pragma unroll
  for (int c = 0; c < 3; c++) t[c] = ldpix<T>(Tp, c, x, y) + sg * gcd[4 * P.res_idx + c];
  int32_t best = INT_MAX;
  int bx = x, by = y;
  for (int dy = -R; dy <= R; dy++) {
    const int ry = clampi(y + dy, 0, H - 1);
    for (int dx = -R; dx <= R; dx++) {
      const int rx = clampi(x + dx, 0, W - 1);
      if (msk[(size_t)ry * ms + rx] == 0) continue;
      const uint32_t d0 = (uint32_t)(t[0] - ldpix<T>(Rp, 0, rx, ry));
      const uint32_t d1 = (uint32_t)(t[1] - ldpix<T>(Rp, 1, rx, ry));
      const uint32_t d2 = (uint32_t)(t[2] - ldpix<T>(Rp, 2, rx, ry));
      const int32_t e = (int32_t)(d0 * d0 * (uint32_t)cw.x + d1 * d1 * (uint32_t)cw.y + d2 * d2 * (uint32_t)cw.z);
      if (e < best) {
        best = e;
        bx = rx;
        by = ry;
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const uint32_t d = (uint32_t)(t[c] - ldpix<T>(Rp, c, bx, by));
    v[c] = (uint64_t)(int64_t)(int32_t)(d * d) * (uint64_t)m;
  }
}

__global__ __launch_bounds__(256) void k_search_masked(const PairDev* __restrict__ pairs, int npairs, int ntiles,
                                                       int R, int4 cw, const int32_t* __restrict__ gcd,
                                                       uint64_t* __restrict__ rows64) {
  const int b = blockIdx.x;
  if (b >= ntiles) return;
  const int pi = find_entry(pairs, npairs, b, [](const PairDev& p) { return p.mtile0; });
  const PairDev& P = pairs[pi];
  const int nx = (P.w + kMaskTileW - 1) / kMaskTileW, ny = (P.h + kMaskTileH - 1) / kMaskTileH;
  const int lt = b - P.mtile0;
  const int dir = lt / (nx * ny), rem = lt - dir * nx * ny;
  const int ty = rem / nx, tx = rem - ty * nx;
  const int lane = threadIdx.x & 63;
  const int y = ty * kMaskTileH + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  if (y >= P.h) return;  // whole wave
  const int x = tx * kMaskTileW + lane;
  uint64_t v[3];
  if (P.p0.bps == 2) {
    if (P.msk_bps == 2) masked_search_px<uint16_t, uint16_t>(P, dir, x, y, R, cw, gcd, v);
    else masked_search_px<uint16_t, uint8_t>(P, dir, x, y, R, cw, gcd, v);
  } else {
    if (P.msk_bps == 2) masked_search_px<uint8_t, uint16_t>(P, dir, x, y, R, cw, gcd, v);
    else masked_search_px<uint8_t, uint8_t>(P, dir, x, y, R, cw, gcd, v);
  }
#pragma unroll
  for (int k = 32; k >= 1; k >>= 1) {
    v[0] += __shfl_xor(v[0], k, 64);
    v[1] += __shfl_xor(v[1], k, 64);
    v[2] += __shfl_xor(v[2], k, 64);
  }
  if (lane < 3) {
    const uint64_t s = lane == 0 ? v[0] : (lane == 1 ? v[1] : v[2]);
    atomicAdd((unsigned long long*)&rows64[P.rows64_off + (int64_t)(dir * 3 + lane) * P.h + y],
              (unsigned long long)s);
  }
}

}  // namespace

hipError_t launch_search_masked(const PairDev* pairs, int npairs, int ntiles, int R, const int32_t cw[4],
                                const int32_t* gcd, uint64_t* rows64, hipStream_t st) {
  if (npairs <= 0 || ntiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_search_masked, dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, R,
                     make_int4(cw[0], cw[1], cw[2], cw[3]), gcd, rows64);
  return hipGetLastError();
}

}  // namespace ivqm
