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
//
// Shift-compensated pictures with a mask (PairDev::scp_off >= 0): the winner's
// reference sample minus the direction's colour shift, clipped to [0, Max]
// (xCorrespPixelShiftSTD.cpp:268-291, GenShftCompRowM), at every pixel whose
// weight is non-zero; the reference leaves masked pixels unwritten, here they
// hold 0. Direction 0 writes SCP1, direction 1 SCP0, as the unmasked search.
#include "ivqm_device.h"
#include "ivqm_device_util.h"

namespace ivqm {

namespace {

typedef unsigned int u32x2m_t __attribute__((ext_vector_type(2)));
typedef short s16x2m_t __attribute__((ext_vector_type(2)));

template <int R> struct MaskTile {
  static constexpr int D = 2 * R + 1, LW = kMaskTileW + 2 * R, LH = kMaskTileH + 2 * R;
  static constexpr int LWP = (LW + 3) & ~3;  // keyed planes' row pitch: 16-byte aligned rows
  static constexpr int KB = D * D <= 16 ? 4 : D * D <= 32 ? 5 : D * D <= 64 ? 6 : 7;  // S = 2^KB > D^2
};
static_assert(kMaskTileH == 16 && kMaskTileW == 64, "4 waves x 4 rows, one column per lane");

// One SCP sample triple of pixel (x, y): clip(r - g, 0, Max) per component, 0 for a masked pixel.
__device__ __forceinline__ void masked_scp_store(const PairDev& P, int dir, int x, int y, bool keep, const int (&r)[3],
                                                 const int (&g)[3], uint16_t* __restrict__ scp) {
  const size_t plane = (size_t)P.scp_pitch * P.h;
  uint16_t* pic = scp + P.scp_off + (size_t)(1 - dir) * 3 * plane + (size_t)y * P.scp_pitch + x;
#pragma unroll
  for (int c = 0; c < 3; c++) pic[c * plane] = keep ? (uint16_t)clampi(r[c] - g[c], 0, P.maxv) : (uint16_t)0;
}

template <typename T, typename TM, int R>
__device__ void masked_tile(const PairDev& P, int dir, int x0, int y0, u32x2m_t (*lds)[MaskTile<R>::LW], int4 cw,
                            const int32_t* __restrict__ gcd, uint64_t* __restrict__ rows64,
                            uint16_t* __restrict__ scp) {
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
      if (P.scp_off >= 0) {
        const int r3[3] = {(int)(c.x & 0xffffu), (int)(c.x >> 16), (int)(c.y & 0xffffu)};
        const int g3[3] = {g0, g1, g2};
        masked_scp_store(P, dir, x, y, true, r3, g3, scp);
      }
    } else if (x < W && P.scp_off >= 0) {
      const int z3[3] = {0, 0, 0};
      masked_scp_store(P, dir, x, y, false, z3, z3, scp);
    }
    const uint64_t s0 = wave_sum_u64_dpp(v[0]), s1 = wave_sum_u64_dpp(v[1]), s2 = wave_sum_u64_dpp(v[2]);
    if (lane < 3) {
      const uint64_t s = lane == 0 ? s0 : (lane == 1 ? s1 : s2);
      atomicAdd((unsigned long long*)&rows64[P.rows64_off + (int64_t)(dir * 3 + lane) * H + y],
                (unsigned long long)s);
    }
  }
}

// NE consecutive ints from a 16-byte aligned LDS address: 16-byte reads plus an 8-byte tail.
template <int NE>
__device__ __forceinline__ void lds_ints(const int* p, int (&v)[NE]) {
  typedef int i32x4m_t __attribute__((ext_vector_type(4)));
  typedef int i32x2m_t __attribute__((ext_vector_type(2)));
  static_assert(NE % 2 == 0, "run length");
  const i32x4m_t* q = reinterpret_cast<const i32x4m_t*>(p);
#pragma unroll
  for (int i = 0; i < NE / 4; i++) {
    const i32x4m_t w = q[i];
    v[4 * i] = w.x;
    v[4 * i + 1] = w.y;
    v[4 * i + 2] = w.z;
    v[4 * i + 3] = w.w;
  }
  if (NE % 4) {
    const i32x2m_t w = *reinterpret_cast<const i32x2m_t*>(q + NE / 4);
    v[NE - 2] = w.x;
    v[NE - 1] = w.y;
  }
}

// Inclusive 64-bit scan within each 16-lane row (row_shr 1, 2, 4, 8; lanes
// shifted in from outside the row read 0): lane 15 of a row holds its total.
__device__ __forceinline__ uint64_t row16_sum_u64_dpp(uint64_t v) {
#define IVQM_ROWQ(ctrl)                                                                                 \
  {                                                                                                    \
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)v, ctrl, 0xf, 0xf, true);  \
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(v >> 32), ctrl, 0xf, 0xf, true); \
    v += (uint64_t)lo | ((uint64_t)hi << 32);                                                          \
  }
  IVQM_ROWQ(0x111)
  IVQM_ROWQ(0x112)
  IVQM_ROWQ(0x114)
  IVQM_ROWQ(0x118)
#undef IVQM_ROWQ
  return v;
}

// Keyed form (host-checked ranges, PairDev::mkey != 0): the unmasked search's
// narrow key S * (Q(r) - 2 sum_c w_c t_c r_c) + offset with the tile-position
// offset ly * D + lx in the entry, so a plain integer min is the first strict
// minimum in raster order (nothing wraps in range, so argmin of the reference's
// E = T + E' is argmin E'). A masked-out candidate's entry holds mkey, which
// exceeds every considered key. As in k_search, a lane owns 4 adjacent pixels
// of one row (16 lanes per row, a wave 4 rows): per candidate row it reads the
// 4 + 2R entries its pixels share once, from four int planes (key base, Y, Cb,
// Cr; 16-byte reads of lanes 16 bytes apart, no bank conflicts), and scores
// 4 x (2R + 1) candidates with three 24-bit multiply-adds and a min each.
// DOT (PairDev::mdot): the Y and Cb planes become one plane of int16 pairs {Y * S, Cb * S} scored by one
// v_dot2 against the pixel's {-2 wY t_Y, -2 wCb t_Cb}: one dot, one 24-bit multiply-add (Cr) and a min per
// candidate, three planes.
template <typename T, typename TM, int R, bool DOT>
__device__ void masked_tile_keyed(const PairDev& P, int dir, int x0, int y0, int* lds, int4 cw,
                                  const int32_t* __restrict__ gcd, uint64_t* __restrict__ rows64,
                                  uint16_t* __restrict__ scp) {
  using MT = MaskTile<R>;
  constexpr int D = MT::D, LW = MT::LW, LH = MT::LH, LWP = MT::LWP, S = 1 << MT::KB, NE = 4 + 2 * R;
  int (*lk)[LWP] = reinterpret_cast<int (*)[LWP]>(lds);
  int (*ly_)[LWP] = reinterpret_cast<int (*)[LWP]>(lds + LH * LWP);
  int (*lb)[LWP] = reinterpret_cast<int (*)[LWP]>(lds + 2 * LH * LWP);  // unused by DOT
  int (*lr)[LWP] = reinterpret_cast<int (*)[LWP]>(lds + (DOT ? 2 : 3) * LH * LWP);
  const int W = P.w, H = P.h;
  const PicDev& Tp = dir ? P.p1 : P.p0;
  const PicDev& Rp = dir ? P.p0 : P.p1;
  const TM* msk = (const TM*)P.msk;
  const int ms = P.msk_stride;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int iy = wv * 4 + (lane >> 4), ix = 4 * (lane & 15);
  const int y = y0 + iy, x = x0 + ix;
  const int sg = dir ? -1 : 1;
  const int g0 = sg * gcd[4 * P.res_idx], g1 = sg * gcd[4 * P.res_idx + 1], g2 = sg * gcd[4 * P.res_idx + 2];
  // Loads through buffer resources over the planes and the mask (row byte offsets scalar where the row is
  // wave-uniform, 32-bit vector offsets otherwise; no 64-bit address arithmetic).
  const int csx = Rp.cs, csy = Rp.csy;
  const __amdgpu_buffer_rsrc_t rY = plane_rsrc<T>(Rp, 0, H), rCb = plane_rsrc<T>(Rp, 1, H), rCr = plane_rsrc<T>(Rp, 2, H);
  const uint32_t sY = (uint32_t)plane_stride(Rp, 0) * (uint32_t)sizeof(T), sC = (uint32_t)plane_stride(Rp, 1) * (uint32_t)sizeof(T);
  const __amdgpu_buffer_rsrc_t rM =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<TM*>(msk), (short)0, H * ms * (int)sizeof(TM), 0x00020000);
  const uint32_t sM = (uint32_t)ms * (uint32_t)sizeof(TM);
  // test samples t = s + sg * GCD and weights of the lane's pixels (weight 0 outside the picture),
  // loaded before the region so both sets of loads are in flight together
  int t[4][3];
  uint32_t m[4];
  {
    const __amdgpu_buffer_rsrc_t tY = plane_rsrc<T>(Tp, 0, H), tCb = plane_rsrc<T>(Tp, 1, H), tCr = plane_rsrc<T>(Tp, 2, H);
    const uint32_t yc = (uint32_t)min(y, H - 1);  // the lane's row (per lane: vector offsets)
    const uint32_t oY = yc * (uint32_t)plane_stride(Tp, 0) * (uint32_t)sizeof(T);
    const uint32_t oC = (yc >> Tp.csy) * (uint32_t)plane_stride(Tp, 1) * (uint32_t)sizeof(T);
    const uint32_t oM = yc * sM;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const uint32_t xk = (uint32_t)min(x + k, W - 1);
      const uint32_t xc = (xk >> Tp.cs) * (uint32_t)sizeof(T);
      t[k][0] = ld_buf<T>(tY, xk * (uint32_t)sizeof(T) + oY, 0) + g0;
      t[k][1] = ld_buf<T>(tCb, xc + oC, 0) + g1;
      t[k][2] = ld_buf<T>(tCr, xc + oC, 0) + g2;
      m[k] = (uint32_t)ld_buf<TM>(rM, xk * (uint32_t)sizeof(TM) + oM, 0);
      if (!(y < H && x + k < W)) m[k] = 0;
    }
  }
  // Region entries: lane = region column (x0 - R + lane, clamped), wave wv stages rows wv, wv + 4, ...; the 2R
  // halo columns 64 .. 64 + 2R - 1 are staged by the first 2R * LH threads.
  const uint32_t uY = (uint32_t)cw.x, uCb = (uint32_t)cw.y, uCr = (uint32_t)cw.z;
  auto put = [&](int ry_, int rx_, uint32_t Y, uint32_t Cb, uint32_t Cr, uint32_t M) {
    // Q(r) = sum w r^2 exactly: w * r < 2^24 and S * Q + off < 2^31 (host-checked)
    const uint32_t q = __umul24(__umul24(uY, Y), Y) + __umul24(__umul24(uCb, Cb), Cb) + __umul24(__umul24(uCr, Cr), Cr);
    lk[ry_][rx_] = M ? (int)(q * (uint32_t)S + (uint32_t)(ry_ * D + rx_)) : P.mkey;
    if (DOT) {
      ly_[ry_][rx_] = (int)((Y * (uint32_t)S) | ((Cb * (uint32_t)S) << 16));
    } else {
      ly_[ry_][rx_] = (int)Y;
      lb[ry_][rx_] = (int)Cb;
    }
    lr[ry_][rx_] = (int)Cr;
  };
  {
    constexpr int NRW = (LH + 3) / 4;
    const uint32_t cx = (uint32_t)clampi(x0 - R + lane, 0, W - 1);
    const uint32_t oYx = cx * (uint32_t)sizeof(T), oCx = (cx >> csx) * (uint32_t)sizeof(T), oMx = cx * (uint32_t)sizeof(TM);
    uint32_t ra[NRW][4];
#pragma unroll
    for (int k = 0; k < NRW; k++) {
      const int ry_ = wv + 4 * k;
      if (ry_ < LH) {
        const uint32_t ry = (uint32_t)clampi(y0 - R + ry_, 0, H - 1);  // wave-uniform
        ra[k][0] = (uint32_t)ld_buf<T>(rY, oYx, ry * sY);
        ra[k][1] = (uint32_t)ld_buf<T>(rCb, oCx, (ry >> csy) * sC);
        ra[k][2] = (uint32_t)ld_buf<T>(rCr, oCx, (ry >> csy) * sC);
        ra[k][3] = (uint32_t)ld_buf<TM>(rM, oMx, ry * sM);
      }
    }
#pragma unroll
    for (int k = 0; k < NRW; k++) {
      const int ry_ = wv + 4 * k;
      if (ry_ < LH) put(ry_, lane, ra[k][0], ra[k][1], ra[k][2], ra[k][3]);
    }
    static_assert(2 * R * LH <= 256, "one pass over the halo entries");
    const int e = (int)threadIdx.x;
    if (e < 2 * R * LH) {
      const int ry_ = e / (2 * R), rx_ = 64 + (e - ry_ * (2 * R));
      const uint32_t hx = (uint32_t)clampi(x0 - R + rx_, 0, W - 1), hy = (uint32_t)clampi(y0 - R + ry_, 0, H - 1);
      const uint32_t Y = (uint32_t)ld_buf<T>(rY, hx * (uint32_t)sizeof(T) + hy * sY, 0);
      const uint32_t oc = (hx >> csx) * (uint32_t)sizeof(T) + (hy >> csy) * sC;
      const uint32_t Cb = (uint32_t)ld_buf<T>(rCb, oc, 0), Cr = (uint32_t)ld_buf<T>(rCr, oc, 0);
      const uint32_t M = (uint32_t)ld_buf<TM>(rM, hx * (uint32_t)sizeof(TM) + hy * sM, 0);
      put(ry_, rx_, Y, Cb, Cr, M);
    }
  }
  __syncthreads();
  // multipliers -2 S w t (|value| < 2^23, host-checked)
  int kY[4], kCb[4], kCr[4], bk[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (DOT) {  // {-2 wY t_Y, -2 wCb t_Cb} as an int16 pair (|value| <= 32767, host-checked)
      kY[k] = (int)(((uint32_t)(-2 * cw.x * t[k][0]) & 0xffffu) | ((uint32_t)(-2 * cw.y * t[k][1]) << 16));
      kCb[k] = 0;
    } else {
      kY[k] = -2 * S * cw.x * t[k][0];
      kCb[k] = -2 * S * cw.y * t[k][1];
    }
    kCr[k] = -2 * S * cw.z * t[k][2];
    bk[k] = INT_MAX;
  }
#pragma unroll 1
  for (int dy = 0; dy < D; dy++) {  // one candidate row at a time (registers)
    int ek[NE], ey[NE], eb[NE], er[NE];
    lds_ints<NE>(&lk[iy + dy][ix], ek);
    lds_ints<NE>(&ly_[iy + dy][ix], ey);
    if (!DOT) lds_ints<NE>(&lb[iy + dy][ix], eb);
    lds_ints<NE>(&lr[iy + dy][ix], er);
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
      for (int dx = 0; dx < D; dx++) {
        int acc = __mul24(kCr[k], er[k + dx]) + ek[k + dx];
        if (DOT) {
          acc = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2m_t, kY[k]), __builtin_bit_cast(s16x2m_t, ey[k + dx]),
                                       acc, false);
        } else {
          acc = __mul24(kY[k], ey[k + dx]) + acc;
          acc = __mul24(kCb[k], eb[k + dx]) + acc;
        }
        bk[k] = min(bk[k], acc);
      }
  }
  uint64_t v[3] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if (m[k] != 0) {
      const uint32_t idx = (uint32_t)(bk[k] - (iy * D + ix + k)) & (uint32_t)(S - 1);
      const uint32_t wdy = idx / (uint32_t)D, wdx = idx - wdy * (uint32_t)D;
      const int ey = iy + (int)wdy, ex = ix + k + (int)wdx;
      int rY, rCb;
      if (DOT) {
        const uint32_t pk = (uint32_t)ly_[ey][ex];
        rY = (int)((pk & 0xffffu) >> MT::KB);
        rCb = (int)((pk >> 16) >> MT::KB);
      } else {
        rY = ly_[ey][ex];
        rCb = lb[ey][ex];
      }
      const int rCr = lr[ey][ex];
      const int d0 = t[k][0] - rY, d1 = t[k][1] - rCb, d2 = t[k][2] - rCr;
      // |d| < 2^15: the squares are positive int32
      v[0] += (uint64_t)(int64_t)(d0 * d0) * (uint64_t)m[k];
      v[1] += (uint64_t)(int64_t)(d1 * d1) * (uint64_t)m[k];
      v[2] += (uint64_t)(int64_t)(d2 * d2) * (uint64_t)m[k];
      if (P.scp_off >= 0) {
        const int r3[3] = {rY, rCb, rCr};
        const int g3[3] = {g0, g1, g2};
        masked_scp_store(P, dir, x + k, y, true, r3, g3, scp);
      }
    } else if (P.scp_off >= 0 && y < H && x + k < W) {
      const int z3[3] = {0, 0, 0};
      masked_scp_store(P, dir, x + k, y, false, z3, z3, scp);
    }
  }
  const uint64_t s0 = row16_sum_u64_dpp(v[0]), s1 = row16_sum_u64_dpp(v[1]), s2 = row16_sum_u64_dpp(v[2]);
  if ((lane & 15) == 15 && y < H) {
    uint64_t* rp = rows64 + P.rows64_off + (int64_t)dir * 3 * H + y;
    atomicAdd((unsigned long long*)rp, (unsigned long long)s0);
    atomicAdd((unsigned long long*)(rp + H), (unsigned long long)s1);
    atomicAdd((unsigned long long*)(rp + 2 * H), (unsigned long long)s2);
  }
}

// 4:2:0 keyed form (PairDev::mdot == 2; host-checked ranges as the packed form): the key of the keyed forms,
// scored as in k_search420. A chroma sample covers a 2 x 2 luma block, so the pixels (2m, 2m + 1) of two rows
// (2n, 2n + 1) share their test chroma and the chroma part of a candidate's score,
//   X = key + dot2({-2 wCb t_Cb, -2 wCr t_Cr}, {Cb * S, Cr * S}),
// computed once per candidate entry and pixel pair (v_dot2 over int16 pairs); a candidate then costs one
// 24-bit multiply-add (luma, -2 S wY t_Y times Y) and half a min3. A lane owns 2 rows x 4 columns, so each
// candidate row's LDS runs and X serve both rows. Block = 4 waves on one 64 x 32 tile of one direction.
// LDS: keys and Y per luma entry, raw {Cb, Cr} per luma entry (the epilogue's winner samples), chroma words
// {Cb * S, Cr * S} at chroma resolution.
template <int R> struct Mask420 {
  static constexpr int D = 2 * R + 1, TW = 64, TH = kMask420TileH, LW = TW + 2 * R, LH = TH + 2 * R;
  static constexpr int R1 = R & 1;
  static constexpr int LPK = (LW + 3) & ~3;          // luma plane pitch (ints): lane runs 16-byte aligned
  static constexpr int LWC = (LW - 1 + R1) / 2 + 1;  // chroma entries of a region row
  static constexpr int LHC = (LH - 1 + R1) / 2 + 1;  // chroma rows of the region
  static constexpr int LPC = (LWC + 1) & ~1;         // chroma plane pitch (ints): lane runs 8-byte aligned
  static constexpr int NE = 4 + 2 * R;               // luma entries of a lane's candidate run
  static constexpr int NC = (NE - 1 + R1) / 2 + 1;   // chroma entries of the run
  static constexpr int KB = MaskTile<R>::KB;
  static constexpr int LDS_INTS = 3 * LH * LPK + LHC * LPC;
  static_assert(TH == 32, "4 waves x 8 rows");
  static_assert(NE % 2 == 0 && NC % 2 == 0, "run lengths");
};

template <typename T, typename TM, int R>
__device__ void masked420_tile(const PairDev& P, int dir, int x0, int y0, int* lds, int4 cw,
                               const int32_t* __restrict__ gcd, uint64_t* __restrict__ rows64,
                               uint16_t* __restrict__ scp) {
  using MT = Mask420<R>;
  constexpr int D = MT::D, LH = MT::LH, LPK = MT::LPK, LPC = MT::LPC, NE = MT::NE, NC = MT::NC, R1 = MT::R1;
  constexpr int S = 1 << MT::KB, NHC = 2 * R;
  int (*lk)[LPK] = reinterpret_cast<int (*)[LPK]>(lds);
  int (*lyv)[LPK] = reinterpret_cast<int (*)[LPK]>(lds + LH * LPK);
  uint32_t (*lcu)[LPK] = reinterpret_cast<uint32_t (*)[LPK]>(lds + 2 * LH * LPK);
  int (*lcw)[LPC] = reinterpret_cast<int (*)[LPC]>(lds + 3 * LH * LPK);
  const int W = P.w, H = P.h;
  const PicDev& Tp = dir ? P.p1 : P.p0;  // direction 0: P0 + GCD searched in P1; 1: P1 - GCD in P0
  const PicDev& Rp = dir ? P.p0 : P.p1;
  const TM* msk = (const TM*)P.msk;
  const int ms = P.msk_stride;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int rp = lane >> 4;
  const int iy0 = 8 * wv + 2 * rp, ix = 4 * (lane & 15);  // the lane's rows iy0, iy0 + 1 and columns ix .. ix + 3
  __builtin_assume(iy0 >= 0 && iy0 < 32);
  const int sg = dir ? -1 : 1;
  const int g0 = sg * gcd[4 * P.res_idx], g1 = sg * gcd[4 * P.res_idx + 1], g2 = sg * gcd[4 * P.res_idx + 2];
  const __amdgpu_buffer_rsrc_t rY = plane_rsrc<T>(Rp, 0, H), rCb = plane_rsrc<T>(Rp, 1, H), rCr = plane_rsrc<T>(Rp, 2, H);
  const uint32_t sY = (uint32_t)plane_stride(Rp, 0) * (uint32_t)sizeof(T), sC = (uint32_t)plane_stride(Rp, 1) * (uint32_t)sizeof(T);
  const __amdgpu_buffer_rsrc_t rM =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<TM*>(msk), (short)0, H * ms * (int)sizeof(TM), 0x00020000);
  const uint32_t sM = (uint32_t)ms * (uint32_t)sizeof(TM);
  // ---- test samples t = s + sg * GCD of the lane's 8 pixels (chroma: one sample per pixel pair, both rows)
  // and their weights (0 outside the picture); loaded before the region so both sets of loads are in flight
  int tY[2][4], tC[2][2];
  uint32_t m[2][4];
  {
    const __amdgpu_buffer_rsrc_t qY = plane_rsrc<T>(Tp, 0, H), qCb = plane_rsrc<T>(Tp, 1, H), qCr = plane_rsrc<T>(Tp, 2, H);
    const uint32_t tsY = (uint32_t)plane_stride(Tp, 0) * (uint32_t)sizeof(T), tsC = (uint32_t)plane_stride(Tp, 1) * (uint32_t)sizeof(T);
#pragma unroll
    for (int r = 0; r < 2; r++) {
      const int y = y0 + iy0 + r;
      const uint32_t yc = (uint32_t)min(y, H - 1);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t xk = (uint32_t)min(x0 + ix + k, W - 1);
        tY[r][k] = ld_buf<T>(qY, xk * (uint32_t)sizeof(T) + yc * tsY, 0) + g0;
        m[r][k] = (y < H && x0 + ix + k < W) ? (uint32_t)ld_buf<TM>(rM, xk * (uint32_t)sizeof(TM) + yc * sM, 0) : 0u;
      }
    }
    const uint32_t ycc = ((uint32_t)min(y0 + iy0, H - 1) >> 1) * tsC;
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const uint32_t xc = ((uint32_t)min(x0 + ix + 2 * h, W - 1) >> 1) * (uint32_t)sizeof(T);
      tC[0][h] = ld_buf<T>(qCb, xc + ycc, 0) + g1;
      tC[1][h] = ld_buf<T>(qCr, xc + ycc, 0) + g2;
    }
  }
  // ---- region: rows y0 - R + ly, columns x0 - R + lx (clamped). Region rows go in pairs (2j - R1, 2j + 1 - R1)
  // of one chroma row j; wave wv stages pairs wv, wv + 4, ...; lane = region column, the 2R halo columns
  // 64 .. 64 + 2R - 1 by the first threads.
  const uint32_t uY = (uint32_t)cw.x, uCb = (uint32_t)cw.y, uCr = (uint32_t)cw.z;
  auto entry = [&](int ly, int lx, uint32_t Y, uint32_t Cb, uint32_t Cr, uint32_t M) {
    // Q(r) = sum w r^2 exactly: w * r < 2^24 and S * Q + off < 2^31 (host-checked)
    const uint32_t q = __umul24(__umul24(uY, Y), Y) + __umul24(__umul24(uCb, Cb), Cb) + __umul24(__umul24(uCr, Cr), Cr);
    lk[ly][lx] = M ? (int)(q * (uint32_t)S + (uint32_t)(ly * D + lx)) : P.mkey;
    lyv[ly][lx] = (int)Y;
    lcu[ly][lx] = Cb | (Cr << 16);
  };
  auto cwriter = [&](int lx) -> bool { return R1 ? (lx == 0 || (lx & 1)) : !(lx & 1); };
  {
    constexpr int NPW = (MT::LHC + 3) / 4;  // row pairs per wave
    const uint32_t cx = (uint32_t)clampi(x0 - R + lane, 0, W - 1);
    const uint32_t oYx = cx * (uint32_t)sizeof(T), oCx = (cx >> 1) * (uint32_t)sizeof(T), oMx = cx * (uint32_t)sizeof(TM);
    uint32_t ra[NPW][6];  // Y, M of the pair's two rows, Cb, Cr
#pragma unroll
    for (int k = 0; k < NPW; k++) {
      const int j = wv + 4 * k;
      if (j < MT::LHC) {
        const int ly0 = 2 * j - R1, ly1 = ly0 + 1;
        const uint32_t ya = (uint32_t)clampi(y0 - R + ly0, 0, H - 1), yb = (uint32_t)clampi(y0 - R + ly1, 0, H - 1);
        if (ly0 >= 0) {
          ra[k][0] = (uint32_t)ld_buf<T>(rY, oYx, ya * sY);
          ra[k][1] = (uint32_t)ld_buf<TM>(rM, oMx, ya * sM);
        }
        if (ly1 < LH) {
          ra[k][2] = (uint32_t)ld_buf<T>(rY, oYx, yb * sY);
          ra[k][3] = (uint32_t)ld_buf<TM>(rM, oMx, yb * sM);
        }
        const uint32_t yc = (ly0 >= 0 ? ya : yb) >> 1;
        ra[k][4] = (uint32_t)ld_buf<T>(rCb, oCx, yc * sC);
        ra[k][5] = (uint32_t)ld_buf<T>(rCr, oCx, yc * sC);
      }
    }
#pragma unroll
    for (int k = 0; k < NPW; k++) {
      const int j = wv + 4 * k;
      if (j < MT::LHC) {
        const int ly0 = 2 * j - R1;
        if (ly0 >= 0) entry(ly0, lane, ra[k][0], ra[k][4], ra[k][5], ra[k][1]);
        if (ly0 + 1 < LH) entry(ly0 + 1, lane, ra[k][2], ra[k][4], ra[k][5], ra[k][3]);
        if (cwriter(lane)) lcw[j][(lane + R1) >> 1] = (int)((ra[k][4] * (uint32_t)S) | ((ra[k][5] * (uint32_t)S) << 16));
      }
    }
    for (int e = (int)threadIdx.x; e < NHC * LH; e += 256) {  // (more than one pass for R = 4)
      const int hy = e / NHC, lx = 64 + (e - hy * NHC);
      const uint32_t hx = (uint32_t)clampi(x0 - R + lx, 0, W - 1), hyy = (uint32_t)clampi(y0 - R + hy, 0, H - 1);
      const uint32_t Y = (uint32_t)ld_buf<T>(rY, hx * (uint32_t)sizeof(T) + hyy * sY, 0);
      const uint32_t oc = (hx >> 1) * (uint32_t)sizeof(T) + (hyy >> 1) * sC;
      const uint32_t Cb = (uint32_t)ld_buf<T>(rCb, oc, 0), Cr = (uint32_t)ld_buf<T>(rCr, oc, 0);
      const uint32_t M = (uint32_t)ld_buf<TM>(rM, hx * (uint32_t)sizeof(TM) + hyy * sM, 0);
      entry(hy, lx, Y, Cb, Cr, M);
      if (cwriter(lx)) lcw[(hy + R1) >> 1][(lx + R1) >> 1] = (int)((Cb * (uint32_t)S) | ((Cr * (uint32_t)S) << 16));
    }
  }
  __syncthreads();
  // ---- multipliers: luma -2 S wY t_Y (|value| < 2^23), chroma {-2 wCb t_Cb, -2 wCr t_Cr} as an int16 pair
  // (|value| <= 32767), host-checked
  int kY[2][4];
  uint32_t mc[2];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int k = 0; k < 4; k++) kY[r][k] = -2 * S * cw.x * tY[r][k];
#pragma unroll
  for (int h = 0; h < 2; h++)
    mc[h] = ((uint32_t)(-2 * cw.y * tC[0][h]) & 0xffffu) | ((uint32_t)(-2 * cw.z * tC[1][h]) << 16);
  int bk[2][4];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int k = 0; k < 4; k++) bk[r][k] = INT_MAX;
  // Candidate rows cr = 0 .. 2R + 1 (region rows iy0 + cr): pixel row 0 takes dy = cr, row 1 dy = cr - 1.
  struct Run {
    int K[NE], Y[NE], C[NC];
  };
  // the run of region row iy0 + cr from its key entry kp (the Y entries follow at a constant distance) and of
  // chroma row (iy0 + cr + R1) >> 1 = (iy0 >> 1) + ((cr + R1) >> 1) (iy0 is even) from cp
  auto load_run = [&](const int* kp, const int* cp, Run& u) {
    lds_ints<NE>(kp, u.K);
    lds_ints<NE>(kp + LH * LPK, u.Y);
#pragma unroll
    for (int i = 0; i < NC / 2; i++) {
      typedef int i32x2m_t __attribute__((ext_vector_type(2)));
      const i32x2m_t w = reinterpret_cast<const i32x2m_t*>(cp)[i];
      u.C[2 * i] = w.x;
      u.C[2 * i + 1] = w.y;
    }
  };
  auto score = [&](const Run& u, bool row0, bool row1) {
#pragma unroll
    for (int h = 0; h < 2; h++) {  // pixel pair h = pixels 2h, 2h + 1 and its entries 2h .. 2h + 1 + 2R
      int X[NE];
#pragma unroll
      for (int e = 2 * h; e < 2 * h + 2 + 2 * R; e++)
        X[e] = __builtin_amdgcn_sdot2(__builtin_bit_cast(s16x2m_t, mc[h]), __builtin_bit_cast(s16x2m_t, u.C[(e + R1) >> 1]),
                                      u.K[e], false);
#pragma unroll
      for (int r = 0; r < 2; r++) {
        if (r == 0 ? !row0 : !row1) continue;
#pragma unroll
        for (int k = 2 * h; k < 2 * h + 2; k++)
#pragma unroll
          for (int dx = 0; dx < D; dx++) bk[r][k] = min(bk[r][k], mad_i24(kY[r][k], u.Y[k + dx], X[k + dx]));
      }
    }
  };
  // (iteration i scores rows 2i + 1, 2i + 2 and loads rows 2i + 2, 2i + 3: the addresses advance by constant
  // steps)
  Run ua, ub;
  const int* kp = &lk[iy0][ix];
  const int* cp = &lcw[iy0 >> 1][ix >> 1];
  load_run(kp, cp, ua);
  load_run(kp + LPK, cp + R1 * LPC, ub);
  score(ua, true, false);
#pragma unroll 1
  for (int i = 0; i < R; i++) {
    kp += 2 * LPK;
    cp += LPC;
    load_run(kp, cp, ua);
    score(ub, true, true);
    load_run(kp + LPK, cp + R1 * LPC, ub);
    score(ua, true, true);
  }
  score(ub, false, true);
  // ---- epilogue, one pixel row at a time: winners, weighted squared differences, SCP; then the row sums
  uint64_t vr[2][3];
#pragma unroll
  for (int r = 0; r < 2; r++) {
    const int iy = iy0 + r, y = y0 + iy;
    uint64_t* v = vr[r];
    v[0] = v[1] = v[2] = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int x = x0 + ix + k;
      if (m[r][k] != 0) {
        const uint32_t idx = (uint32_t)(bk[r][k] - (iy * D + ix + k)) & (uint32_t)(S - 1);
        const uint32_t wdy = min(idx / (uint32_t)D, (uint32_t)(2 * R)), wdx = min(idx - (idx / (uint32_t)D) * (uint32_t)D, (uint32_t)(2 * R));
        const int ey = iy + (int)wdy, ex = ix + k + (int)wdx;
        const int rYv = lyv[ey][ex];
        const uint32_t cu = lcu[ey][ex];
        const int rCb = (int)(cu & 0xffffu), rCr = (int)(cu >> 16);
        const int d0 = tY[r][k] - rYv, d1 = tC[0][k >> 1] - rCb, d2 = tC[1][k >> 1] - rCr;
        // |d| < 2^15: the squares are positive int32
        v[0] += (uint64_t)(int64_t)(d0 * d0) * (uint64_t)m[r][k];
        v[1] += (uint64_t)(int64_t)(d1 * d1) * (uint64_t)m[r][k];
        v[2] += (uint64_t)(int64_t)(d2 * d2) * (uint64_t)m[r][k];
        if (P.scp_off >= 0) {
          const int r3[3] = {rYv, rCb, rCr};
          const int g3[3] = {g0, g1, g2};
          masked_scp_store(P, dir, x, y, true, r3, g3, scp);
        }
      } else if (P.scp_off >= 0 && y < H && x < W) {
        const int z3[3] = {0, 0, 0};
        masked_scp_store(P, dir, x, y, false, z3, z3, scp);
      }
    }
  }
  // Row sums of both rows in one reduction: even lanes carry row iy0, odd lanes row iy0 + 1 (each lane hands
  // its partner the other row's value), then 64-bit sums over the lanes of one parity within each 16-lane row
  // (row_shr 2, 4, 8): lane 14 ends with row iy0's total, lane 15 with row iy0 + 1's.
  const bool odd = lane & 1;
  uint64_t s[3];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const uint64_t own = odd ? vr[1][c] : vr[0][c], give = odd ? vr[0][c] : vr[1][c];
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)give, 0xB1, 0xf, 0xf, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(give >> 32), 0xB1, 0xf, 0xf, true);
    s[c] = own + ((uint64_t)lo | ((uint64_t)hi << 32));
  }
#define IVQM_ROWP(ctrl)                                                                                     \
  _Pragma("unroll") for (int c = 0; c < 3; c++) {                                                          \
    const uint32_t lo = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)s[c], ctrl, 0xf, 0xf, true);         \
    const uint32_t hi = (uint32_t)__builtin_amdgcn_mov_dpp((int)(uint32_t)(s[c] >> 32), ctrl, 0xf, 0xf, true); \
    s[c] += (uint64_t)lo | ((uint64_t)hi << 32);                                                           \
  }
  IVQM_ROWP(0x112) IVQM_ROWP(0x114) IVQM_ROWP(0x118)
#undef IVQM_ROWP
  const int yr = y0 + iy0 + (odd ? 1 : 0);
  if ((lane & 15) >= 14 && yr < H) {
    uint64_t* rsp = rows64 + P.rows64_off + (int64_t)dir * 3 * H + yr;
    atomicAdd((unsigned long long*)rsp, (unsigned long long)s[0]);
    atomicAdd((unsigned long long*)(rsp + H), (unsigned long long)s[1]);
    atomicAdd((unsigned long long*)(rsp + 2 * H), (unsigned long long)s[2]);
  }
}

// Form of a pair's masked search: 0 reference form, 1 keyed (narrow), 2 keyed (packed dot), 3 keyed 4:2:0.
__device__ __forceinline__ int masked_form(const PairDev& P) {
  return P.mkey == 0 ? 0 : (P.mdot == 2 ? 3 : (P.mdot ? 2 : 1));
}

// One launch per form present in the batch (the forms' register needs differ: one kernel would run
// every form at the largest); tiles of pairs of other forms exit at once.
template <int R, int F>
__global__ __launch_bounds__(256) void k_search_masked(const PairDev* __restrict__ pairs, int npairs, int ntiles,
                                                       int4 cw, const int32_t* __restrict__ gcd,
                                                       uint64_t* __restrict__ rows64, uint16_t* __restrict__ scp) {
  using MT = MaskTile<R>;
  // keyed forms: int planes of LH x LWP (4:2:0 form: its planes); the reference form: LH x LW 8-byte entries
  constexpr int NI = F == 3 ? Mask420<R>::LDS_INTS : (F == 2 ? 3 : 4) * MT::LH * MT::LWP;
  __shared__ __attribute__((aligned(16))) int ldsi[NI];
  const int b = blockIdx.x;
  if (b >= ntiles) return;
  const int pi = find_entry(pairs, npairs, b, [](const PairDev& p) { return p.mtile0; });
  const PairDev& P = pairs[pi];
  if (masked_form(P) != F) return;  // block-uniform
  constexpr int TH = F == 3 ? kMask420TileH : kMaskTileH;
  const int nx = (P.w + kMaskTileW - 1) / kMaskTileW, ny = (P.h + TH - 1) / TH;
  const int lt = b - P.mtile0;
  const int dir = lt / (nx * ny), rem = lt - dir * nx * ny;
  const int ty = rem / nx, tx = rem - ty * nx;
  const int x0 = tx * kMaskTileW, y0 = ty * TH;
  if (F == 3) {
    if (P.p0.bps == 2) {
      if (P.msk_bps == 2) masked420_tile<uint16_t, uint16_t, R>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
      else masked420_tile<uint16_t, uint8_t, R>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
    } else {
      if (P.msk_bps == 2) masked420_tile<uint8_t, uint16_t, R>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
      else masked420_tile<uint8_t, uint8_t, R>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
    }
  } else if (F == 0) {
    auto lds = reinterpret_cast<u32x2m_t (*)[MT::LW]>(ldsi);
    if (P.p0.bps == 2) {
      if (P.msk_bps == 2) masked_tile<uint16_t, uint16_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64, scp);
      else masked_tile<uint16_t, uint8_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64, scp);
    } else {
      if (P.msk_bps == 2) masked_tile<uint8_t, uint16_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64, scp);
      else masked_tile<uint8_t, uint8_t, R>(P, dir, x0, y0, lds, cw, gcd, rows64, scp);
    }
  } else {
    constexpr bool DOT = F == 2;
    if (P.p0.bps == 2) {
      if (P.msk_bps == 2) masked_tile_keyed<uint16_t, uint16_t, R, DOT>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
      else masked_tile_keyed<uint16_t, uint8_t, R, DOT>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
    } else {
      if (P.msk_bps == 2) masked_tile_keyed<uint8_t, uint16_t, R, DOT>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
      else masked_tile_keyed<uint8_t, uint8_t, R, DOT>(P, dir, x0, y0, ldsi, cw, gcd, rows64, scp);
    }
  }
}

template <int R>
hipError_t masked_r(const PairDev* pairs, int npairs, int ntiles, int4 w, const int32_t* gcd, uint64_t* rows64,
                    uint16_t* scp, uint32_t forms, hipStream_t st) {
  if (forms & 1u) hipLaunchKernelGGL((k_search_masked<R, 0>), dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64, scp);
  if (forms & 2u) hipLaunchKernelGGL((k_search_masked<R, 1>), dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64, scp);
  if (forms & 4u) hipLaunchKernelGGL((k_search_masked<R, 2>), dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64, scp);
  if (forms & 8u) hipLaunchKernelGGL((k_search_masked<R, 3>), dim3(ntiles), dim3(256), 0, st, pairs, npairs, ntiles, w, gcd, rows64, scp);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_search_masked(const PairDev* pairs, int npairs, int ntiles, int R, const int32_t cw[4],
                                const int32_t* gcd, uint64_t* rows64, uint16_t* scp, uint32_t forms, hipStream_t st) {
  if (npairs <= 0 || ntiles <= 0 || !forms) return hipSuccess;
  const int4 w = make_int4(cw[0], cw[1], cw[2], cw[3]);
  switch (R) {
    case 1: return masked_r<1>(pairs, npairs, ntiles, w, gcd, rows64, scp, forms, st);
    case 2: return masked_r<2>(pairs, npairs, ntiles, w, gcd, rows64, scp, forms, st);
    case 3: return masked_r<3>(pairs, npairs, ntiles, w, gcd, rows64, scp, forms, st);
    case 4: return masked_r<4>(pairs, npairs, ntiles, w, gcd, rows64, scp, forms, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ivqm
