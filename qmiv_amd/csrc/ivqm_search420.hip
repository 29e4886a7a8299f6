// k_search420: the corresponding-pixel-shift search of 4:2:0 pairs in keyed
// form (libs/pmbb-ivqm/src/xCorrespPixelShiftSSE.cpp:18-127 semantics: for each
// test pixel t = Tst + gcs, the first strict minimum in dy-outer raster order of
// E = sum_c w_c (t_c - r_c)^2 over the (2R+1)^2 candidates; per-row sums of the
// winner's per-component squared differences; SCP = clip(r_best - gcs, 0, Max)).
//
// Keyed score (host-checked ranges, as the 4:4:4 keyed variants of k_search):
// E = T + E', T = sum_c w_c t_c^2 constant per pixel, E' = Q(r) - 2 sum_c w_c t_c r_c.
// Nothing wraps, so argmin E = argmin E'; the key Sk * E' + off with off an
// offset that grows with the candidate's raster position makes a plain integer
// min equal to "first strict minimum". With C0 = Max + max|gcd| >= t and
// t' = C0 - t >= 0, E' = Q'(r) + 2 sum_c w_c t'_c r_c with
// Q'(r) = sum_c w_c r_c (r_c - 2 C0) per reference entry: every multiplier is a
// non-negative value: luma products are v_mad_u32_u24, chroma products
// v_dot2_u32_u16 (exact mod 2^32; the final key fits int32).
//
// What 4:2:0 allows: a chroma sample covers a 2 x 2 luma block, so
//  * the two pixels (2m, 2m + 1) of a row share their test chroma, and the
//    chroma part of a candidate's score, X = key + dot2(2 w t'_chroma, r_chroma),
//    is computed once for both; a candidate then costs one v_mad_u32_u24
//    (luma) and half a v_min3;
//  * a lane owns 2 rows x 4 columns (8 pixels, two 2 x 2 blocks of one test
//    chroma each) and the X values of a candidate row serve both pixel rows;
//  * chroma lives in LDS at chroma resolution, luma as a key plane (Sk * Q' +
//    off, int32) and a sample plane (Y, uint32).
// Per candidate row a lane reads NE keys, NE luma samples and NC chroma words
// and makes 2 (2R + 2) v_dot2 for X, then 4 (2R + 1) luma multiply-adds and
// min3 folds per pixel row.
//
// Block = 2 directions x TH / 8 waves on one 64 x TH tile (TH = 16 or 32, the
// tuning knob "search tile rows"): the first half of the waves runs direction 0
// (R2T: Tst = P0 + gcd searched in P1), the second half direction 1 (T2R:
// Tst = P1 - gcd searched in P0). Each direction stages its reference region
// once; test samples are read from the other direction's region. Outputs match
// k_search: uint32 row sums (atomics, modular), SCP pictures or fused IV-SSIM
// sub-block moments, optional winner map.
#include "ivqm_device.h"
#include "ivqm_device_util.h"

#include <limits.h>
#include <stddef.h>

namespace ivqm {

namespace {

// Key layout per search range: a region entry (ly, lx) gets off = ly * DP + lx, DP the power of two
// >= 2R + 1, so a pixel's candidate (dy, dx) carries its pixel's constant plus dy * DP + dx < Sk.
// Sy = 2^YB scales the stored chroma samples (u16 operands of v_dot2); F = Sk / Sy scales the chroma
// multipliers.
template <int R, int TH> struct S420 {
  static constexpr int D = 2 * R + 1;
  static constexpr int LOGDP = R <= 1 ? 2 : (R <= 3 ? 3 : 4);
  static constexpr int DP = 1 << LOGDP;
  static constexpr int KB = R <= 1 ? 4 : (R <= 3 ? 6 : 8);
  static constexpr int YB = R <= 1 ? 4 : 5;
  static constexpr int FB = KB - YB;
  static constexpr int R1 = R & 1;
  static constexpr int TW = 64, LW = TW + 2 * R, LH = TH + 2 * R;
  static constexpr int LPK = (LW + 3) & ~3;          // key row pitch (ints): lane runs 16-byte aligned
  static constexpr int LWC = (LW - 1 + R1) / 2 + 1;  // chroma entries of a region row
  static constexpr int LHC = (LH - 1 + R1) / 2 + 1;  // chroma rows of the region
  static constexpr int LPC = (LWC + 1) & ~1;         // chroma row pitch (ints): lane runs 8-byte aligned
  static constexpr int NE = 4 + 2 * R;               // luma entries of a lane's candidate run
  static constexpr int NC = (NE - 1 + R1) / 2 + 1;   // chroma entries of the run
  static constexpr int WPD = TH / 8;                 // waves per direction
  static constexpr int NT = 2 * WPD * 64;            // threads per block (16 TH)
  static_assert(NT == 16 * TH, "block size");
  static_assert(DP * 2 * R + 2 * R < (1 << KB), "candidate offsets must stay below the key scale");
  static_assert(NE % 2 == 0 && NC % 2 == 0, "run lengths");
  static_assert(TH == 16 || TH == 32, "tile height");
};

template <int R, int TH> struct Lds420 {
  using P = S420<R, TH>;
  int key[2][P::LH][P::LPK];
  unsigned int ys[2][P::LH][P::LPK];     // Y
  unsigned int cu[2][P::LH][P::LPK];     // {Cb, Cr} u16 pair of the entry's chroma sample (epilogue)
  int cw[2][P::LHC][P::LPC];             // {Cb * Sy, Cr * Sy} u16 pairs at chroma resolution
};

typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

// N dwords from a 16-byte aligned LDS address (16-byte reads and an 8-byte tail).
template <int N>
__device__ __forceinline__ void lds_read16(const int* p, int (&w)[N]) {
  static_assert(N % 2 == 0, "even run");
#pragma unroll
  for (int i = 0; i < N / 4; i++) {
    const i32x4 v = reinterpret_cast<const i32x4*>(p)[i];
    w[4 * i] = v.x; w[4 * i + 1] = v.y; w[4 * i + 2] = v.z; w[4 * i + 3] = v.w;
  }
  if (N % 4) {
    const i32x2 v = *reinterpret_cast<const i32x2*>(p + 4 * (N / 4));
    w[N - 2] = v.x; w[N - 1] = v.y;
  }
}
// N dwords from an 8-byte aligned LDS address.
template <int N>
__device__ __forceinline__ void lds_read8(const int* p, int (&w)[N]) {
#pragma unroll
  for (int i = 0; i < N / 2; i++) {
    const i32x2 v = reinterpret_cast<const i32x2*>(p)[i];
    w[2 * i] = v.x; w[2 * i + 1] = v.y;
  }
  if (N % 2) w[N - 1] = p[N - 1];
}

__device__ __forceinline__ void dpp_row16_add3(uint32_t v[3]) {
#define IVQM_R16(ctrl)                                                                        \
  {                                                                                           \
    const uint32_t a0 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[0], ctrl, 0xf, 0xf, true); \
    const uint32_t a1 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[1], ctrl, 0xf, 0xf, true); \
    const uint32_t a2 = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v[2], ctrl, 0xf, 0xf, true); \
    v[0] += a0; v[1] += a1; v[2] += a2;                                                       \
  }
  IVQM_R16(0x111) IVQM_R16(0x112) IVQM_R16(0x114) IVQM_R16(0x118)
#undef IVQM_R16
}

// x' + y' after v_permlane16_swap: rows (0, 1) of x summed in row 0, of y in row 1; rows (2, 3) likewise.
__device__ __forceinline__ uint32_t swap16_sum(uint32_t x, uint32_t y) {
  const auto r = __builtin_amdgcn_permlane16_swap(x, y, false, false);
  return r[0] + r[1];
}

template <typename T, int R, int TH>
__device__ __forceinline__ void search420_tile(const PairDev& Pr, const PairDev* __restrict__ pe, int x0, int y0,
                                               Lds420<R, TH>& L, int wY, int wCb, int wCr,
                                               const int32_t* __restrict__ gcd, uint32_t* __restrict__ rows,
                                               uint16_t* __restrict__ scp, uint8_t* __restrict__ best) {
  using P = S420<R, TH>;
  constexpr int D = P::D, NE = P::NE, NC = P::NC, LH = P::LH, KB = P::KB, YB = P::YB, FB = P::FB, R1 = P::R1;
  constexpr int LOGDP = P::LOGDP, DP = P::DP;
  const int W = Pr.w, H = Pr.h;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int dir = wv / P::WPD, wd = wv - dir * P::WPD;  // direction, wave within it (wave-uniform)
  const int tl = (int)threadIdx.x - dir * P::WPD * 64;  // thread within the direction
  const PicDev& Ref = dir ? Pr.p0 : Pr.p1;
  const int32_t* g4 = gcd + 4 * Pr.res_idx;
  const int g0 = dir ? -g4[0] : g4[0];
  const int g1 = dir ? -g4[1] : g4[1];
  const int g2 = dir ? -g4[2] : g4[2];
  // C0 >= every test value t = s + g (uniform)
  const int C0 = Pr.maxv + max(max(abs(g0), abs(g1)), abs(g2));

  // ---- fill: reference region rows y0 - R .. y0 + TH + R - 1, columns x0 - R .. x0 + 63 + R (clamped) ----
  // Region rows go in pairs (2j - R1, 2j + 1 - R1) of one chroma row j; wave wd stages pairs wd, wd + WPD, ...
  // and lane = region column (the 2R halo columns are staged by threads < 2R * LH).
  constexpr int NPT = P::LHC;                       // row pairs = chroma rows
  constexpr int NPW = (NPT + P::WPD - 1) / P::WPD;  // row pairs per wave
  const int cxl = clampi(x0 - R + lane, 0, W - 1);
  const uint32_t ol = (uint32_t)cxl * (uint32_t)sizeof(T), oc = (uint32_t)(cxl >> 1) * (uint32_t)sizeof(T);
  const __amdgpu_buffer_rsrc_t rs0 = plane_rsrc<T>(Ref, 0, H), rs1 = plane_rsrc<T>(Ref, 1, H),
                               rs2 = plane_rsrc<T>(Ref, 2, H);
  const uint32_t st0 = (uint32_t)plane_stride(Ref, 0) * (uint32_t)sizeof(T);
  const uint32_t st1 = (uint32_t)plane_stride(Ref, 1) * (uint32_t)sizeof(T);
  const uint32_t st2 = (uint32_t)plane_stride(Ref, 2) * (uint32_t)sizeof(T);
  int ra[NPW][4];  // luma of the pair's two rows, Cb, Cr
#pragma unroll
  for (int k = 0; k < NPW; k++) {
    const int j = wd + P::WPD * k;
    const int ly0 = 2 * j - R1, ly1 = ly0 + 1;
    if (j < NPT) {
      const uint32_t ya = (uint32_t)clampi(y0 - R + ly0, 0, H - 1), yb = (uint32_t)clampi(y0 - R + ly1, 0, H - 1);
      if (ly0 >= 0) ra[k][0] = ld_buf<T>(rs0, ol, ya * st0);
      if (ly1 < LH) ra[k][1] = ld_buf<T>(rs0, ol, yb * st0);
      const uint32_t yc = (ly0 >= 0 ? ya : yb) >> 1;
      ra[k][2] = ld_buf<T>(rs1, oc, yc * st1);
      ra[k][3] = ld_buf<T>(rs2, oc, yc * st2);
    }
  }
  constexpr int NH = 2 * R * LH;                  // halo entries (columns 64 .. 64 + 2R - 1)
  constexpr int NHK = (NH + P::WPD * 64 - 1) / (P::WPD * 64);
  int rh[NHK][3];
#pragma unroll
  for (int k = 0; k < NHK; k++) {
    const int e = tl + P::WPD * 64 * k;
    rh[k][0] = rh[k][1] = rh[k][2] = 0;
    if (e < NH) {
      const int hy = e / (2 * R), hx = e - hy * (2 * R);
      const int x = clampi(x0 - R + 64 + hx, 0, W - 1), y = clampi(y0 - R + hy, 0, H - 1);
      rh[k][0] = ldpix<T>(Ref, 0, x, y);
      rh[k][1] = ldpix<T>(Ref, 1, x, y);
      rh[k][2] = ldpix<T>(Ref, 2, x, y);
    }
  }
  // Entries: key = Sk * Q'(r) + ly * DP + lx, Y, chroma pair {Cb, Cr} and chroma word {Cb * Sy, Cr * Sy}
  // (the pair shifted; Sy * Max < 2^16 host-checked) written by one lane per chroma column.
  // Q'(r) = sum_c r_c t_c with t_c = w_c (r_c - 2 C0) = r_c w_c - 2 C0 w_c: 24-bit multiply-adds (host-checked:
  // samples below 2^12, weights below 128, 2 C0 below 2^14, so |t_c| < 2^21), exact mod 2^32 as the keys are.
  // The chroma part is shared by the two rows of a pair; each row adds its luma part in one multiply-add.
  const int c2 = 2 * C0;
  const int kY = -wY * c2, kCb = -wCb * c2, kCr = -wCr * c2;
  auto qc_of = [&](int cb, int cr) -> int {
    return mad_i24(cb, mad_i24_s(cb, wCb, kCb), __mul24(cr, mad_i24_s(cr, wCr, kCr)));
  };
  auto q_of = [&](int y, int qc) -> int { return mad_i24(y, mad_i24_s(y, wY, kY), qc); };
  auto cwriter = [&](int lx) -> bool { return R1 ? (lx == 0 || (lx & 1)) : !(lx & 1); };
#pragma unroll
  for (int k = 0; k < NPW; k++) {
    const int j = wd + P::WPD * k;
    if (j < NPT) {
      const int qc = qc_of(ra[k][2], ra[k][3]);
      const uint32_t cu = (uint32_t)ra[k][2] | ((uint32_t)ra[k][3] << 16);
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int ly = 2 * j - R1 + h;
        if (ly >= 0 && ly < LH) {
          L.key[dir][ly][lane] = (int)(((uint32_t)q_of(ra[k][h], qc) << KB) + (uint32_t)(ly * DP + lane));
          L.ys[dir][ly][lane] = (uint32_t)ra[k][h];
          L.cu[dir][ly][lane] = cu;
        }
      }
      if (cwriter(lane)) L.cw[dir][j][(lane + R1) >> 1] = (int)(cu << YB);
    }
  }
#pragma unroll
  for (int k = 0; k < NHK; k++) {
    const int e = tl + P::WPD * 64 * k;
    if (e < NH) {
      const int hy = e / (2 * R), hx = e - hy * (2 * R), lx = 64 + hx;
      const uint32_t cu = (uint32_t)rh[k][1] | ((uint32_t)rh[k][2] << 16);
      L.key[dir][hy][lx] = (int)(((uint32_t)q_of(rh[k][0], qc_of(rh[k][1], rh[k][2])) << KB) + (uint32_t)(hy * DP + lx));
      L.ys[dir][hy][lx] = (uint32_t)rh[k][0];
      L.cu[dir][hy][lx] = cu;
      if (cwriter(lx)) L.cw[dir][(hy + R1) >> 1][(lx + R1) >> 1] = (int)(cu << YB);
    }
  }
  __syncthreads();

  // ---- this lane's pixels: tile rows iy0, iy0 + 1 (iy0 even), tile columns ix .. ix + 3 ----
  const int rp = lane >> 4;
  const int iy0 = 8 * wd + 2 * rp, ix = 4 * (lane & 15);
  __builtin_assume(iy0 >= 0 && iy0 < TH);  // (small non-negative operands: 24-bit multiplies in the addressing)
  // Test samples: the other direction's region at (iy + R, ix + k + R) holds the test picture at the
  // clamped position; both rows and each pixel pair (2h, 2h + 1) share one chroma sample.
  auto test_y = [&](int r, int k) -> int { return (int)L.ys[1 - dir][iy0 + r + R][ix + k + R]; };
  auto test_c = [&](int h) -> uint32_t { return (uint32_t)L.cw[1 - dir][(iy0 + R + R1) >> 1][(ix + 2 * h + R + R1) >> 1]; };
  // Multipliers (host-checked): luma 2 Sk wY t' < 2^24, chroma {2 F wCb t'_Cb, 2 F wCr t'_Cr} u16.
  // (t'_c = C0 - g_c - s_c >= 0, below 2^14; the masks state the u16 ranges for the 24-bit multiplies; the
  // chroma pair is packed with a byte permute)
  const uint32_t kmY = ((uint32_t)(2 * wY) << KB) & 0xffffu, kmCb = ((uint32_t)(2 * wCb) << FB) & 0xffffu,
                 kmCr = ((uint32_t)(2 * wCr) << FB) & 0xffffu;
  // (uniform constants read as such, so that they are folded once rather than per sample)
  const int c0y = __builtin_amdgcn_readfirstlane(C0 - g0), c0b = __builtin_amdgcn_readfirstlane(C0 - g1),
            c0r = __builtin_amdgcn_readfirstlane(C0 - g2);
  uint32_t mY[2][4], mc[2];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int k = 0; k < 4; k++) {
      mY[r][k] = kmY * ((uint32_t)(c0y - test_y(r, k)) & 0xffffu);
      __builtin_assume(mY[r][k] < (1u << 24));  // host-checked: 2 Sk wY 2 A < 2^24
    }
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t w = test_c(h);
    const int tcb = (int)((w & 0xffffu) >> YB), tcr = (int)(w >> (16 + YB));
    mc[h] = __builtin_amdgcn_perm(kmCr * ((uint32_t)(c0r - tcr) & 0xffffu), kmCb * ((uint32_t)(c0b - tcb) & 0xffffu),
                                  0x05040100u);
  }

  int bk[2][4];
#pragma unroll
  for (int r = 0; r < 2; r++)
#pragma unroll
    for (int k = 0; k < 4; k++) bk[r][k] = INT_MAX;
  // Candidate rows cr = 0 .. 2R + 1 (region rows iy0 + cr): pixel row 0 takes dy = cr, row 1 dy = cr - 1.
  // Row 0 alone scores cr = 0 and row 1 alone cr = 2R + 1; the 2R rows between score both pixel rows
  // with one body (the keys carry the candidate's position, so the body does not depend on cr) in a
  // rolled loop over pairs of rows with two buffers: the next row's LDS reads are in flight while a row
  // is scored, and the live registers stay those of two rows.
  struct Run {
    int K[NE], Y[NE], C[NC];
  };
  // The run of region row iy0 + cr from its key entry kp (the Y entries follow at a constant distance) and of
  // chroma row ci = (iy0 + cr + R1) >> 1 = (iy0 >> 1) + ((cr + R1) >> 1) (iy0 is even) from cp.
  constexpr int YD = 2 * LH * P::LPK;  // ints from a key entry to its Y entry
  using LT = Lds420<R, TH>;
  static_assert(offsetof(LT, ys) - offsetof(LT, key) == YD * sizeof(int), "layout");
  auto load_run = [&](const int* kp, const int* cp, Run& u) {
    lds_read16<NE>(kp, u.K);
    lds_read16<NE>(kp + YD, u.Y);
    lds_read8<NC>(cp, u.C);
  };
  auto score = [&](const Run& u, bool row0, bool row1) {
#pragma unroll
    for (int h = 0; h < 2; h++) {  // pixel pair h = pixels 2h, 2h + 1 and its entries 2h .. 2h + 1 + 2R
      uint32_t X[NE];
#pragma unroll
      for (int e = 2 * h; e < 2 * h + 2 + 2 * R; e++)
        X[e] = __builtin_amdgcn_udot2(__builtin_bit_cast(u16x2, mc[h]), __builtin_bit_cast(u16x2, u.C[(e + R1) >> 1]),
                                      (uint32_t)u.K[e], false);
#pragma unroll
      for (int r = 0; r < 2; r++) {
        if (r == 0 ? !row0 : !row1) continue;
#pragma unroll
        for (int k = 2 * h; k < 2 * h + 2; k++)
#pragma unroll
          for (int dx = 0; dx < D; dx++) {
            const int e = k + dx;
            const int v = (int)(__umul24(mY[r][k], (uint32_t)u.Y[e]) + X[e]);
            bk[r][k] = min(bk[r][k], v);
          }
      }
    }
  };
  // (iteration i scores rows cr = 2i + 1, 2i + 2 and loads rows 2i + 2, 2i + 3, chroma rows
  // (iy0 >> 1) + 1 + i and (iy0 >> 1) + 1 + i + R1: the loop's addresses advance by constant steps)
  Run ua, ub;
  const int* kp = &L.key[dir][iy0][ix];
  const int* cp = &L.cw[dir][iy0 >> 1][ix >> 1];
  load_run(kp, cp, ua);
  load_run(kp + P::LPK, cp + R1 * P::LPC, ub);
  score(ua, true, false);
#pragma unroll 1
  for (int i = 0; i < R; i++) {
    kp += 2 * P::LPK;
    cp += P::LPC;
    load_run(kp, cp, ua);
    score(ub, true, true);
    load_run(kp + P::LPK, cp + R1 * P::LPC, ub);
    score(ua, true, true);
  }
  score(ub, false, true);

  // the epilogue's rows run after the search, one at a time, and read the test samples again rather than
  // keeping them in registers through the search (bounds the live registers)
  __builtin_amdgcn_sched_barrier(0);
  asm volatile("" ::: "memory");
  // The epilogue reads its pair fields (pe) here, after the search: loaded with the fields the fill needs
  // they would stay live through it and be spilled to vector-register lanes (VALU reloads).
  // ---- epilogue, one pixel row of the lane at a time: winners, squared differences, row sums, winner map,
  // SCP / fused moments ----
  // The winner of pixel (iy0 + r, ix + k) is region entry (iy0 + r + ddy, ix + k + ddx) with
  // u = ddy * DP + ddx = (key - pixel constant) mod Sk; its Y and chroma words sit at the entry's offset
  // 4 (u + (LPK - DP) (u >> LOGDP)) from the pixel's own entry.
  const bool edge = x0 + 64 > W || y0 + TH > H;  // block-uniform
  const bool want_scp = pe->scp_off >= 0, want_mom = pe->mom_off >= 0;  // block-uniform
  const int gc[3] = {g0, g1, g2};
  // clip(r - g, 0, Max) = (r +sat ca) -sat cb with bias = 2^16 - 1 - Max and g clamped to [-Max, Max]
  // (which leaves the clip unchanged for r in [0, Max]): g >= 0: ca = bias (no saturation as r <= Max),
  // cb = bias + g <= 2^16 - 1, giving max(r - g, 0); g < 0: ca = bias - g <= 2^16 - 1 (the sum saturates
  // where r - g > Max), cb = bias, giving min(r - g, Max)
  u16x2 ca[3], cb[3];
  {
    const int maxv = pe->maxv, bias = 0xffff - maxv;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int g = clampi(gc[c], -maxv, maxv);
      ca[c] = (unsigned short)(g < 0 ? bias - g : bias);
      cb[c] = (unsigned short)(g < 0 ? bias : bias + g);
    }
  }
  // test chroma of pixel pair h, per component, in both halves (shared by both rows of the lane): the
  // samples (tca, the a samples of the moments) and plus g (tcp)
  uint32_t tcp[2][2];
  u16x2 tca[2][2];
#pragma unroll
  for (int h = 0; h < 2; h++) {
    const uint32_t w = L.cu[1 - dir][iy0 + R][ix + 2 * h + R];
    tca[0][h] = (unsigned short)(w & 0xffffu);
    tca[1][h] = (unsigned short)(w >> 16);
    tcp[0][h] = __builtin_bit_cast(uint32_t, tca[0][h] + (unsigned short)g1);
    tcp[1][h] = __builtin_bit_cast(uint32_t, tca[1][h] + (unsigned short)g2);
  }
  // fused IV-SSIM partials over the lane's 8 pixels: per component sum a and sum b as u16 pairs (pixels
  // 2h, 2h + 1), S_aa, S_bb, S_ab; a = the test picture's sample (t - g), b = the SCP sample
  u16x2 sa[3], sb[3];
  uint32_t saa[3], sbb[3], sab[3];
  // MS-SSIM level 1 of the SCP (pe->scp1_off >= 0, with the moments): each 2 x 2 quad of the lane's
  // pixels (rows iy0, iy0 + 1, columns ix + 2h, ix + 2h + 1) summed per component as a u16 pair
  const bool want_scp1 = pe->scp1_off >= 0;  // block-uniform
  u16x2 qs[3][2];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    sa[c] = sb[c] = (unsigned short)0;
    saa[c] = sbb[c] = sab[c] = 0u;
    qs[c][0] = qs[c][1] = (unsigned short)0;
  }
  constexpr int UMAX = DP * 2 * R + 2 * R;
  uint32_t qr[2][3];  // the lane's squared differences per row and component (row sums after the rows)
#pragma unroll
  for (int r = 0; r < 2; r++) {
    __builtin_amdgcn_sched_barrier(0);
    const int y = y0 + iy0 + r;
    const unsigned int* ybase = &L.ys[dir][iy0 + r][ix];
    const unsigned int* cbase = &L.cu[dir][iy0 + r][ix];
    const uint32_t nb = (uint32_t)(-((iy0 + r) * DP + ix));  // minus the row's pixel constant at k = 0
    uint32_t wy[4], wc[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      // exact keys give u <= UMAX; the clamp keeps the reads inside the region regardless
      const uint32_t u = min(((uint32_t)bk[r][k] + nb + (uint32_t)(-k)) & ((1u << KB) - 1u), (uint32_t)UMAX);
      const uint32_t o = u + (uint32_t)(P::LPK - DP) * (u >> LOGDP);  // in entries
      wy[k] = ybase[k + o];
      wc[k] = cbase[k + o];
    }
    // u16 pairs of pixels (2h, 2h + 1): test + g and winner samples per component
    uint32_t tp[3][2], wp[3][2];
    u16x2 tya[2];  // test luma samples (the a samples of the moments)
#pragma unroll
    for (int h = 0; h < 2; h++) {
      const unsigned int* ty = &L.ys[1 - dir][iy0 + r + R][ix + 2 * h + R];
      tya[h] = u16x2{(unsigned short)ty[0], (unsigned short)ty[1]};
      tp[0][h] = __builtin_bit_cast(uint32_t, tya[h] + (unsigned short)g0);
      tp[1][h] = tcp[0][h];
      tp[2][h] = tcp[1][h];
      wp[0][h] = __builtin_amdgcn_perm(wy[2 * h + 1], wy[2 * h], 0x05040100u);
      wp[1][h] = __builtin_amdgcn_perm(wc[2 * h + 1], wc[2 * h], 0x05040100u);
      wp[2][h] = __builtin_amdgcn_perm(wc[2 * h + 1], wc[2 * h], 0x07060302u);
    }
    // |t - r| <= Max + |gcd| < 2^15 (host-checked): int16 differences, v_dot2 adds their squares
    uint32_t q[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      s16x2 d[2];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        d[h] = __builtin_bit_cast(s16x2, tp[c][h]) - __builtin_bit_cast(s16x2, wp[c][h]);
        if (edge) {
          const int x = x0 + ix + 2 * h;
          const uint32_t msk = (y < H && x < W ? 0xffffu : 0u) | (y < H && x + 1 < W ? 0xffff0000u : 0u);
          d[h] = __builtin_bit_cast(s16x2, __builtin_bit_cast(uint32_t, d[h]) & msk);
        }
      }
      q[c] = (uint32_t)__builtin_amdgcn_sdot2(d[1], d[1], __builtin_amdgcn_sdot2(d[0], d[0], 0, false), false);
    }
#pragma unroll
    for (int c = 0; c < 3; c++) qr[r][c] = q[c];
    if (pe->best_off >= 0 && y < H) {
      typedef __attribute__((address_space(1))) uint8_t GU8;
      const int x = x0 + ix;
      GU8* bm = (GU8*)best + pe->best_off + (size_t)dir * W * H + (size_t)y * W + x;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const uint32_t u = min((uint32_t)(bk[r][k] - ((iy0 + r) * DP + ix + k)) & ((1u << KB) - 1u), (uint32_t)UMAX);
        if (x + k < W) bm[k] = (uint8_t)((u >> LOGDP) * D + (u & (DP - 1)));
      }
    }
    if (want_scp || want_mom) {
      // SCP samples clip(r - g, 0, Max) for r in [0, Max] as two saturating u16 operations with uniform
      // operands: (r +sat ca) -sat cb (see ca, cb)
      u16x2 sv[3][2];
#pragma unroll
      for (int c = 0; c < 3; c++)
#pragma unroll
        for (int h = 0; h < 2; h++)
          sv[c][h] = __builtin_elementwise_sub_sat(
              __builtin_elementwise_add_sat(__builtin_bit_cast(u16x2, wp[c][h]), ca[c]), cb[c]);
      if (want_scp && y < H) {
        typedef __attribute__((address_space(1))) uint8_t GU8;
        typedef __attribute__((address_space(1))) u32x2 GU32x2;
        const size_t plane_bytes = (size_t)pe->scp_pitch * H * 2;
        GU8* pic = (GU8*)(scp + pe->scp_off) + (size_t)(1 - dir) * 3 * plane_bytes;
        const uint32_t eo = ((uint32_t)y * (uint32_t)pe->scp_pitch + (uint32_t)(x0 + ix)) * 2u;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const u32x2 o = {__builtin_bit_cast(uint32_t, sv[c][0]), __builtin_bit_cast(uint32_t, sv[c][1])};
          __builtin_nontemporal_store(o, (GU32x2*)(pic + c * plane_bytes + eo));  // read once, by the SSIM pass
        }
      }
      if (want_mom) {
        // at most 12-bit samples (host-checked): no u16 half overflows over 16 pixels. The chroma a samples
        // are the same in both rows: their S_a and S_aa are taken once, after the rows.
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const u16x2 ta = c ? tca[c - 1][h] : tya[h];
            if (c == 0) {
              sa[c] += ta;
              saa[c] = __builtin_amdgcn_udot2(ta, ta, saa[c], false);
            }
            sb[c] += sv[c][h];
            sbb[c] = __builtin_amdgcn_udot2(sv[c][h], sv[c][h], sbb[c], false);
            sab[c] = __builtin_amdgcn_udot2(ta, sv[c][h], sab[c], false);
          }
        if (want_scp1) {
#pragma unroll
          for (int c = 0; c < 3; c++)
#pragma unroll
            for (int h = 0; h < 2; h++) qs[c][h] += sv[c][h];
        }
      }
    }
  }
  // Row sums of both rows in one reduction: even lanes carry row iy0, odd lanes row iy0 + 1 (each lane hands
  // its partner the other row's value), then sums over the lanes of one parity within each 16-lane row:
  // lane 14 ends with row iy0's total, lane 15 with row iy0 + 1's.
  {
    const bool odd = lane & 1;
    uint32_t v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const uint32_t own = odd ? qr[1][c] : qr[0][c], give = odd ? qr[0][c] : qr[1][c];
      v[c] = own + (uint32_t)__builtin_amdgcn_mov_dpp((int)give, 0xB1, 0xf, 0xf, true);  // quad_perm [1, 0, 3, 2]
    }
#define IVQM_R16P(ctrl)                                                                    \
    {                                                                                      \
      const uint32_t a0 = (uint32_t)__builtin_amdgcn_mov_dpp((int)v[0], ctrl, 0xf, 0xf, true); \
      const uint32_t a1 = (uint32_t)__builtin_amdgcn_mov_dpp((int)v[1], ctrl, 0xf, 0xf, true); \
      const uint32_t a2 = (uint32_t)__builtin_amdgcn_mov_dpp((int)v[2], ctrl, 0xf, 0xf, true); \
      v[0] += a0; v[1] += a1; v[2] += a2;                                                  \
    }
    IVQM_R16P(0x112) IVQM_R16P(0x114) IVQM_R16P(0x118)  // row_shr 2, 4, 8
#undef IVQM_R16P
    const int yr = y0 + iy0 + (odd ? 1 : 0);
    if ((lane & 15) >= 14 && yr < H) {
      uint32_t* rsp = rows + pe->rows_off + (size_t)dir * 3 * H + yr;
      atomicAdd(rsp, v[0]);
      atomicAdd(rsp + H, v[1]);
      atomicAdd(rsp + 2 * H, v[2]);
    }
  }
  if (want_scp1) {
    // level 1 = 2 x 2 rounded mean (xPixelOpsSTD.cpp:50-67, odd sizes truncated) at (x1 + h, y1)
    const int W1 = W >> 1, H1 = H >> 1;
    const int x1 = (x0 + ix) >> 1, y1 = (y0 + iy0) >> 1;
    if (y1 < H1 && x1 < W1) {
      typedef __attribute__((address_space(1))) uint16_t GU16;
      typedef __attribute__((address_space(1))) uint32_t GU32;
      const int64_t plane = (int64_t)pe->scp1_pitch * H1;
      GU16* pic = (GU16*)(scp + pe->scp1_off) + (int64_t)(1 - dir) * 3 * plane + (int64_t)y1 * pe->scp1_pitch + x1;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const uint32_t v0 = ((uint32_t)qs[c][0].x + qs[c][0].y + 2u) >> 2;
        const uint32_t v1 = ((uint32_t)qs[c][1].x + qs[c][1].y + 2u) >> 2;
        if (x1 + 1 < W1)  // x1 even, the pitch even: a 4-byte aligned pair
          __builtin_nontemporal_store(v0 | (v1 << 16), (GU32*)(pic + c * plane));
        else
          pic[c * plane] = (uint16_t)v0;
      }
    }
  }
  if (want_mom) {
    // Fused IV-SSIM: a 4 x 4 sub-block's rows (0, 1) and (2, 3) are lanes l and l ^ 16. Per lane and
    // component {S_a | S_b << 16, S_aa, S_bb, S_ab}; the lane-pair reduction leaves component 0 and words
    // 0, 2 of component 2 in even 16-lane rows, component 1 and words 1, 3 of component 2 in odd rows.
#pragma unroll
    for (int c = 1; c < 3; c++) {  // the chroma a samples: both rows alike
      sa[c] = (tca[c - 1][0] + tca[c - 1][1]) << (unsigned short)1;
      saa[c] = __builtin_amdgcn_udot2(tca[c - 1][1], tca[c - 1][1],
                                      __builtin_amdgcn_udot2(tca[c - 1][0], tca[c - 1][0], 0u, false), false) << 1;
    }
    uint32_t m[3][4];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const u16x2 lo = {sa[c].x, sb[c].x}, hi = {sa[c].y, sb[c].y};
      m[c][0] = __builtin_bit_cast(uint32_t, lo + hi);
      m[c][1] = saa[c];
      m[c][2] = sbb[c];
      m[c][3] = sab[c];
    }
    uint32_t f[4], e2[2];
#pragma unroll
    for (int i = 0; i < 4; i++) f[i] = swap16_sum(m[0][i], m[1][i]);
    e2[0] = swap16_sum(m[2][0], m[2][1]);
    e2[1] = swap16_sum(m[2][2], m[2][3]);
    const int odd = rp & 1;
    const int sby = (y0 >> 2) + 2 * wd + (rp >> 1), sbx = (x0 >> 2) + (lane & 15);
    if (sby < pe->mom_rows && sbx < pe->mom_cols) {
      typedef __attribute__((address_space(1))) u32x4 GU32x4;
      typedef __attribute__((address_space(1))) uint32_t GU32;
      GU32x4* base = (GU32x4*)(scp + pe->mom_off);
      const int64_t cstride = (int64_t)pe->mom_rows * pe->mom_pitch;
      const int64_t sb = (int64_t)sby * pe->mom_pitch + sbx;
      const u32x4 fin = {f[0], f[1], f[2], f[3]};
      __builtin_nontemporal_store(fin, base + ((int64_t)dir * 3 + odd) * cstride + sb);  // read once, by the SSIM pass
      GU32* c2p = (GU32*)(base + ((int64_t)dir * 3 + 2) * cstride + sb);
      __builtin_nontemporal_store(e2[0], c2p + odd);
      __builtin_nontemporal_store(e2[1], c2p + 2 + odd);
    }
  }
}

template <int R, int TH>
__global__ __launch_bounds__(16 * TH) void k_search420(const PairDev* __restrict__ pairs, int npairs,
                                                                int ntiles, int wY, int wCb, int wCr,
                                                                const int32_t* __restrict__ gcd,
                                                                uint32_t* __restrict__ rows, uint16_t* __restrict__ scp,
                                                                uint8_t* __restrict__ best) {
  __shared__ __attribute__((aligned(16))) Lds420<R, TH> lds;
  // XCD-aware order: XCD g runs a contiguous run of tiles, whose halos meet in its L2.
  const int per = (ntiles + kXcds - 1) / kXcds;
  const int tile = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (tile >= ntiles) return;  // padding of the last XCD's run (uniform per block)
  const int pi = find_entry(pairs, npairs, tile, [](const PairDev& p) { return p.tile0b; });
  const PairDev P = pairs[pi];
  const int lt = tile - P.tile0b;
  const int ntx = (P.w + 63) / 64;
  const int ty = lt / ntx, tx = lt - ty * ntx;
  if (P.p0.bps == 2)
    search420_tile<uint16_t, R, TH>(P, pairs + pi, tx * 64, ty * TH, lds, wY, wCb, wCr, gcd, rows, scp, best);
  else
    search420_tile<uint8_t, R, TH>(P, pairs + pi, tx * 64, ty * TH, lds, wY, wCb, wCr, gcd, rows, scp, best);
}

template <int R, int TH>
hipError_t search420_launch(const PairDev* pairs, int npairs, int ntiles, const int32_t cw[4], const int32_t* gcd,
                            uint32_t* rows, uint16_t* scp, uint8_t* best, hipStream_t st) {
  const int per = (ntiles + kXcds - 1) / kXcds;
  hipLaunchKernelGGL((k_search420<R, TH>), dim3(kXcds * per), dim3(S420<R, TH>::NT), 0, st, pairs, npairs, ntiles,
                     cw[0], cw[1], cw[2], gcd, rows, scp, best);
  return hipGetLastError();
}

template <int R>
hipError_t search420_r(const PairDev* pairs, int npairs, int ntiles, int th, const int32_t cw[4], const int32_t* gcd,
                       uint32_t* rows, uint16_t* scp, uint8_t* best, hipStream_t st) {
  if (th == 16) return search420_launch<R, 16>(pairs, npairs, ntiles, cw, gcd, rows, scp, best, st);
  if (th == 32) return search420_launch<R, 32>(pairs, npairs, ntiles, cw, gcd, rows, scp, best, st);
  return hipErrorInvalidValue;
}

}  // namespace

int search420_key_bits(int R) { return R <= 1 ? 4 : (R <= 3 ? 6 : 8); }
int search420_sample_bits(int R) { return R <= 1 ? 4 : 5; }

hipError_t launch_search420(const PairDev* pairs, int npairs, int ntiles, int R, int th, const int32_t cw[4],
                            const int32_t* gcd, uint32_t* rows, uint16_t* scp, uint8_t* best, hipStream_t st) {
  if (npairs <= 0 || ntiles <= 0) return hipSuccess;
  switch (R) {
    case 1: return search420_r<1>(pairs, npairs, ntiles, th, cw, gcd, rows, scp, best, st);
    case 2: return search420_r<2>(pairs, npairs, ntiles, th, cw, gcd, rows, scp, best, st);
    case 3: return search420_r<3>(pairs, npairs, ntiles, th, cw, gcd, rows, scp, best, st);
    case 4: return search420_r<4>(pairs, npairs, ntiles, th, cw, gcd, rows, scp, best, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace ivqm
