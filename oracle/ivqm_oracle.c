/*
 * ORACLE - test infrastructure only.
 *
 * Plain-C, scalar restatement of QMIV's metric hot path (pmbb-ivqm over
 * pmbb-core), used as the parity checker for the HIP implementation. Only the
 * tests, the smoke check and the CPU-baseline leg of bench.py may load it; the
 * product path never does. Built with -O2 -ffp-contract=off so every FP64
 * operation is rounded exactly as written.
 *
 * Pictures are 4:4:4 planar uint16 planes, tightly packed (stride = width).
 * The reference pads pictures with a Nearest-extended margin of
 * roundup(R, 4) samples (apps/QMIV/src/xAppQMIV.cpp:353,
 * libs/pmbb-core/src/xMarginOps.h:40-73); every read the search makes lies
 * inside that margin, so the oracle clamps coordinates instead.
 *
 * Each function names the reference lines it follows. How each one is pinned
 * (golden vectors from the compiled reference, closed forms, brute force) is
 * listed in tests/test_oracle.py; no function here is left unpinned.
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_EXPORT __attribute__((visibility("default")))

static inline int32_t or_clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* ---------------------------------------------------------------------------
 * Kahan-Babuska-Neumaier summation: libs/pmbb-core/src/xKBNS.h:36-52 (scalar),
 * xKBNS.cpp:14-26 (SSE twin, same arithmetic), xKBNS.cpp:31-54 (4-lane).
 * ------------------------------------------------------------------------- */
typedef struct { double s, c; } or_kbns;
static inline void or_kbns_acc(or_kbns* k, double v) {
  double t = k->s + v;
  if (fabs(k->s) >= fabs(v)) k->c += ((k->s - t) + v);
  else                       k->c += ((v - t) + k->s);
  k->s = t;
}
static inline double or_kbns_result(const or_kbns* k) { return k->s + k->c; }

OR_EXPORT double or_kbns_sum(const double* x, int64_t n) {
  or_kbns k = {0.0, 0.0};
  for (int64_t i = 0; i < n; i++) or_kbns_acc(&k, x[i]);
  return or_kbns_result(&k);
}

/* ---------------------------------------------------------------------------
 * Rounding helpers: libs/pmbb-core/src/xHelpersFLT.h:22-23 (x86 branch):
 * value + copysign(0.5, value) in the operand's precision, then truncation.
 * ------------------------------------------------------------------------- */
OR_EXPORT int32_t or_round_f32(float v) { float a = v + copysignf(0.5f, v); return (int32_t)a; }
OR_EXPORT int32_t or_round_f64(double v) { double a = v + copysign(0.5, v); return (int32_t)a; }

/* ---------------------------------------------------------------------------
 * 4:2:0 -> 4:4:4 nearest replication: libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49.
 * 2x2 rounded-mean downsampling: xPixelOpsSTD.cpp:50-67 (dst dims = src >> 1).
 * ------------------------------------------------------------------------- */
OR_EXPORT void or_upsample_hv(uint16_t* dst, const uint16_t* src, int32_t dst_w, int32_t dst_h) {
  const int32_t sw = dst_w >> 1;
  for (int32_t y = 0; y < dst_h; y++)
    for (int32_t x = 0; x < dst_w; x++) dst[(size_t)y * dst_w + x] = src[(size_t)(y >> 1) * sw + (x >> 1)];
}
/* 4:2:2 -> 4:4:4: xPixelOpsSTD::UpsampleH (libs/pmbb-core/src/xPixelOpsSTD.cpp:107-120), each
 * chroma sample replicated into two horizontally adjacent positions (xSeq.cpp:266-280). */
OR_EXPORT void or_upsample_h(uint16_t* dst, const uint16_t* src, int32_t dst_w, int32_t dst_h) {
  const int32_t sw = dst_w >> 1;
  for (int32_t y = 0; y < dst_h; y++)
    for (int32_t x = 0; x < dst_w; x++) dst[(size_t)y * dst_w + x] = src[(size_t)y * sw + (x >> 1)];
}
OR_EXPORT void or_downsample_hv(uint16_t* dst, const uint16_t* src, int32_t src_w, int32_t src_h) {
  const int32_t dw = src_w >> 1, dh = src_h >> 1;
  for (int32_t y = 0; y < dh; y++)
    for (int32_t x = 0; x < dw; x++) {
      const uint16_t* s0 = src + (size_t)(2 * y) * src_w + 2 * x;
      const uint16_t* s1 = s0 + src_w;
      dst[(size_t)y * dw + x] = (uint16_t)(((int32_t)s0[0] + s0[1] + s1[0] + s1[1] + 2) >> 2);
    }
}

/* ---------------------------------------------------------------------------
 * Distortion: SD = sum(Tst - Ref), libs/pmbb-core/src/xDistortionSTD.cpp:12-22;
 * on x86-64-v4 builds the <=14-bit SD accumulates in int32 lanes and returns a
 * sign-extended int32 (xDistortionAVX512.cpp:14-73, dispatch
 * xDistortionAVX512.h:25); wrap32 != 0 reproduces that.
 * SSD = sum((Tst - Ref)^2) in uint64: xDistortionSTD.cpp:34-44 (exact; the
 * AVX-512 14-bit kernel xDistortionAVX512.cpp:202-258 is exact too).
 * ------------------------------------------------------------------------- */
OR_EXPORT int64_t or_sd(const uint16_t* tst, const uint16_t* ref, int64_t n, int32_t wrap32) {
  int64_t sd = 0;
  for (int64_t i = 0; i < n; i++) sd += (int32_t)tst[i] - (int32_t)ref[i];
  if (wrap32) sd = (int64_t)(int32_t)(uint32_t)(uint64_t)sd;
  return sd;
}
OR_EXPORT uint64_t or_ssd(const uint16_t* tst, const uint16_t* ref, int64_t n) {
  uint64_t s = 0;
  for (int64_t i = 0; i < n; i++) { int64_t d = (int64_t)tst[i] - (int64_t)ref[i]; s += (uint64_t)(d * d); }
  return s;
}
/* Per-row SSD of one plane: the row loop of xWSPSNR::xCalcCmpWSPSNR,
 * libs/pmbb-ivqm/src/xWSPSNR.cpp:91-98. */
OR_EXPORT void or_row_ssd(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, uint64_t* rows) {
  for (int32_t y = 0; y < h; y++) rows[y] = or_ssd(tst + (size_t)y * w, ref + (size_t)y * w, w);
}

/* ---------------------------------------------------------------------------
 * Global colour difference: libs/pmbb-ivqm/src/xGlobClrDiff.cpp:23-56.
 * CalcGlobalColorDiff(Tst, Ref): S = SD(Tst, Ref) per component, mean over
 * Area in FP64, round half away, clip to +-round_f32(coef_f32 * Max_f32).
 * The QMIV app passes (P1, P0) through xGlobClrDiffProc (xGlobClrDiff.h:41),
 * i.e. GCD_R2T = round(mean(P1 - P0)).
 * ------------------------------------------------------------------------- */
OR_EXPORT void or_gcd(const uint16_t* const tst[3], const uint16_t* const ref[3], int32_t w, int32_t h,
                      int32_t bit_depth, const float coef[4], int32_t wrap32, int32_t gcd[4]) {
  const int32_t maxv = (1 << bit_depth) - 1;
  const double area = (double)((int64_t)w * h);
  for (int c = 0; c < 4; c++) {
    const int32_t maxdiff = or_round_f32(coef[c] * (float)maxv);
    int64_t sd = c < 3 ? or_sd(tst[c], ref[c], (int64_t)w * h, wrap32) : 0;
    double avg = (double)sd / area;
    int32_t g = or_round_f64(avg);
    gcd[c] = g < -maxdiff ? -maxdiff : (g > maxdiff ? maxdiff : g);
  }
}

/* ---------------------------------------------------------------------------
 * Corresponding-pixel-shift search, x86 default (SSE) semantics:
 * libs/pmbb-ivqm/src/xCorrespPixelShiftSSE.cpp:18-70 (distance rows) and
 * :75-127 (shift-compensated picture). Portable twin:
 * xCorrespPixelShiftSTD.cpp:89-146, :249-267.
 *   t_c = Tst[y][x] + gcs_c (int32)
 *   for dy in [-R,R] (outer), dx in [-R,R] (inner):
 *     D_c = (t_c - Ref[y+dy][x+dx]_c)^2 (int32, wrapping), E = sum_c w_c D_c
 *     (int32, wrapping, the 4th lane included), keep the first strict minimum.
 *   row_c += D_best_c (int32, wrapping); row is returned sign-extended to
 *   uint64 (xCorrespPixelShiftSSE.cpp:38-40, libs/pmbb-core/src/xVec.h:270).
 *   SCP_c = clip(Ref_best_c - gcs_c, 0, Max) (xCorrespPixelShiftSSE.cpp:94-95).
 * rows: H x 4 uint64 (may be NULL). scp: 3 planes (may be NULL).
 * best: optional W*H bytes, raster index (dy+R)*(2R+1)+(dx+R) of the winner.
 * ------------------------------------------------------------------------- */
OR_EXPORT void or_cps(const uint16_t* const tst[3], const uint16_t* const ref[3], int32_t w, int32_t h,
                      int32_t bit_depth, const int32_t gcs[4], int32_t R, const int32_t cw[4], uint64_t* rows,
                      uint16_t* const scp[3], uint8_t* best) {
  const int32_t maxv = (1 << bit_depth) - 1;
  /* rows are independent: threads over y change no result */
#pragma omp parallel for schedule(dynamic, 16)
  for (int32_t y = 0; y < h; y++) {
    uint32_t row[4] = {0, 0, 0, 0};
    for (int32_t x = 0; x < w; x++) {
      int32_t t[4];
      for (int c = 0; c < 3; c++) t[c] = (int32_t)tst[c][(size_t)y * w + x] + gcs[c];
      t[3] = gcs[3];
      int32_t best_err = INT32_MAX;
      uint32_t best_d[4] = {0, 0, 0, 0};
      int32_t best_r[4] = {0, 0, 0, 0};
      int32_t best_i = 0, idx = 0;
      for (int32_t dy = -R; dy <= R; dy++) {
        const int32_t ry = or_clampi(y + dy, 0, h - 1);
        for (int32_t dx = -R; dx <= R; dx++, idx++) {
          const int32_t rx = or_clampi(x + dx, 0, w - 1);
          int32_t r[4];
          for (int c = 0; c < 3; c++) r[c] = ref[c][(size_t)ry * w + rx];
          r[3] = 0;
          uint32_t d[4], e = 0;
          for (int c = 0; c < 4; c++) {
            uint32_t diff = (uint32_t)(t[c] - r[c]);
            d[c] = diff * diff;
            e += d[c] * (uint32_t)cw[c];
          }
          if ((int32_t)e < best_err) {
            best_err = (int32_t)e;
            memcpy(best_d, d, sizeof(d));
            memcpy(best_r, r, sizeof(r));
            best_i = idx;
          }
        }
      }
      for (int c = 0; c < 4; c++) row[c] += best_d[c];
      if (scp)
        for (int c = 0; c < 3; c++) {
          int32_t v = best_r[c] - gcs[c];
          scp[c][(size_t)y * w + x] = (uint16_t)(v < 0 ? 0 : (v > maxv ? maxv : v));
        }
      if (best) best[(size_t)y * w + x] = (uint8_t)best_i;
    }
    if (rows)
      for (int c = 0; c < 4; c++) rows[(size_t)y * 4 + c] = (uint64_t)(int64_t)(int32_t)row[c];
  }
}

/* ---------------------------------------------------------------------------
 * PSNR from SSD: libs/pmbb-ivqm/src/xPSNR.cpp:186-192 (and :137-151).
 * ------------------------------------------------------------------------- */
OR_EXPORT double or_psnr_from_ssd(double ssd, int64_t area, int32_t bit_depth) {
  const uint64_t maxv = (uint64_t)((1 << bit_depth) - 1);
  const uint64_t MAX = (uint64_t)area * (maxv * maxv);
  return ssd > 0 ? 10.0 * log10((double)MAX / ssd) : DBL_MAX;
}

/* ERP weights: libs/pmbb-ivqm/src/xWeightedSpherically.cpp:13-29. */
OR_EXPORT void or_ws_weights(int32_t h, int32_t lat_range_deg, double* wts, double* correction) {
  const double eh = 180.0 * (double)h / (double)lat_range_deg;
  const double eo = (eh - h) / 2.0;
  const double pi = 3.14159265358979323846;
  for (int32_t y = 0; y < h; y++) wts[y] = cos((y + eo - (eh / 2 - 0.5)) * pi / eh);
  if (correction) *correction = (double)h / or_kbns_sum(wts, h);
}

/* Component-weighted mean: (v * w).getSum() * (1 / sum(w)),
 * libs/pmbb-ivqm/src/xIVPSNR.cpp:120-123, xMetricQMIV.cpp:91-97. */
static double or_weighted_mean4(const double v[4], const int32_t cw[4]) {
  const double inv = 1.0 / (double)(cw[0] + cw[1] + cw[2] + cw[3]);
  return (v[0] * (double)cw[0] + v[1] * (double)cw[1] + v[2] * (double)cw[2] + v[3] * (double)cw[3]) * inv;
}

/* ---------------------------------------------------------------------------
 * Asymmetric IV-PSNR quality from distance rows:
 * libs/pmbb-ivqm/src/xIVPSNR.cpp:99-125. Non-ERP: uint64 std::accumulate of
 * the rows (wrapping). ERP: 4-lane KBNS of (double)row[y] * w[y] in y order,
 * without the WS correction factor. PSNR per component of (err > 0 ? err : 1),
 * then the component-weighted mean.
 * ------------------------------------------------------------------------- */
OR_EXPORT double or_ivpsnr_quality(const uint64_t* rows, int32_t w, int32_t h, int32_t bit_depth, const double* erp_w,
                                   const int32_t cw_avg[4]) {
  double err[4] = {0, 0, 0, 0};
  if (erp_w) {
    or_kbns k[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
    for (int32_t y = 0; y < h; y++)
      for (int c = 0; c < 4; c++) or_kbns_acc(&k[c], (double)rows[(size_t)y * 4 + c] * erp_w[y]);
    for (int c = 0; c < 4; c++) err[c] = or_kbns_result(&k[c]);
  } else {
    uint64_t s[4] = {0, 0, 0, 0};
    for (int32_t y = 0; y < h; y++)
      for (int c = 0; c < 4; c++) s[c] += rows[(size_t)y * 4 + c];
    for (int c = 0; c < 4; c++) err[c] = (double)s[c];
  }
  double q[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++) q[c] = or_psnr_from_ssd(err[c] > 0 ? err[c] : 1.0, (int64_t)w * h, bit_depth);
  return or_weighted_mean4(q, cw_avg);
}

/* ---------------------------------------------------------------------------
 * Full IV-PSNR of a picture pair (P0 = test, P1 = reference) as the QMIV app
 * computes it: xAppQMIV.cpp:927-933 (GCD), :1022-1043 -> xIVPSNR.cpp:45-57.
 * R2T searches P0 + GCD in P1, T2R searches P1 - GCD in P0; result = min.
 * ------------------------------------------------------------------------- */
OR_EXPORT double or_ivpsnr(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                           int32_t bit_depth, const int32_t gcd_r2t[4], int32_t R, const int32_t cw_search[4],
                           const int32_t cw_avg[4], const double* erp_w, double* r2t_out, double* t2r_out,
                           uint64_t* rows_r2t, uint64_t* rows_t2r) {
  int32_t ng[4];
  for (int c = 0; c < 4; c++) ng[c] = -gcd_r2t[c];
  uint64_t* a = rows_r2t ? rows_r2t : (uint64_t*)malloc((size_t)h * 4 * sizeof(uint64_t));
  uint64_t* b = rows_t2r ? rows_t2r : (uint64_t*)malloc((size_t)h * 4 * sizeof(uint64_t));
  or_cps(p0, p1, w, h, bit_depth, gcd_r2t, R, cw_search, a, NULL, NULL);
  or_cps(p1, p0, w, h, bit_depth, ng, R, cw_search, b, NULL, NULL);
  double r2t = or_ivpsnr_quality(a, w, h, bit_depth, erp_w, cw_avg);
  double t2r = or_ivpsnr_quality(b, w, h, bit_depth, erp_w, cw_avg);
  if (!rows_r2t) free(a);
  if (!rows_t2r) free(b);
  if (r2t_out) *r2t_out = r2t;
  if (t2r_out) *t2r_out = t2r;
  return t2r < r2t ? t2r : r2t;
}

/* Shift-compensated pictures as the app generates them:
 * xAppQMIV.cpp:934-948 -> xShftCompPic.cpp:22-31,63-79.
 * scp1 = SCP from (Ref = P1, Tst = P0, +GCD); scp0 = SCP from (Ref = P0, Tst = P1, -GCD). */
OR_EXPORT void or_gen_scp(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                          int32_t bit_depth, const int32_t gcd_r2t[4], int32_t R, const int32_t cw_search[4],
                          uint16_t* const scp0[3], uint16_t* const scp1[3]) {
  int32_t ng[4];
  for (int c = 0; c < 4; c++) ng[c] = -gcd_r2t[c];
  or_cps(p0, p1, w, h, bit_depth, gcd_r2t, R, cw_search, NULL, scp1, NULL);
  or_cps(p1, p0, w, h, bit_depth, ng, R, cw_search, NULL, scp0, NULL);
}

/* ---------------------------------------------------------------------------
 * WS-PSNR per component: libs/pmbb-ivqm/src/xWSPSNR.cpp:81-107
 * (err = KBNS_y((double)RowSSD[y] * w[y]) * correction). Without ERP it is
 * plain PSNR (xWSPSNR.cpp:24-27 -> xPSNR.cpp:30-36). SSD of 0 yields DBL_MAX,
 * as the library does with fake values disabled.
 * ------------------------------------------------------------------------- */
OR_EXPORT void or_wspsnr(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                         int32_t bit_depth, int32_t erp, int32_t lat_range_deg, double out[3]) {
  uint64_t* rows = (uint64_t*)malloc((size_t)h * sizeof(uint64_t));
  double* wts = (double*)malloc((size_t)h * sizeof(double));
  double corr = 1.0;
  if (erp) or_ws_weights(h, lat_range_deg, wts, &corr);
  for (int c = 0; c < 3; c++) {
    or_row_ssd(p1[c], p0[c], w, h, rows);
    if (erp) {
      or_kbns k = {0, 0};
      for (int32_t y = 0; y < h; y++) or_kbns_acc(&k, (double)rows[y] * wts[y]);
      out[c] = or_psnr_from_ssd(or_kbns_result(&k) * corr, (int64_t)w * h, bit_depth);
    } else {
      uint64_t s = 0;
      for (int32_t y = 0; y < h; y++) s += rows[y];
      out[c] = or_psnr_from_ssd((double)s, (int64_t)w * h, bit_depth);
    }
  }
  free(rows);
  free(wts);
}

/* ---------------------------------------------------------------------------
 * Structural similarity.
 *
 * Window modes (libs/pmbb-ivqm/src/xSSIM.h:24-31):
 *   0 RegularGaussianFlt  11x11 centred, FP64 sums of float32 Gaussian weights
 *                         (xStructSimSTD.cpp:19-88)
 *   1 RegularGaussianInt  11x11 centred, int64 sums of Q18 weights
 *                         (xStructSimSTD.cpp:89-154)
 *   2 RegularAveraged     11x11 centred, plain int64 sums (xStructSimSTD.cpp:155-193)
 *   3 BlockGaussianInt    wnd x wnd (8 or 16) from the top-left, Q18 weights
 *                         (xStructSimSTD.cpp:353-419)
 *   4 BlockAveraged       wnd x wnd from the top-left, plain sums
 *                         (xStructSimSTD.cpp:420-461)
 * Integer moments are exact int64 here; the reference's STD kernels square in
 * int32 for the averaged and block-Gaussian modes, which differs only above
 * 15-bit samples (its SIMD kernels are limited to 15 bits by 16-bit madd).
 *
 * Border extension (libs/pmbb-core/src/xCommonDefCORE.h:322-331): regular
 * windows with an extension mode other than None run over every sample and
 * read the picture extended as xMarginOps.h:40-215 extends it (Constant uses
 * 65535, the value xPicP::extend passes: libs/pmbb-core/src/xPic.cpp:94-101).
 * ------------------------------------------------------------------------- */
enum { OR_RGLR_FLT = 0, OR_RGLR_INT = 1, OR_RGLR_AVG = 2, OR_BLCK_INT = 3, OR_BLCK_AVG = 4 };
enum { OR_EXT_NONE = 0, OR_EXT_NEAREST = 1, OR_EXT_REFLECT = 2, OR_EXT_MIRROR = 3, OR_EXT_CONSTANT = 4, OR_EXT_ZERO = 5 };

typedef struct {
  int32_t mode, ext, wnd, stride;
} or_ssim_cfg;

/* Window weights.
 * Float (xStructSimConsts.h:51-66, sigma c_Sigma = 1.5 at :35): the 11x11
 * Gaussian normalised to sum 1 in FP64, each weight rounded to float32.
 * Q18 integer tables (c_FltrIntShift = 18, :95-98): round(2^18 * normalised
 * Gaussian), plus one at the listed (row, column) positions --
 *   11x11 sigma 1.5 (:100-113), 8x8 sigma 1.5 (:134-144), 16x16 sigma 2.4
 *   (:146-164); 8x8 and 16x16 are centred between samples. */
static double or_w_flt11[121], or_w_int11[121], or_w_int8[64], or_w_int16[256];
static int or_w_ready = 0;

static void or_gauss(int32_t n, double sigma, double* g) {
  const double c = 0.5 * (double)(n - 1);
  double s = 0.0;
  for (int32_t i = 0; i < n; i++)
    for (int32_t j = 0; j < n; j++) {
      const double v = exp(-((i - c) * (i - c) + (j - c) * (j - c)) / (2.0 * sigma * sigma));
      g[i * n + j] = v;
      s += v;
    }
  for (int32_t k = 0; k < n * n; k++) g[k] /= s;
}
static void or_q18(int32_t n, double sigma, const int32_t* plus1, int32_t nplus, double* out) {
  or_gauss(n, sigma, out);
  for (int32_t k = 0; k < n * n; k++) out[k] = floor(out[k] * 262144.0 + 0.5);
  for (int32_t k = 0; k < nplus; k++) out[plus1[2 * k] * n + plus1[2 * k + 1]] += 1.0;
}
static void or_weights_init(void) {
  static const int32_t p11[] = {3, 4, 3, 6, 4, 3, 4, 7, 5, 4, 5, 6, 6, 3, 6, 7, 7, 4, 7, 6};
  static const int32_t p8[] = {2, 2, 2, 5, 5, 2, 5, 5};
  static const int32_t p16[] = {0, 3, 0, 12, 3, 0, 3, 15, 12, 0, 12, 15, 15, 3, 15, 12};
  if (or_w_ready) return;
  or_gauss(11, 1.5, or_w_flt11);
  for (int32_t k = 0; k < 121; k++) or_w_flt11[k] = (double)(float)or_w_flt11[k];
  or_q18(11, 1.5, p11, 10, or_w_int11);
  or_q18(8, 1.5, p8, 4, or_w_int8);
  or_q18(16, 2.4, p16, 8, or_w_int16);
  or_w_ready = 1;
}
/* which: 0 = float 11x11, 1 = Q18 11x11, 2 = Q18 8x8, 3 = Q18 16x16 */
OR_EXPORT void or_ssim_weights(int32_t which, double* out) {
  or_weights_init();
  const double* src = which == 0 ? or_w_flt11 : which == 1 ? or_w_int11 : which == 2 ? or_w_int8 : or_w_int16;
  const int32_t n = which == 2 ? 64 : which == 3 ? 256 : 121;
  memcpy(out, src, (size_t)n * sizeof(double));
}

/* Sample (x, y) of a w x h plane extended by `ext` (one fold for Reflect and
 * Mirror: overhangs up to the plane size). */
static inline int32_t or_ext_sample(const uint16_t* p, int32_t w, int32_t h, int32_t x, int32_t y, int32_t ext) {
  if (x < 0 || x >= w || y < 0 || y >= h) {
    switch (ext) {
      case OR_EXT_NEAREST:  /* ( a a a | a b c d | d d d ) */
        x = or_clampi(x, 0, w - 1);
        y = or_clampi(y, 0, h - 1);
        break;
      case OR_EXT_REFLECT:  /* ( c b a | a b c d | d c b ) */
        x = x < 0 ? -x - 1 : (x >= w ? 2 * w - 1 - x : x);
        y = y < 0 ? -y - 1 : (y >= h ? 2 * h - 1 - y : y);
        break;
      case OR_EXT_MIRROR:   /* ( d c b | a b c d | c b a ) */
        x = x < 0 ? -x : (x >= w ? 2 * w - 2 - x : x);
        y = y < 0 ? -y : (y >= h ? 2 * h - 2 - y : y);
        break;
      case OR_EXT_CONSTANT:
        return 65535;
      default:
        return 0;
    }
  }
  return p[(size_t)y * w + x];
}

static double or_ssim_ratio(double ar, double at, double vr, double vt, double cv, double c1, double c2,
                            int32_t calc_l) {
  const double cs = (2 * cv + c2) / (vr + vt + c2);
  if (calc_l) {
    const double l = (2 * ar * at + c1) / (ar * ar + at * at + c1);
    return l * cs;
  }
  return cs;
}

/* One window at (x, y): the top-left sample for block modes, the centre for
 * regular modes. */
static double or_ssim_win(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t x, int32_t y,
                          const or_ssim_cfg* p, double c1, double c2, int32_t calc_l) {
  const int32_t regular = p->mode <= OR_RGLR_AVG;
  const int32_t n = regular ? 11 : p->wnd;
  const int32_t x0 = regular ? x - 5 : x, y0 = regular ? y - 5 : y;
  if (p->mode == OR_RGLR_FLT) {
    double sr = 0, st = 0, srr = 0, stt = 0, srt = 0;
    for (int32_t dy = 0; dy < n; dy++)
      for (int32_t dx = 0; dx < n; dx++) {
        const double r = or_ext_sample(ref, w, h, x0 + dx, y0 + dy, p->ext);
        const double t = or_ext_sample(tst, w, h, x0 + dx, y0 + dy, p->ext);
        const double c = or_w_flt11[dy * 11 + dx];
        sr += r * c;
        st += t * c;
        srr += (r * r) * c;
        stt += (t * t) * c;
        srt += (r * t) * c;
      }
    return or_ssim_ratio(sr, st, srr - sr * sr, stt - st * st, srt - sr * st, c1, c2, calc_l);
  }
  const double* wt = p->mode == OR_RGLR_INT ? or_w_int11
                     : p->mode == OR_BLCK_INT ? (n == 8 ? or_w_int8 : or_w_int16) : NULL;
  int64_t sr = 0, st = 0, srr = 0, stt = 0, srt = 0;
  for (int32_t dy = 0; dy < n; dy++)
    for (int32_t dx = 0; dx < n; dx++) {
      const int64_t r = or_ext_sample(ref, w, h, x0 + dx, y0 + dy, p->ext);
      const int64_t t = or_ext_sample(tst, w, h, x0 + dx, y0 + dy, p->ext);
      const int64_t c = wt ? (int64_t)wt[dy * n + dx] : 1;
      sr += r * c;
      st += t * c;
      srr += r * r * c;
      stt += t * t * c;
      srt += r * t * c;
    }
  const double k = wt ? (double)1.0 / (double)262144 : (p->mode == OR_RGLR_AVG ? (double)1.0 / (double)121
                                                                                : (double)1.0 / (double)(n * n));
  const double ar = (double)sr * k, at = (double)st * k;
  return or_ssim_ratio(ar, at, (double)srr * k - ar * ar, (double)stt * k - at * at, (double)srt * k - ar * at, c1,
                       c2, calc_l);
}

/* Block-averaged window with explicit strides (kept for the tests' brute-force checks). */
OR_EXPORT double or_ssim_window(const uint16_t* tst, const uint16_t* ref, int32_t stride_t, int32_t stride_r,
                                int32_t wnd, double c1, double c2, int32_t calc_l) {
  int64_t sr = 0, st = 0, srr = 0, stt = 0, srt = 0;
  for (int32_t y = 0; y < wnd; y++) {
    for (int32_t x = 0; x < wnd; x++) {
      int64_t r = ref[x], t = tst[x];
      sr += r; st += t; srr += r * r; stt += t * t; srt += r * t;
    }
    ref += stride_r;
    tst += stride_t;
  }
  const double inv = (double)1.0 / (double)(wnd * wnd);
  const double ar = (double)sr * inv, at = (double)st * inv;
  return or_ssim_ratio(ar, at, (double)srr * inv - ar * ar, (double)stt * inv - at * at, (double)srt * inv - ar * at,
                       c1, c2, calc_l);
}

/* Number of windows along one axis: libs/pmbb-ivqm/src/xSSIM.cpp:381-404. */
OR_EXPORT int32_t or_ssim_num_blocks(int32_t len, int32_t wnd, int32_t stride) {
  int32_t cnt = 0;
  for (int32_t i = 0; i < len - wnd + 1; i += stride) cnt++;
  return cnt;
}
/* Number of MS-SSIM scales: libs/pmbb-ivqm/src/xSSIM.cpp:405-417. */
OR_EXPORT int32_t or_ms_num_scales(int32_t w, int32_t h) {
  for (int32_t i = 0; i < 5; i++) {
    const int32_t lim = 32 << i;
    if (w < lim || h < lim) return i;
  }
  return 5;
}

/* Window grid: xSSIM::xInitLoopRanges, libs/pmbb-ivqm/src/xSSIM.cpp:161-196.
 * Regular: centres from FilterRange (5) to size - 5, or over the whole plane
 * with a border extension; block: top-left corners 0 .. size - wnd. */
OR_EXPORT void or_ssim_grid(int32_t w, int32_t h, int32_t mode, int32_t ext, int32_t wnd, int32_t stride,
                            int32_t out[4]) {
  if (mode <= OR_RGLR_AVG) {
    const int32_t b = ext != OR_EXT_NONE ? 0 : 5;
    out[0] = b;
    out[1] = (w - 2 * b + stride - 1) / stride;
    out[2] = b;
    out[3] = (h - 2 * b + stride - 1) / stride;
  } else {
    out[0] = 0;
    out[1] = or_ssim_num_blocks(w, wnd, stride);
    out[2] = 0;
    out[3] = or_ssim_num_blocks(h, wnd, stride);
  }
}

/* Window-row sums of one plane: row loop of xSSIM::xCalcRowSSIM,
 * libs/pmbb-ivqm/src/xSSIM.cpp:238-254 (KBNS over x), task loop :219-231.
 * rowsums receives one value per window row (NumUnitY entries). */
OR_EXPORT void or_ssim_rowsums_ex(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t mode,
                                  int32_t ext, int32_t wnd, int32_t stride, double c1, double c2, int32_t calc_l,
                                  double* rowsums) {
  const or_ssim_cfg p = {mode, ext, wnd, stride};
  int32_t g[4];
  or_weights_init();
  or_ssim_grid(w, h, mode, ext, wnd, stride, g);
  /* window rows are independent: threads over rows change no result */
#pragma omp parallel for schedule(dynamic, 8)
  for (int32_t j = 0; j < g[3]; j++) {
    const int32_t y = g[2] + j * stride;
    or_kbns k = {0, 0};
    for (int32_t i = 0; i < g[1]; i++) or_kbns_acc(&k, or_ssim_win(tst, ref, w, h, g[0] + i * stride, y, &p, c1, c2, calc_l));
    rowsums[j] = or_kbns_result(&k);
  }
}
OR_EXPORT void or_ssim_rowsums(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t wnd,
                               int32_t stride, double c1, double c2, int32_t calc_l, double* rowsums) {
  or_ssim_rowsums_ex(tst, ref, w, h, OR_BLCK_AVG, OR_EXT_NONE, wnd, stride, c1, c2, calc_l, rowsums);
}

/* SSIM of one plane: xSSIM::xCalcCmpSSIM, libs/pmbb-ivqm/src/xSSIM.cpp:213-237
 * (KBNS over rows, / NumUnitY*NumUnitX). With erp_w, the IV-SSIM weighting of
 * xIVSSIM::xCalcCmpSSIM, libs/pmbb-ivqm/src/xIVSSIM.cpp:148-165: the row of
 * windows at y is scaled by w[y] (regular, y = centre) or w[y + wnd/2]
 * (block, y = top), the picture sum is divided by
 * NumActive * (KBNS(weights) / NumUnitY). */
OR_EXPORT double or_ssim_plane_ex(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t mode,
                                  int32_t ext, int32_t wnd, int32_t stride, double c1, double c2, int32_t calc_l,
                                  const double* erp_w) {
  int32_t g[4];
  or_ssim_grid(w, h, mode, ext, wnd, stride, g);
  const int32_t nx = g[1], ny = g[3];
  if (nx <= 0 || ny <= 0) return NAN;
  double* rs = (double*)malloc((size_t)ny * sizeof(double));
  or_ssim_rowsums_ex(tst, ref, w, h, mode, ext, wnd, stride, c1, c2, calc_l, rs);
  double res;
  if (erp_w) {
    or_kbns wk = {0, 0};
    for (int32_t j = 0; j < ny; j++) {
      const int32_t y = g[2] + j * stride;
      const double wt = erp_w[mode <= OR_RGLR_AVG ? y : y + (wnd >> 1)];
      rs[j] = rs[j] * wt;
      or_kbns_acc(&wk, wt);
    }
    const double corr = or_kbns_result(&wk) / ny;
    res = or_kbns_sum(rs, ny) / ((double)((int64_t)ny * nx) * corr);
  } else {
    res = or_kbns_sum(rs, ny) / (double)((int64_t)ny * nx);
  }
  free(rs);
  return res;
}
OR_EXPORT double or_ssim_plane(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t wnd,
                               int32_t stride, double c1, double c2, int32_t calc_l, const double* erp_w) {
  return or_ssim_plane_ex(tst, ref, w, h, OR_BLCK_AVG, OR_EXT_NONE, wnd, stride, c1, c2, calc_l, erp_w);
}

static void or_ssim_consts(int32_t bit_depth, double* c1, double* c2) {
  const double maxv = (double)((1 << bit_depth) - 1);
  *c1 = (0.01 * maxv) * (0.01 * maxv); /* xSSIM.cpp:50-52 */
  *c2 = (0.03 * maxv) * (0.03 * maxv);
}

/* Per-component SSIM of a pair (xSSIM::calcPicSSIM, xSSIM.h:83). */
OR_EXPORT void or_ssim_ex(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                          int32_t bit_depth, int32_t mode, int32_t ext, int32_t wnd, int32_t stride, double out[3]) {
  double c1, c2;
  or_ssim_consts(bit_depth, &c1, &c2);
  for (int c = 0; c < 3; c++) out[c] = or_ssim_plane_ex(p0[c], p1[c], w, h, mode, ext, wnd, stride, c1, c2, 1, NULL);
}
OR_EXPORT void or_ssim(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                       int32_t bit_depth, int32_t wnd, int32_t stride, double out[3]) {
  or_ssim_ex(p0, p1, w, h, bit_depth, OR_BLCK_AVG, OR_EXT_NONE, wnd, stride, out);
}

/* Multi-scale exponents: libs/pmbb-ivqm/src/xStructSimConsts.h:42-49. */
static const double or_ms_exp[5][5] = {
    {1.0000, 0.0000, 0.0000, 0.0000, 0.0000}, {0.1356, 0.8644, 0.0000, 0.0000, 0.0000},
    {0.0711, 0.4530, 0.4760, 0.0000, 0.0000}, {0.0517, 0.3295, 0.3462, 0.2726, 0.0000},
    {0.0448, 0.2856, 0.3001, 0.2363, 0.1333}};

/* MS-SSIM per component: libs/pmbb-ivqm/src/xSSIM.cpp:115-141. Scale 0 on the
 * full picture, scale i on the 2x2 rounded-mean pyramid level i (extended by
 * the same border mode, :128); luminance only at the last scale; product of
 * ReLU(score_i)^exp_i. No ERP weighting. */
OR_EXPORT void or_msssim_ex(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                            int32_t bit_depth, int32_t mode, int32_t ext, int32_t wnd, int32_t stride,
                            double out[3]) {
  double c1, c2;
  or_ssim_consts(bit_depth, &c1, &c2);
  const int32_t ns = or_ms_num_scales(w, h);
  for (int c = 0; c < 3; c++) {
    double sc[5] = {0, 0, 0, 0, 0};
    const uint16_t* a = p0[c];
    const uint16_t* b = p1[c];
    uint16_t *la = NULL, *lb = NULL;
    int32_t cw = w, ch = h;
    for (int32_t i = 0; i < ns; i++) {
      if (i > 0) {
        uint16_t* na = (uint16_t*)malloc((size_t)(cw >> 1) * (ch >> 1) * 2 + 2);
        uint16_t* nb = (uint16_t*)malloc((size_t)(cw >> 1) * (ch >> 1) * 2 + 2);
        or_downsample_hv(na, a, cw, ch);
        or_downsample_hv(nb, b, cw, ch);
        free(la); free(lb);
        la = na; lb = nb; a = na; b = nb;
        cw >>= 1; ch >>= 1;
      }
      sc[i] = or_ssim_plane_ex(a, b, cw, ch, mode, ext, wnd, stride, c1, c2, i == ns - 1, NULL);
    }
    free(la); free(lb);
    double v = 1.0;
    for (int32_t i = 0; i < ns; i++) v *= pow(sc[i] > 0 ? sc[i] : 0.0, or_ms_exp[ns - 1][i]);
    out[c] = v;
  }
}
OR_EXPORT void or_msssim(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                         int32_t bit_depth, int32_t wnd, int32_t stride, double out[3]) {
  or_msssim_ex(p0, p1, w, h, bit_depth, OR_BLCK_AVG, OR_EXT_NONE, wnd, stride, out);
}

/* IV-SSIM: libs/pmbb-ivqm/src/xIVSSIM.cpp:31-61 as called from
 * xAppQMIV.cpp:1074-1078: T2R = SSIM(P0, SCP1), R2T = SSIM(P1, SCP0), each the
 * component-weighted mean, result = min. ERP weights apply when erp_w != NULL. */
OR_EXPORT double or_ivssim_ex(const uint16_t* const p0[3], const uint16_t* const p1[3],
                              const uint16_t* const scp0[3], const uint16_t* const scp1[3], int32_t w, int32_t h,
                              int32_t bit_depth, int32_t mode, int32_t ext, int32_t wnd, int32_t stride,
                              const int32_t cw_avg[4], const double* erp_w, double* r2t_out, double* t2r_out) {
  double c1, c2;
  or_ssim_consts(bit_depth, &c1, &c2);
  double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++) {
    a[c] = or_ssim_plane_ex(p0[c], scp1[c], w, h, mode, ext, wnd, stride, c1, c2, 1, erp_w);
    b[c] = or_ssim_plane_ex(p1[c], scp0[c], w, h, mode, ext, wnd, stride, c1, c2, 1, erp_w);
  }
  const double t2r = or_weighted_mean4(a, cw_avg), r2t = or_weighted_mean4(b, cw_avg);
  if (r2t_out) *r2t_out = r2t;
  if (t2r_out) *t2r_out = t2r;
  return r2t < t2r ? r2t : t2r;
}
OR_EXPORT double or_ivssim(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* const scp0[3],
                           const uint16_t* const scp1[3], int32_t w, int32_t h, int32_t bit_depth, int32_t wnd,
                           int32_t stride, const int32_t cw_avg[4], const double* erp_w, double* r2t_out,
                           double* t2r_out) {
  return or_ivssim_ex(p0, p1, scp0, scp1, w, h, bit_depth, OR_BLCK_AVG, OR_EXT_NONE, wnd, stride, cw_avg, erp_w,
                      r2t_out, t2r_out);
}

/* IV-MS-SSIM: libs/pmbb-ivqm/src/xIVSSIM.cpp:62-92 (MS-SSIM of (P0, SCP1) and
 * (P1, SCP0), weighted means, min; no ERP weighting). */
OR_EXPORT double or_ivmsssim_ex(const uint16_t* const p0[3], const uint16_t* const p1[3],
                                const uint16_t* const scp0[3], const uint16_t* const scp1[3], int32_t w, int32_t h,
                                int32_t bit_depth, int32_t mode, int32_t ext, int32_t wnd, int32_t stride,
                                const int32_t cw_avg[4]) {
  double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
  or_msssim_ex(p0, scp1, w, h, bit_depth, mode, ext, wnd, stride, a);
  or_msssim_ex(p1, scp0, w, h, bit_depth, mode, ext, wnd, stride, b);
  const double t2r = or_weighted_mean4(a, cw_avg), r2t = or_weighted_mean4(b, cw_avg);
  return r2t < t2r ? r2t : t2r;
}
OR_EXPORT double or_ivmsssim(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* const scp0[3],
                             const uint16_t* const scp1[3], int32_t w, int32_t h, int32_t bit_depth, int32_t wnd,
                             int32_t stride, const int32_t cw_avg[4]) {
  return or_ivmsssim_ex(p0, p1, scp0, scp1, w, h, bit_depth, OR_BLCK_AVG, OR_EXT_NONE, wnd, stride, cw_avg);
}

/* ---------------------------------------------------------------------------
 * Mask mode (apps/QMIV/src/xAppQMIV.cpp:902-906, :930, :953, :971, :999,
 * :1026-1029). The mask is a luma-resolution weight plane M (0..MaxM,
 * BitDepthM 8..14); chroma of 4:2:0 pictures is compared at 4:4:4, so every
 * sum below runs over the 4:4:4 planes with M at the same position.
 * ------------------------------------------------------------------------- */
/* xPixelOpsSTD::CountNonZero (libs/pmbb-core/src/xPixelOpsSTD.cpp). */
OR_EXPORT int64_t or_count_nonzero(const uint16_t* m, int64_t n) {
  int64_t k = 0;
  for (int64_t i = 0; i < n; i++) k += m[i] != 0;
  return k;
}
/* xDistortionSTD::CalcWeightedSD (libs/pmbb-core/src/xDistortionSTD.cpp:45-56):
 * sum of ((int32)T - (int32)R) * (int32)M, the product in int32. */
OR_EXPORT int64_t or_weighted_sd(const uint16_t* tst, const uint16_t* ref, const uint16_t* msk, int64_t n) {
  int64_t sd = 0;
  for (int64_t i = 0; i < n; i++) sd += (int32_t)(((int32_t)tst[i] - (int32_t)ref[i]) * (int32_t)msk[i]);
  return sd;
}
/* xDistortionSTD::CalcWeightedSSD (:57-68): sum of (uint64)(int32 square) * M;
 * the square is int32 (wraps for |d| > 46340) and sign-extends to uint64. */
static uint64_t or_wssd_term(int32_t d, uint16_t m) {
  const int32_t sq = (int32_t)((uint32_t)d * (uint32_t)d);
  return (uint64_t)(int64_t)sq * (uint64_t)m;
}
OR_EXPORT uint64_t or_weighted_ssd(const uint16_t* tst, const uint16_t* ref, const uint16_t* msk, int64_t n) {
  uint64_t s = 0;
  for (int64_t i = 0; i < n; i++) s += or_wssd_term((int32_t)tst[i] - (int32_t)ref[i], msk[i]);
  return s;
}

/* xGlobClrDiff::CalcGlobalColorDiffM (libs/pmbb-ivqm/src/xGlobClrDiff.cpp:60-91)
 * as xGlobClrDiffProc calls it (xGlobClrDiff.h:42 swaps the pictures):
 * SD_c = sum (P1 - P0) * M, avg = SD / (NumNonMasked * MaxM), round half
 * away, clip to the FP32 MaxDiff. */
OR_EXPORT void or_gcd_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk, int32_t w,
                             int32_t h, int32_t bit_depth, int32_t bit_depth_m, const float coef[4], int32_t gcd[4]) {
  const int64_t n = (int64_t)w * h;
  const int64_t nnm = or_count_nonzero(msk, n);
  const int32_t maxv = (1 << bit_depth) - 1, maxm = (1 << bit_depth_m) - 1;
  for (int c = 0; c < 4; c++) {
    const int32_t maxdiff = or_round_f32(coef[c] * (float)maxv);
    const int64_t sd = c < 3 ? or_weighted_sd(p1[c], p0[c], msk, n) : 0;
    const double avg = (double)sd / (double)(nnm * (int64_t)maxm);
    int32_t g = or_round_f64(avg);
    gcd[c] = or_clampi(g, -maxdiff, maxdiff);
  }
}

/* Masked PSNR family as the app computes it:
 *   MSE_c  = SSDM_c / (NNM * MaxM)                        (xPSNR.cpp:153-163)
 *   PSNR_c = 10 log10(NNM * MaxP^2 * MaxM / SSDM_c), DBL_MAX for SSDM 0
 *            (xPSNR.cpp:164-179) where calcPicPSNRM passes the MASK bit depth
 *            as the picture's (xPSNR.cpp:60: MaxP = MaxM), and so does
 *            WS-PSNR without ERP (xWSPSNR.cpp:61-64 calls calcPicPSNRM);
 *   WS-PSNR with ERP: rows of weighted SSD, KBNS of row * w[y], times the
 *            correction, then the same formula with the picture's own bit
 *            depth (xWSPSNR.cpp:108-136).
 * SSDM_c = sum (P0 - P1)^2 * M. psnr_bd is the bit depth used for MaxP. */
OR_EXPORT double or_psnr_masked(double ssdm, int64_t nnm, int32_t psnr_bd, int32_t bit_depth_m) {
  const int64_t maxp = (1 << psnr_bd) - 1, maxm = (1 << bit_depth_m) - 1;
  const int64_t mx = nnm * (maxp * maxp) * maxm;
  return ssdm > 0 ? 10.0 * log10((double)mx / ssdm) : DBL_MAX;
}
OR_EXPORT void or_psnr_family_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk,
                                     int32_t w, int32_t h, int32_t bit_depth, int32_t bit_depth_m, const double* erp_w,
                                     double correction, double mse[3], double psnr[3], double wspsnr[3]) {
  const int64_t n = (int64_t)w * h;
  const int64_t nnm = or_count_nonzero(msk, n);
  const int64_t maxm = (1 << bit_depth_m) - 1;
  for (int c = 0; c < 3; c++) {
    const uint64_t ssdm = or_weighted_ssd(p0[c], p1[c], msk, n);
    mse[c] = (double)ssdm / (double)(nnm * maxm);
    psnr[c] = or_psnr_masked((double)ssdm, nnm, bit_depth_m, bit_depth_m);
    if (erp_w) {
      or_kbns k = {0, 0};
      for (int32_t y = 0; y < h; y++) {
        const size_t o = (size_t)y * w;
        or_kbns_acc(&k, (double)or_weighted_ssd(p1[c] + o, p0[c] + o, msk + o, w) * erp_w[y]);
      }
      wspsnr[c] = or_psnr_masked(or_kbns_result(&k) * correction, nnm, bit_depth, bit_depth_m);
    } else {
      wspsnr[c] = psnr[c];
    }
  }
}

/* Masked corresponding-pixel search of one direction:
 * xCorrespPixelShiftSTD::CalcDistAsymmetricRowM / FindBestPixelWithinBlockM
 * (libs/pmbb-ivqm/src/xCorrespPixelShiftSTD.cpp:151-215). Test pixels with
 * M = 0 are skipped; candidates with M = 0 are skipped (M outside the
 * picture is the Nearest-extended mask, xAppQMIV.cpp:899); the winner is the
 * first strict minimum of the int32 (wrapping) weighted error; the row sum is
 * uint64 of (uint64)(int32 square) * M per component. */
static void or_masked_best(const uint16_t* const ref[3], const uint16_t* msk, int32_t w, int32_t h, int32_t x,
                           int32_t y, const int32_t t[3], int32_t R, const int32_t cw[4], int32_t* bx, int32_t* by) {
  int32_t best_err = INT32_MAX;
  *bx = -1;
  *by = -1;
  for (int32_t dy = -R; dy <= R; dy++) {
    const int32_t ry = or_clampi(y + dy, 0, h - 1);
    for (int32_t dx = -R; dx <= R; dx++) {
      const int32_t rx = or_clampi(x + dx, 0, w - 1);
      if (msk[(size_t)ry * w + rx] == 0) continue;
      uint32_t e = 0;
      for (int c = 0; c < 3; c++) {
        const uint32_t d = (uint32_t)(t[c] - (int32_t)ref[c][(size_t)ry * w + rx]);
        e += d * d * (uint32_t)cw[c];
      }
      if ((int32_t)e < best_err) { best_err = (int32_t)e; *bx = rx; *by = ry; }
    }
  }
}

OR_EXPORT void or_cps_masked(const uint16_t* const tst[3], const uint16_t* const ref[3], const uint16_t* msk,
                             int32_t w, int32_t h, const int32_t gcs[4], int32_t R, const int32_t cw[4],
                             uint64_t* rows) {
  for (int32_t y = 0; y < h; y++) {
    uint64_t row[4] = {0, 0, 0, 0};
    for (int32_t x = 0; x < w; x++) {
      const uint16_t m = msk[(size_t)y * w + x];
      if (m == 0) continue;
      int32_t t[3], bx, by;
      for (int c = 0; c < 3; c++) t[c] = (int32_t)tst[c][(size_t)y * w + x] + gcs[c];
      or_masked_best(ref, msk, w, h, x, y, t, R, cw, &bx, &by);
      for (int c = 0; c < 3; c++) row[c] += or_wssd_term(t[c] - (int32_t)ref[c][(size_t)by * w + bx], m);
    }
    for (int c = 0; c < 4; c++) rows[(size_t)y * 4 + c] = row[c];
  }
}

/* Shift-compensated pictures with a mask: xShftCompPic::GenShftCompPicsM
 * (libs/pmbb-ivqm/src/xShftCompPic.cpp:32-42, :80-117) ->
 * xCorrespPixelShiftSTD::GenShftCompRowM (xCorrespPixelShiftSTD.cpp:268-291):
 * for every pixel with M != 0, the masked search's winner in Ref minus the
 * shift, clipped to [0, Max]; pixels with M = 0 are left as they are.
 * scp1 = from (Ref = P1, Tst = P0, +GCD); scp0 = from (Ref = P0, Tst = P1, -GCD). */
static void or_scp_masked_dir(const uint16_t* const tst[3], const uint16_t* const ref[3], const uint16_t* msk,
                              int32_t w, int32_t h, int32_t bit_depth, const int32_t gcs[4], int32_t R,
                              const int32_t cw[4], uint16_t* const dst[3]) {
  const int32_t maxv = (1 << bit_depth) - 1;
  for (int32_t y = 0; y < h; y++)
    for (int32_t x = 0; x < w; x++) {
      if (msk[(size_t)y * w + x] == 0) continue;
      int32_t t[3], bx, by;
      for (int c = 0; c < 3; c++) t[c] = (int32_t)tst[c][(size_t)y * w + x] + gcs[c];
      or_masked_best(ref, msk, w, h, x, y, t, R, cw, &bx, &by);
      for (int c = 0; c < 3; c++)
        dst[c][(size_t)y * w + x] = (uint16_t)or_clampi((int32_t)ref[c][(size_t)by * w + bx] - gcs[c], 0, maxv);
    }
}

OR_EXPORT void or_gen_scp_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk,
                                 int32_t w, int32_t h, int32_t bit_depth, const int32_t gcd_r2t[4], int32_t R,
                                 const int32_t cw_search[4], uint16_t* const scp0[3], uint16_t* const scp1[3]) {
  int32_t ng[4];
  for (int c = 0; c < 4; c++) ng[c] = -gcd_r2t[c];
  or_scp_masked_dir(p0, p1, msk, w, h, bit_depth, gcd_r2t, R, cw_search, scp1);
  or_scp_masked_dir(p1, p0, msk, w, h, bit_depth, ng, R, cw_search, scp0);
}

/* IV-PSNR with a mask: xIVPSNRM::calcPicIVPSNRM(TstI, RefI, Msk, NNM, GCD_R2T)
 * (libs/pmbb-ivqm/src/xIVPSNR.cpp:152-208). Its labels are swapped against
 * the unmasked calcPicIVPSNR (:45-55): here R2T is P1 - GCD searched in P0
 * and T2R is P0 + GCD searched in P1 (:161-162). The component error is the
 * uint64 row total (or, ERP, the 4-lane KBNS of (double)row * w[y], no
 * correction); PSNR_c = 10 log10(NNM * MaxP^2 * MaxM / max(err, ->1 if 0)). */
OR_EXPORT double or_ivpsnr_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk,
                                  int32_t w, int32_t h, int32_t bit_depth, int32_t bit_depth_m,
                                  const int32_t gcd_r2t[4], int32_t R, const int32_t cw_search[4],
                                  const int32_t cw_avg[4], const double* erp_w, double* r2t_out, double* t2r_out) {
  const int64_t nnm = or_count_nonzero(msk, (int64_t)w * h);
  int32_t ng[4];
  for (int c = 0; c < 4; c++) ng[c] = -gcd_r2t[c];
  uint64_t* rows = (uint64_t*)malloc((size_t)h * 4 * sizeof(uint64_t));
  double q[2];
  for (int dir = 0; dir < 2; dir++) {
    /* dir 0 = R2T: Tst = P0 (+GCD) against Ref = P1; dir 1 = T2R: P1 (-GCD) against P0 */
    if (dir == 0) or_cps_masked(p0, p1, msk, w, h, gcd_r2t, R, cw_search, rows);
    else or_cps_masked(p1, p0, msk, w, h, ng, R, cw_search, rows);
    double err[4] = {0, 0, 0, 0};
    if (erp_w) {
      or_kbns k[4] = {{0, 0}, {0, 0}, {0, 0}, {0, 0}};
      for (int32_t y = 0; y < h; y++)
        for (int c = 0; c < 4; c++) or_kbns_acc(&k[c], (double)rows[(size_t)y * 4 + c] * erp_w[y]);
      for (int c = 0; c < 4; c++) err[c] = or_kbns_result(&k[c]);
    } else {
      uint64_t tot[4] = {0, 0, 0, 0};
      for (int32_t y = 0; y < h; y++)
        for (int c = 0; c < 4; c++) tot[c] += rows[(size_t)y * 4 + c];
      for (int c = 0; c < 4; c++) err[c] = (double)tot[c];
    }
    double v[4] = {0, 0, 0, 0};
    for (int c = 0; c < 3; c++) v[c] = or_psnr_masked(err[c] > 0 ? err[c] : 1.0, nnm, bit_depth, bit_depth_m);
    q[dir] = or_weighted_mean4(v, cw_avg);
  }
  free(rows);
  if (r2t_out) *r2t_out = q[1];
  if (t2r_out) *t2r_out = q[0];
  return q[1] < q[0] ? q[1] : q[0];
}

/* ---------------------------------------------------------------------------
 * Colour-space conversion, fixed point with 16 fractional bits, as the
 * reference dispatches it (xColorSpace.h: the _I32 kernels):
 * xColorSpaceSTD::ConvertRGB2YCbCr_I32 / ConvertYCbCr2RGB_I32
 * (libs/pmbb-core/src/xColorSpaceSTD.cpp), coefficients
 * xColorSpaceCoeffYCbCr::c_RGB2YCbCr_I32 / c_YCbCr2RGB_I32
 * (xColorSpaceCoeff.h:59-73) for BT.601 (= SMPTE170M, JPEG), BT.709,
 * SMPTE240M, BT.2020. The int32 arithmetic wraps for 16-bit samples (the
 * products and iy << 16 exceed 2^31); it is computed here in uint32 and
 * shifted as int32, which is what the reference's code does on x86.
 * Planes are converted in place; the pictures are 4:4:4.
 * ------------------------------------------------------------------------- */
/* [space][row][col]: RGB -> YCbCr rows Y, Cb, Cr over R, G, B. */
static const int32_t or_rgb2ycc[4][3][3] = {
    {{19595, 38470, 7471}, {-11058, -21710, 32768}, {32768, -27439, -5329}},  /* BT.601 */
    {{13933, 46871, 4732}, {-7509, -25259, 32768}, {32768, -29763, -3005}},   /* BT.709 */
    {{13894, 45941, 5701}, {-7609, -25159, 32768}, {32768, -29150, -3618}},   /* SMPTE240M */
    {{17216, 44434, 3886}, {-9151, -23617, 32768}, {32768, -30133, -2635}}};  /* BT.2020 */
/* YCbCr -> RGB: R = Y + rv Cr, G = Y + gu Cb + gv Cr, B = Y + bu Cb; {rv, gu, gv, bu}. */
static const int32_t or_ycc2rgb[4][4] = {
    {91881, -22553, -46802, 116130}, {103206, -12276, -30679, 121609},
    {103285, -14852, -31236, 119669}, {96639, -10784, -37444, 123299}};

static inline int32_t or_w32(uint32_t v) { return (int32_t)v; }
static inline int32_t or_clipu(int32_t v, int32_t mx) { return v < 0 ? 0 : (v > mx ? mx : v); }

OR_EXPORT void or_rgb2ycbcr(uint16_t* c0, uint16_t* c1, uint16_t* c2, int64_t n, int32_t bit_depth, int32_t space) {
  const int32_t(*k)[3] = or_rgb2ycc[space];
  const int32_t mid = 1 << (bit_depth - 1), mx = (1 << bit_depth) - 1;
  for (int64_t i = 0; i < n; i++) {
    const uint32_t r = c0[i], g = c1[i], b = c2[i];
    const int32_t ty = or_w32((uint32_t)k[0][0] * r + (uint32_t)k[0][1] * g + (uint32_t)k[0][2] * b + 32768u) >> 16;
    const int32_t tu = or_w32((uint32_t)k[1][0] * r + (uint32_t)k[1][1] * g + (b << 15) + 32768u) >> 16;
    const int32_t tv = or_w32((r << 15) + (uint32_t)k[2][1] * g + (uint32_t)k[2][2] * b + 32768u) >> 16;
    c0[i] = (uint16_t)or_clipu(ty, mx);
    c1[i] = (uint16_t)or_clipu(tu + mid, mx);
    c2[i] = (uint16_t)or_clipu(tv + mid, mx);
  }
}
OR_EXPORT void or_ycbcr2rgb(uint16_t* c0, uint16_t* c1, uint16_t* c2, int64_t n, int32_t bit_depth, int32_t space) {
  const int32_t* k = or_ycc2rgb[space];
  const int32_t mid = 1 << (bit_depth - 1), mx = (1 << bit_depth) - 1;
  for (int64_t i = 0; i < n; i++) {
    const int32_t iy = c0[i], iu = (int32_t)c1[i] - mid, iv = (int32_t)c2[i] - mid;
    const uint32_t sy = ((uint32_t)iy << 16) + 32768u;
    const int32_t r = or_w32(sy + (uint32_t)k[0] * (uint32_t)iv) >> 16;
    const int32_t g = or_w32(sy + (uint32_t)k[1] * (uint32_t)iu + (uint32_t)k[2] * (uint32_t)iv) >> 16;
    const int32_t b = or_w32(sy + (uint32_t)k[3] * (uint32_t)iu) >> 16;
    c0[i] = (uint16_t)or_clipu(r, mx);
    c1[i] = (uint16_t)or_clipu(g, mx);
    c2[i] = (uint16_t)or_clipu(b, mx);
  }
}

/* Component-weighted mean exported for the picture value of per-component
 * metrics (xMetricQMIV.cpp:91-97 uses v0*w0 + v1*w1 + v2*w2). */
OR_EXPORT double or_pic_value(const double v[3], const int32_t cw[4]) {
  const double inv = 1.0 / (double)(cw[0] + cw[1] + cw[2] + cw[3]);
  return (v[0] * cw[0] + v[1] * cw[1] + v[2] * cw[2]) * inv;
}
