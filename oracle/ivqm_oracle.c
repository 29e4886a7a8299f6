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
 * Block-averaged SSIM window: libs/pmbb-ivqm/src/xStructSimSTD.cpp:420-461,
 * with exact int64 moments (the FP64/int64 oracle; the reference's own STD
 * kernel squares in int32, which differs only above 15-bit samples).
 * ------------------------------------------------------------------------- */
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
  double avg_r = (double)sr * inv;
  double avg_t = (double)st * inv;
  double var_r = (double)srr * inv - avg_r * avg_r;
  double var_t = (double)stt * inv - avg_t * avg_t;
  double cov = (double)srt * inv - avg_r * avg_t;
  double cs = (2 * cov + c2) / (var_r + var_t + c2);
  if (calc_l) {
    double l = (2 * avg_r * avg_t + c1) / (avg_r * avg_r + avg_t * avg_t + c1);
    return l * cs;
  }
  return cs;
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

/* Window-row sums of one plane: row loop of xSSIM::xCalcRowSSIM,
 * libs/pmbb-ivqm/src/xSSIM.cpp:238-254 (KBNS over x), task loop :219-231.
 * rowsums receives one value per window row (NumUnitY entries). */
OR_EXPORT void or_ssim_rowsums(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t wnd,
                               int32_t stride, double c1, double c2, int32_t calc_l, double* rowsums) {
  const int32_t ny = h < wnd ? 0 : (h - wnd) / stride + 1;
  /* window rows are independent: threads over rows change no result */
#pragma omp parallel for schedule(dynamic, 8)
  for (int32_t j = 0; j < ny; j++) {
    const int32_t y = j * stride;
    or_kbns k = {0, 0};
    for (int32_t x = 0; x < w - wnd + 1; x += stride)
      or_kbns_acc(&k, or_ssim_window(tst + (size_t)y * w + x, ref + (size_t)y * w + x, w, w, wnd, c1, c2, calc_l));
    rowsums[j] = or_kbns_result(&k);
  }
}

/* SSIM of one plane: xSSIM::xCalcCmpSSIM, libs/pmbb-ivqm/src/xSSIM.cpp:213-237
 * (KBNS over rows, / NumUnitY*NumUnitX). With erp_w, the IV-SSIM weighting of
 * xIVSSIM::xCalcCmpSSIM, libs/pmbb-ivqm/src/xIVSSIM.cpp:148-165: row j is
 * scaled by w[y_j + wnd/2], the picture sum is divided by
 * NumActive * (KBNS(weights) / NumUnitY). */
OR_EXPORT double or_ssim_plane(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t wnd,
                               int32_t stride, double c1, double c2, int32_t calc_l, const double* erp_w) {
  const int32_t nx = or_ssim_num_blocks(w, wnd, stride), ny = or_ssim_num_blocks(h, wnd, stride);
  if (nx <= 0 || ny <= 0) return NAN;
  double* rs = (double*)malloc((size_t)ny * sizeof(double));
  or_ssim_rowsums(tst, ref, w, h, wnd, stride, c1, c2, calc_l, rs);
  double res;
  if (erp_w) {
    or_kbns wk = {0, 0};
    for (int32_t j = 0; j < ny; j++) {
      const double wt = erp_w[j * stride + (wnd >> 1)];
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

static void or_ssim_consts(int32_t bit_depth, double* c1, double* c2) {
  const double maxv = (double)((1 << bit_depth) - 1);
  *c1 = (0.01 * maxv) * (0.01 * maxv); /* xSSIM.cpp:50-52 */
  *c2 = (0.03 * maxv) * (0.03 * maxv);
}

/* Per-component SSIM of a pair (xSSIM::calcPicSSIM, xSSIM.h:83). */
OR_EXPORT void or_ssim(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                       int32_t bit_depth, int32_t wnd, int32_t stride, double out[3]) {
  double c1, c2;
  or_ssim_consts(bit_depth, &c1, &c2);
  for (int c = 0; c < 3; c++) out[c] = or_ssim_plane(p0[c], p1[c], w, h, wnd, stride, c1, c2, 1, NULL);
}

/* Multi-scale exponents: libs/pmbb-ivqm/src/xStructSimConsts.h:42-49. */
static const double or_ms_exp[5][5] = {
    {1.0000, 0.0000, 0.0000, 0.0000, 0.0000}, {0.1356, 0.8644, 0.0000, 0.0000, 0.0000},
    {0.0711, 0.4530, 0.4760, 0.0000, 0.0000}, {0.0517, 0.3295, 0.3462, 0.2726, 0.0000},
    {0.0448, 0.2856, 0.3001, 0.2363, 0.1333}};

/* MS-SSIM per component: libs/pmbb-ivqm/src/xSSIM.cpp:115-141. Scale 0 on the
 * full picture, scale i on the 2x2 rounded-mean pyramid level i; luminance only
 * at the last scale; product of ReLU(score_i)^exp_i. No ERP weighting. */
OR_EXPORT void or_msssim(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h,
                         int32_t bit_depth, int32_t wnd, int32_t stride, double out[3]) {
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
      sc[i] = or_ssim_plane(a, b, cw, ch, wnd, stride, c1, c2, i == ns - 1, NULL);
    }
    free(la); free(lb);
    double v = 1.0;
    for (int32_t i = 0; i < ns; i++) v *= pow(sc[i] > 0 ? sc[i] : 0.0, or_ms_exp[ns - 1][i]);
    out[c] = v;
  }
}

/* IV-SSIM: libs/pmbb-ivqm/src/xIVSSIM.cpp:31-61 as called from
 * xAppQMIV.cpp:1074-1078: T2R = SSIM(P0, SCP1), R2T = SSIM(P1, SCP0), each the
 * component-weighted mean, result = min. ERP weights apply when erp_w != NULL. */
OR_EXPORT double or_ivssim(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* const scp0[3],
                           const uint16_t* const scp1[3], int32_t w, int32_t h, int32_t bit_depth, int32_t wnd,
                           int32_t stride, const int32_t cw_avg[4], const double* erp_w, double* r2t_out,
                           double* t2r_out) {
  double c1, c2;
  or_ssim_consts(bit_depth, &c1, &c2);
  double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++) {
    a[c] = or_ssim_plane(p0[c], scp1[c], w, h, wnd, stride, c1, c2, 1, erp_w);
    b[c] = or_ssim_plane(p1[c], scp0[c], w, h, wnd, stride, c1, c2, 1, erp_w);
  }
  const double t2r = or_weighted_mean4(a, cw_avg), r2t = or_weighted_mean4(b, cw_avg);
  if (r2t_out) *r2t_out = r2t;
  if (t2r_out) *t2r_out = t2r;
  return r2t < t2r ? r2t : t2r;
}

/* IV-MS-SSIM: libs/pmbb-ivqm/src/xIVSSIM.cpp:62-92 (MS-SSIM of (P0, SCP1) and
 * (P1, SCP0), weighted means, min; no ERP weighting). */
OR_EXPORT double or_ivmsssim(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* const scp0[3],
                             const uint16_t* const scp1[3], int32_t w, int32_t h, int32_t bit_depth, int32_t wnd,
                             int32_t stride, const int32_t cw_avg[4]) {
  double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
  or_msssim(p0, scp1, w, h, bit_depth, wnd, stride, a);
  or_msssim(p1, scp0, w, h, bit_depth, wnd, stride, b);
  const double t2r = or_weighted_mean4(a, cw_avg), r2t = or_weighted_mean4(b, cw_avg);
  return r2t < t2r ? r2t : t2r;
}

/* Component-weighted mean exported for the picture value of per-component
 * metrics (xMetricQMIV.cpp:91-97 uses v0*w0 + v1*w1 + v2*w2). */
OR_EXPORT double or_pic_value(const double v[3], const int32_t cw[4]) {
  const double inv = 1.0 / (double)(cw[0] + cw[1] + cw[2] + cw[3]);
  return (v[0] * cw[0] + v[1] * cw[1] + v[2] * cw[2]) * inv;
}
