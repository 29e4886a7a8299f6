// Host FP64 epilogue of the metric hot path.
//
// The device produces exact integers (row distances, row SSDs, SD totals) and
// FP64 SSIM totals; the floating-point steps that decide the printed dB
// digits run here, on the host, in the reference's operation order, with the
// host libm (log10, cos, pow), without contraction, so PSNR, WS-PSNR and
// IV-PSNR match the reference bit for bit.
#pragma once
#include <float.h>
#include <math.h>
#include <stdint.h>

#include <vector>

#include "../../include/ivqm_hip.h"

#pragma clang fp contract(off)

namespace ivqm {
namespace host {

// Kahan-Babuska-Neumaier: libs/pmbb-core/src/xKBNS.h:36-52 (and the SSE / AVX
// twins in xKBNS.cpp:14-26, :85-96, which compute the same values).
struct Kbns {
  double s = 0.0, c = 0.0;
  void acc(double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += ((s - t) + v);
    else c += ((v - t) + s);
    s = t;
  }
  double result() const { return s + c; }
};

// libs/pmbb-core/src/xHelpersFLT.h:22 (value + copysign(0.5f, value), truncate).
inline int32_t round_f32(float v) {
  const float a = v + copysignf(0.5f, v);
  return (int32_t)a;
}

// libs/pmbb-ivqm/src/xPSNR.cpp:186-192.
inline double psnr_from_ssd(double ssd, int64_t area, int bit_depth) {
  const uint64_t maxv = (uint64_t)((1 << bit_depth) - 1);
  const uint64_t MAX = (uint64_t)area * (maxv * maxv);
  return ssd > 0 ? 10.0 * log10((double)MAX / ssd) : DBL_MAX;
}

// libs/pmbb-ivqm/src/xWeightedSpherically.cpp:13-29: h row weights into w; returns the correction
// factor h / sum(w).
inline double ws_weights_into(int h, int lat_range_deg, double* w) {
  const double eh = 180.0 * (double)h / (double)lat_range_deg;
  const double eo = (eh - h) / 2.0;
  const double pi = 3.14159265358979323846;
  for (int y = 0; y < h; y++) w[y] = cos((y + eo - (eh / 2 - 0.5)) * pi / eh);
  Kbns k;
  for (int y = 0; y < h; y++) k.acc(w[y]);
  return (double)h / k.result();
}
inline void ws_weights(int h, int lat_range_deg, std::vector<double>& w, double& correction) {
  w.resize(h);
  correction = ws_weights_into(h, lat_range_deg, w.data());
}

// (v * w).getSum() * (1 / sum(w)): libs/pmbb-ivqm/src/xIVPSNR.cpp:120-123.
inline double weighted_mean4(const double v[4], const int32_t cw[4]) {
  const double inv = 1.0 / (double)(cw[0] + cw[1] + cw[2] + cw[3]);
  return (v[0] * (double)cw[0] + v[1] * (double)cw[1] + v[2] * (double)cw[2] + v[3] * (double)cw[3]) * inv;
}
// Picture value of a per-component metric: apps/QMIV/src/xMetricQMIV.cpp:91-97.
inline double pic_value(const double v[4], const int32_t cw[4]) {
  const double inv = 1.0 / (double)(cw[0] + cw[1] + cw[2] + cw[3]);
  return (v[0] * cw[0] + v[1] * cw[1] + v[2] * cw[2]) * inv;
}

// Asymmetric IV-PSNR quality from per-row distances (device uint32 rows,
// [3][h]; the reference's int32 row sums are sign-extended to uint64,
// libs/pmbb-ivqm/src/xCorrespPixelShiftSSE.cpp:38-40):
// libs/pmbb-ivqm/src/xIVPSNR.cpp:99-125.
inline double ivpsnr_quality(const uint32_t* rows, int w, int h, int bit_depth, const double* erp_w,
                             const int32_t cw_avg[4]) {
  double err[4] = {0, 0, 0, 0};
  if (erp_w) {
    Kbns k[3];
    for (int y = 0; y < h; y++)
      for (int c = 0; c < 3; c++) k[c].acc((double)(uint64_t)(int64_t)(int32_t)rows[(size_t)c * h + y] * erp_w[y]);
    for (int c = 0; c < 3; c++) err[c] = k[c].result();
  } else {
    uint64_t s[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++)
      for (int y = 0; y < h; y++) s[c] += (uint64_t)(int64_t)(int32_t)rows[(size_t)c * h + y];
    for (int c = 0; c < 3; c++) err[c] = (double)s[c];
  }
  double q[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++) q[c] = psnr_from_ssd(err[c] > 0 ? err[c] : 1.0, (int64_t)w * h, bit_depth);
  return weighted_mean4(q, cw_avg);
}

// The PSNR family of one unmasked pair from the device's exact integers; r.ssd and r.exact are set.
// row_ssd: [3][plane rows] SSD rows (read for WS-PSNR with ERP weights); rows: the search's [2][3][H]
// row sums (IV-PSNR); erp_w / erp_corr: row weights and correction of height H, or null without ERP.
// cs / csy: chroma column / row shifts of the pair's pictures (4:2:0 1 / 1, 4:2:2 1 / 0, 4:4:4 0 / 0).
inline void psnr_family(uint32_t m, int W, int H, int bd, int cs, int csy, bool fake_exact, const int32_t cwa[4],
                        const uint64_t* row_ssd, const uint32_t* rows, const double* erp_w, double erp_corr,
                        ivqm_frame_result& r) {
  const int64_t area = (int64_t)W * H;
  const double fake = psnr_from_ssd(1.0, area, bd);
  if (m & (IVQM_M_PSNR | IVQM_M_MSE)) {
    // libs/pmbb-ivqm/src/xPSNR.cpp:131-151; exact components -> fake PSNR (xAppQMIV.cpp:974-981)
    for (int k = 0; k < 4; k++) { r.mse[k] = DBL_MAX; r.psnr[k] = DBL_MAX; }
    for (int k = 0; k < 3; k++) {
      const double ssd = (double)r.ssd[k];
      r.mse[k] = ssd / (double)area;
      r.psnr[k] = psnr_from_ssd(ssd, area, bd);
      if (fake_exact && r.exact[k]) r.psnr[k] = fake;
    }
    r.psnr_pic = pic_value(r.psnr, cwa);
  }
  if (m & IVQM_M_WSPSNR) {
    for (int k = 0; k < 4; k++) r.wspsnr[k] = DBL_MAX;
    if (!erp_w) {
      for (int k = 0; k < 3; k++) r.wspsnr[k] = psnr_from_ssd((double)r.ssd[k], area, bd);
    } else {
      // libs/pmbb-ivqm/src/xWSPSNR.cpp:81-107; a subsampled chroma row counts 2^cs times per 4:4:4 row,
      // and serves 2^csy of them
      for (int k = 0; k < 3; k++) {
        Kbns kb;
        for (int y = 0; y < H; y++) {
          const uint64_t row = k ? ((uint64_t)1 << cs) * row_ssd[(size_t)k * H + (y >> csy)] : row_ssd[y];
          kb.acc((double)row * erp_w[y]);
        }
        const double err = kb.result() * erp_corr;
        r.wspsnr[k] = psnr_from_ssd(err, area, bd);
        if (fake_exact && err == 0) r.wspsnr[k] = fake;
      }
    }
    for (int k = 0; k < 3; k++)
      if (fake_exact && r.exact[k]) r.wspsnr[k] = fake;
    r.wspsnr_pic = pic_value(r.wspsnr, cwa);
  }
  if (m & IVQM_M_IVPSNR) {
    r.ivpsnr_r2t = ivpsnr_quality(rows, W, H, bd, erp_w, cwa);
    r.ivpsnr_t2r = ivpsnr_quality(rows + (size_t)3 * H, W, H, bd, erp_w, cwa);
    r.ivpsnr = (r.ivpsnr_t2r < r.ivpsnr_r2t) ? r.ivpsnr_t2r : r.ivpsnr_r2t;  // xIVPSNR.cpp:55 (xMin = std::min)
  }
}

// Masked PSNR: libs/pmbb-ivqm/src/xPSNR.cpp:164-179 / :199-206,
// MAX = NumNonMasked * MaxPic^2 * MaxMsk in int64.
inline double psnr_masked(double ssdm, int64_t nnm, int bit_depth_pic, int bit_depth_msk) {
  const int64_t maxp = (1 << bit_depth_pic) - 1, maxm = (1 << bit_depth_msk) - 1;
  const int64_t MAX = nnm * (maxp * maxp) * maxm;
  return ssdm > 0 ? 10.0 * log10((double)MAX / ssdm) : DBL_MAX;
}

// Masked asymmetric IV-PSNR quality from uint64 rows [3][h]:
// libs/pmbb-ivqm/src/xIVPSNR.cpp:172-199.
inline double ivpsnr_quality_masked(const uint64_t* rows, int h, int64_t nnm, int bit_depth, int bit_depth_msk,
                                    const double* erp_w, const int32_t cw_avg[4]) {
  double err[4] = {0, 0, 0, 0};
  if (erp_w) {
    Kbns k[3];
    for (int y = 0; y < h; y++)
      for (int c = 0; c < 3; c++) k[c].acc((double)rows[(size_t)c * h + y] * erp_w[y]);
    for (int c = 0; c < 3; c++) err[c] = k[c].result();
  } else {
    uint64_t s[3] = {0, 0, 0};
    for (int c = 0; c < 3; c++)
      for (int y = 0; y < h; y++) s[c] += rows[(size_t)c * h + y];
    for (int c = 0; c < 3; c++) err[c] = (double)s[c];
  }
  double q[4] = {0, 0, 0, 0};
  for (int c = 0; c < 3; c++) q[c] = psnr_masked(err[c] > 0 ? err[c] : 1.0, nnm, bit_depth, bit_depth_msk);
  return weighted_mean4(q, cw_avg);
}

}  // namespace host
}  // namespace ivqm
