// ORACLE - test infrastructure only.
//
// C-ABI driver over the reference pmbb libraries compiled from their own
// sources (oracle/Makefile builds it into oracle/_ref/). It is used to write
// golden vectors, to cross-check the C restatement (oracle/ivqm_oracle.c) and
// as the CPU baseline that bench.py times. Nothing in the product links it.
//
// Every entry point takes 4:4:4 planar uint16 planes, tightly packed, wraps
// them into the reference's own picture types (xPicP with a Nearest-extended
// margin of roundup(R, 4), xPicI interleaved) and calls the reference classes
// the way the QMIV app does (apps/QMIV/src/xAppQMIV.cpp:680-744 for the
// processor set-up, :927-1108 for the per-frame calls).
#include <chrono>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "xGlobClrDiff.h"
#include "xIVPSNR.h"
#include "xIVSSIM.h"
#include "xKBNS.h"
#include "xPSNR.h"
#include "xPic.h"
#include "xPixelOps.h"
#include "xSSIM.h"
#include "xShftCompPic.h"
#include "xThreadPool.h"
#include "xWSPSNR.h"
#include "xDistortion.h"
#include "xColorSpace.h"

using namespace PMBB_NAMESPACE;

namespace {

int32 margin_for(int32 R) { return xRoundUpToNearestMultiple(R, 2); }

void load_pic(xPicP& pic, const uint16_t* const planes[3], int32 w, int32 h, int32 bd, int32 margin) {
  pic.create(int32V2(w, h), bd, margin);
  for (int c = 0; c < 3; c++) {
    uint16* dst = pic.getAddr((eCmp)c);
    for (int32 y = 0; y < h; y++) std::memcpy(dst + (size_t)y * pic.getStride(), planes[c] + (size_t)y * w, (size_t)w * 2);
  }
  pic.extend(eMrgExt::Nearest);
}
void store_pic(uint16_t* const planes[3], const xPicP& pic) {
  const int32 w = pic.getWidth(), h = pic.getHeight();
  for (int c = 0; c < 3; c++) {
    const uint16* src = pic.getAddr((eCmp)c);
    for (int32 y = 0; y < h; y++) std::memcpy(planes[c] + (size_t)y * w, src + (size_t)y * pic.getStride(), (size_t)w * 2);
  }
}
int32V4 v4(const int32_t a[4]) { return int32V4(a[0], a[1], a[2], a[3]); }

// A pool-less thread-pool interface executes stored tasks inline
// (libs/pmbb-core/src/xThreadPool.cpp:267-285).
struct Inline {
  tThPI tpi;
};

// Exposes the protected per-row distance computation of xIVPSNR.
struct IVPSNRRows : public xIVPSNR {
  void rows(const xPicI* tst, const xPicI* ref, const int32V4& gcs, uint64_t* out) {
    const int32 h = ref->getHeight();
    for (int32 y = 0; y < h; y++) {
      uint64V4 r = tCPS::CalcDistAsymmetricRow(tst, ref, y, gcs, m_SearchRange, m_CmpWeightsSearch);
      for (int c = 0; c < 4; c++) out[(size_t)y * 4 + c] = r[c];
    }
  }
};

}  // namespace

extern "C" {

int ref_upsample_hv(uint16_t* dst, const uint16_t* src, int32_t dst_w, int32_t dst_h) {
  xPixelOps::UpsampleHV(dst, src, dst_w, dst_w >> 1, dst_w, dst_h);
  return 0;
}
int ref_upsample_h(uint16_t* dst, const uint16_t* src, int32_t dst_w, int32_t dst_h) {
  xPixelOps::UpsampleH(dst, src, dst_w, dst_w >> 1, dst_w, dst_h);
  return 0;
}
int ref_downsample_hv(uint16_t* dst, const uint16_t* src, int32_t src_w, int32_t src_h) {
  xPixelOps::DownsampleHV(dst, src, src_w >> 1, src_w, src_w >> 1, src_h >> 1);
  return 0;
}
int64_t ref_sd(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t bd) {
  return xDistortion::CalcSD(tst, ref, w, w, w, h, bd);
}
uint64_t ref_ssd(const uint16_t* tst, const uint16_t* ref, int32_t w, int32_t h, int32_t bd) {
  return xDistortion::CalcSSD(tst, ref, w, w, w, h, bd);
}

// xGlobClrDiffProc::CalcGlobalColorDiff(P0, P1) (xAppQMIV.cpp:931).
int ref_gcd(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd,
            const float coef[4], int32_t R, int32_t out[4]) {
  xPicP a, b;
  load_pic(a, p0, w, h, bd, margin_for(R));
  load_pic(b, p1, w, h, bd, margin_for(R));
  xGlobClrDiffProc proc;
  proc.setUnntcbCoef(flt32V4(coef[0], coef[1], coef[2], coef[3]));
  int32V4 g = proc.CalcGlobalColorDiff(&a, &b);
  for (int c = 0; c < 4; c++) out[c] = g[c];
  return 0;
}

// Per-row distances of one search direction on interleaved pictures (the SSE
// kernel the x86 build dispatches to, xCorrespPixelShift.h:44).
int ref_cps_rows(const uint16_t* const tst[3], const uint16_t* const ref[3], int32_t w, int32_t h, int32_t bd,
                 const int32_t gcs[4], int32_t R, const int32_t cw[4], uint64_t* rows) {
  xPicP a, b;
  load_pic(a, tst, w, h, bd, margin_for(R));
  load_pic(b, ref, w, h, bd, margin_for(R));
  xPicI ai(int32V2(w, h), bd, margin_for(R)), bi(int32V2(w, h), bd, margin_for(R));
  ai.rearrangeFromPlanar(&a);
  bi.rearrangeFromPlanar(&b);
  IVPSNRRows p;
  p.setSearchRange(R);
  p.setCmpWeightsSearch(v4(cw));
  p.rows(&ai, &bi, v4(gcs), rows);
  return 0;
}

// xShftCompPicProc::GenShftCompPics on interleaved pictures + rearrangeToPlanar
// (xAppQMIV.cpp:934-948).
int ref_scp(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd,
            const int32_t gcd[4], int32_t R, const int32_t cw[4], uint16_t* const scp0[3], uint16_t* const scp1[3]) {
  const int32 m = margin_for(R);
  xPicP a, b, s0(int32V2(w, h), bd, m), s1(int32V2(w, h), bd, m);
  load_pic(a, p0, w, h, bd, m);
  load_pic(b, p1, w, h, bd, m);
  xPicI ai(int32V2(w, h), bd, m), bi(int32V2(w, h), bd, m), si0(int32V2(w, h), bd, m), si1(int32V2(w, h), bd, m);
  ai.rearrangeFromPlanar(&a);
  bi.rearrangeFromPlanar(&b);
  xShftCompPicProc proc;
  proc.setSearchRange(R);
  proc.setCmpWeightsSearch(v4(cw));
  proc.GenShftCompPics(&si1, &si0, &bi, &ai, v4(gcd));
  si0.rearrangeToPlanar(&s0);
  si1.rearrangeToPlanar(&s1);
  store_pic(scp0, s0);
  store_pic(scp1, s1);
  return 0;
}

int ref_psnr(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd, double out[4]) {
  xPicP a, b;
  load_pic(a, p0, w, h, bd, 4);
  load_pic(b, p1, w, h, bd, 4);
  Inline in;
  xPSNR p;
  p.bindThrdPoolIntf(&in.tpi);
  flt64V4 r = p.calcPicPSNR(&a, &b);
  for (int c = 0; c < 4; c++) out[c] = r[c];
  return 0;
}

int ref_wspsnr(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd, int32_t erp,
               int32_t lat, double out[4]) {
  xPicP a, b;
  load_pic(a, p0, w, h, bd, 4);
  load_pic(b, p1, w, h, bd, 4);
  Inline in;
  xWSPSNR p;
  p.bindThrdPoolIntf(&in.tpi);
  p.initRowBuffers(h);
  if (erp) p.initWS(true, w, h, bd, 360, lat);
  flt64V4 r = p.calcPicWSPSNR(&a, &b);
  for (int c = 0; c < 4; c++) out[c] = r[c];
  return 0;
}

// xIVPSNR::calcPicIVPSNR(I0, I1, GCD_R2T) (xAppQMIV.cpp:1031).
int ref_ivpsnr(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd,
               const int32_t gcd[4], int32_t R, const int32_t cws[4], const int32_t cwa[4], int32_t erp, int32_t lat,
               double* out, double* r2t, double* t2r) {
  const int32 m = margin_for(R);
  xPicP a, b;
  load_pic(a, p0, w, h, bd, m);
  load_pic(b, p1, w, h, bd, m);
  xPicI ai(int32V2(w, h), bd, m), bi(int32V2(w, h), bd, m);
  ai.rearrangeFromPlanar(&a);
  bi.rearrangeFromPlanar(&b);
  Inline in;
  xIVPSNR p;
  p.setSearchRange(R);
  p.setCmpWeightsSearch(v4(cws));
  p.setCmpWeightsAverage(v4(cwa));
  p.bindThrdPoolIntf(&in.tpi);
  p.initRowBuffers(h);
  if (erp) p.initWS(true, w, h, bd, 360, lat);
  double lr = 0, lt = 0;
  p.setDebugCallbackQAP([&](flt64 x, flt64 y) { lr = x; lt = y; });
  *out = p.calcPicIVPSNR(&ai, &bi, v4(gcd));
  if (r2t) *r2t = lr;
  if (t2r) *t2r = lt;
  return 0;
}

namespace {
void setup_ssim(xIVSSIM& p, tThPI* tpi, int32 w, int32 h, int32 bd, int32 R, bool ms, int32 wnd, int32 stride,
                const int32_t cwa[4], int32 erp, int32 lat) {
  p.create(int32V2(w, h), bd, margin_for(R), ms);
  p.setSearchRange(R);
  p.setCmpWeightsAverage(v4(cwa));
  p.setStructSimParams(xSSIM::eMode::BlockAveraged, eMrgExt::None, wnd, stride);
  p.bindThrdPoolIntf(tpi);
  p.initRowBuffers(h);
  if (erp) p.initWS(true, w, h, bd, 360, lat);
}
}  // namespace

// xSSIM::calcPicSSIM / calcPicMSSSIM (xAppQMIV.cpp:1050-1072).
int ref_ssim(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd, int32_t wnd,
             int32_t stride, int32_t ms, double out[4]) {
  xPicP a, b;
  load_pic(a, p0, w, h, bd, 4);
  load_pic(b, p1, w, h, bd, 4);
  Inline in;
  const int32_t cwa[4] = {4, 1, 1, 0};
  xIVSSIM p;
  setup_ssim(p, &in.tpi, w, h, bd, 2, ms != 0, wnd, stride, cwa, 0, 180);
  flt64V4 r = ms ? p.calcPicMSSSIM(&a, &b) : p.calcPicSSIM(&a, &b);
  for (int c = 0; c < 4; c++) out[c] = r[c];
  p.destroy();
  return 0;
}

// xIVSSIM::calcPicIVSSIM / calcPicIVMSSSIM (xAppQMIV.cpp:1074-1106).
int ref_ivssim(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* const scp0[3],
               const uint16_t* const scp1[3], int32_t w, int32_t h, int32_t bd, int32_t wnd, int32_t stride,
               const int32_t cwa[4], int32_t erp, int32_t lat, int32_t ms, double* out) {
  xPicP a, b, s0, s1;
  load_pic(a, p0, w, h, bd, 4);
  load_pic(b, p1, w, h, bd, 4);
  load_pic(s0, scp0, w, h, bd, 4);
  load_pic(s1, scp1, w, h, bd, 4);
  Inline in;
  xIVSSIM p;
  setup_ssim(p, &in.tpi, w, h, bd, 2, ms != 0, wnd, stride, cwa, erp, lat);
  *out = ms ? p.calcPicIVMSSSIM(&a, &b, &s0, &s1) : p.calcPicIVSSIM(&a, &b, &s0, &s1);
  p.destroy();
  return 0;
}

// Any structural-similarity mode (xSSIM.h:24-31) and border extension
// (xCommonDefCORE.h:322-331), on pictures with a margin of `margin` samples.
// With ext != None the pictures are extended by that mode before the call, as
// xAppQMIV::addStructSimMargs does (xAppQMIV.cpp:918-925); the library extends
// its MS-SSIM sub-pictures itself (xSSIM.cpp:128).
namespace {
void load_ext(xPicP& pic, const uint16_t* const planes[3], int32 w, int32 h, int32 bd, int32 margin, int32 ext) {
  load_pic(pic, planes, w, h, bd, margin);
  if (ext != 0) pic.extend((eMrgExt)ext);
}
}  // namespace

int ref_ssim_ex(const uint16_t* const p0[3], const uint16_t* const p1[3], int32_t w, int32_t h, int32_t bd,
                int32_t mode, int32_t ext, int32_t wnd, int32_t stride, int32_t margin, int32_t ms, double out[4]) {
  xPicP a, b;
  load_ext(a, p0, w, h, bd, margin, ext);
  load_ext(b, p1, w, h, bd, margin, ext);
  Inline in;
  const int32_t cwa[4] = {4, 1, 1, 0};
  xIVSSIM p;
  p.create(int32V2(w, h), bd, margin, ms != 0);
  p.setCmpWeightsAverage(v4(cwa));
  p.setStructSimParams((xSSIM::eMode)mode, (eMrgExt)ext, wnd, stride);
  p.bindThrdPoolIntf(&in.tpi);
  p.initRowBuffers(h);
  flt64V4 r = ms ? p.calcPicMSSSIM(&a, &b) : p.calcPicSSIM(&a, &b);
  for (int c = 0; c < 4; c++) out[c] = r[c];
  p.destroy();
  return 0;
}

int ref_ivssim_ex(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* const scp0[3],
                  const uint16_t* const scp1[3], int32_t w, int32_t h, int32_t bd, int32_t mode, int32_t ext,
                  int32_t wnd, int32_t stride, int32_t margin, const int32_t cwa[4], int32_t erp, int32_t lat,
                  int32_t ms, double* out, double* r2t, double* t2r) {
  xPicP a, b, s0, s1;
  load_ext(a, p0, w, h, bd, margin, ext);
  load_ext(b, p1, w, h, bd, margin, ext);
  load_ext(s0, scp0, w, h, bd, margin, ext);
  load_ext(s1, scp1, w, h, bd, margin, ext);
  Inline in;
  xIVSSIM p;
  p.create(int32V2(w, h), bd, margin, ms != 0);
  p.setCmpWeightsAverage(v4(cwa));
  p.setStructSimParams((xSSIM::eMode)mode, (eMrgExt)ext, wnd, stride);
  p.bindThrdPoolIntf(&in.tpi);
  p.initRowBuffers(h);
  if (erp) p.initWS(true, w, h, bd, 360, lat);
  double lr = 0, lt = 0;
  p.setDebugCallbackQAP([&](flt64 x, flt64 y) { lr = x; lt = y; });
  *out = ms ? p.calcPicIVMSSSIM(&a, &b, &s0, &s1) : p.calcPicIVSSIM(&a, &b, &s0, &s1);
  if (r2t) *r2t = lr;
  if (t2r) *t2r = lt;
  p.destroy();
  return 0;
}


// Mask mode, as the app runs it (xAppQMIV.cpp:899-906 extends every input
// picture, the mask included, with Nearest margins and counts the non-zero
// mask samples; :930, :953, :971, :999, :1026-1029 call the masked
// processors). The mask is one luma-resolution plane of bit depth bdm.
namespace {
struct Masked {
  xPicP a, b, m;
  int32 nnm = 0;
  Masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk, int32 w, int32 h, int32 bd,
         int32 bdm, int32 margin) {
    load_pic(a, p0, w, h, bd, margin);
    load_pic(b, p1, w, h, bd, margin);
    const uint16_t* mp[3] = {msk, msk, msk};
    load_pic(m, mp, w, h, bdm, margin);
    nnm = xPixelOps::CountNonZero(m.getAddr(eCmp::LM), m.getStride(), m.getWidth(), m.getHeight());
  }
};
}  // namespace

int ref_gcd_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk, int32_t w, int32_t h,
                   int32_t bd, int32_t bdm, int32_t R, const float coef[4], int32_t out[4]) {
  Masked M(p0, p1, msk, w, h, bd, bdm, margin_for(R));
  Inline in;
  xGlobClrDiffProc g;
  g.setUnntcbCoef(flt32V4(coef[0], coef[1], coef[2], coef[3]));
  g.bindThrdPoolIntf(&in.tpi);
  int32V4 r = g.CalcGlobalColorDiffM(&M.a, &M.b, &M.m, M.nnm);
  for (int c = 0; c < 4; c++) out[c] = r[c];
  return M.nnm;
}

int ref_psnr_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk, int32_t w,
                    int32_t h, int32_t bd, int32_t bdm, int32_t erp, int32_t lat, double mse[4], double psnr[4],
                    double wspsnr[4]) {
  Masked M(p0, p1, msk, w, h, bd, bdm, 4);
  Inline in;
  xIVPSNRM p;
  p.bindThrdPoolIntf(&in.tpi);
  p.initRowBuffers(h);
  if (erp) p.initWS(true, w, h, bd, 360, lat);
  flt64V4 a = p.calcPicMSEM(&M.a, &M.b, &M.m, M.nnm);
  flt64V4 b = p.calcPicPSNRM(&M.a, &M.b, &M.m, M.nnm);
  flt64V4 c = p.calcPicWSPSNRM(&M.a, &M.b, &M.m, M.nnm);
  for (int k = 0; k < 4; k++) { mse[k] = a[k]; psnr[k] = b[k]; wspsnr[k] = c[k]; }
  return M.nnm;
}

// xShftCompPicProc::GenShftCompPicsM(SCI1, SCI0, I1, I0, Msk, GCD_R2T) + rearrangeToPlanar
// (xAppQMIV.cpp:937-942). The destination pictures start filled with `fill`, which the
// reference keeps at pixels whose mask sample is 0.
int ref_scp_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk, int32_t w,
                   int32_t h, int32_t bd, const int32_t gcd[4], int32_t R, const int32_t cw[4], int32_t fill,
                   uint16_t* const scp0[3], uint16_t* const scp1[3]) {
  const int32 mg = margin_for(R);
  Masked M(p0, p1, msk, w, h, bd, bd, mg);
  xPicI ai(int32V2(w, h), bd, mg), bi(int32V2(w, h), bd, mg), si0(int32V2(w, h), bd, mg), si1(int32V2(w, h), bd, mg);
  ai.rearrangeFromPlanar(&M.a);
  bi.rearrangeFromPlanar(&M.b);
  xPicP f(int32V2(w, h), bd, mg), s0(int32V2(w, h), bd, mg), s1(int32V2(w, h), bd, mg);
  for (int c = 0; c < 3; c++) {
    uint16* dst = f.getAddr((eCmp)c);
    for (int32 y = 0; y < h; y++)
      for (int32 x = 0; x < w; x++) dst[(size_t)y * f.getStride() + x] = (uint16)fill;
  }
  si0.rearrangeFromPlanar(&f);
  si1.rearrangeFromPlanar(&f);
  Inline in;
  xShftCompPicProc proc;
  proc.setSearchRange(R);
  proc.setCmpWeightsSearch(v4(cw));
  proc.bindThrdPoolIntf(&in.tpi);
  proc.GenShftCompPicsM(&si1, &si0, &bi, &ai, &M.m, v4(gcd));
  si0.rearrangeToPlanar(&s0);
  si1.rearrangeToPlanar(&s1);
  store_pic(scp0, s0);
  store_pic(scp1, s1);
  return M.nnm;
}

int ref_ivpsnr_masked(const uint16_t* const p0[3], const uint16_t* const p1[3], const uint16_t* msk, int32_t w,
                      int32_t h, int32_t bd, int32_t bdm, const int32_t gcd[4], int32_t R, const int32_t cws[4],
                      const int32_t cwa[4], int32_t erp, int32_t lat, double* out, double* r2t, double* t2r) {
  const int32 mg = margin_for(R);
  Masked M(p0, p1, msk, w, h, bd, bdm, mg);
  xPicI ai(int32V2(w, h), bd, mg), bi(int32V2(w, h), bd, mg);
  ai.rearrangeFromPlanar(&M.a);
  bi.rearrangeFromPlanar(&M.b);
  Inline in;
  xIVPSNRM p;
  p.setSearchRange(R);
  p.setCmpWeightsSearch(v4(cws));
  p.setCmpWeightsAverage(v4(cwa));
  p.bindThrdPoolIntf(&in.tpi);
  p.initRowBuffers(h);
  if (erp) p.initWS(true, w, h, bd, 360, lat);
  double lr = 0, lt = 0;
  p.setDebugCallbackQAP([&](flt64 x, flt64 y) { lr = x; lt = y; });
  *out = p.calcPicIVPSNRM(&ai, &bi, &M.m, M.nnm, v4(gcd));
  if (r2t) *r2t = lr;
  if (t2r) *t2r = lt;
  return M.nnm;
}

// xColorSpace::ConvertRGB2YCbCr / ConvertYCbCr2RGB in place on 4:4:4 planes (xAppQMIV.cpp:856-876
// calls them the same way); dir 0 = YCbCr -> RGB, 1 = RGB -> YCbCr; space = eClrSpcLC 0..3.
int ref_convert_colour(uint16_t* c0, uint16_t* c1, uint16_t* c2, int32_t w, int32_t h, int32_t bd, int32_t dir,
                       int32_t space) {
  if (dir == 0) xColorSpace::ConvertYCbCr2RGB(c0, c1, c2, c0, c1, c2, w, w, w, h, bd, (eClrSpcLC)space);
  else xColorSpace::ConvertRGB2YCbCr(c0, c1, c2, c0, c1, c2, w, w, w, h, bd, (eClrSpcLC)space);
  return 0;
}

// ---------------------------------------------------------------------------
// CPU baseline: replays the per-frame stage order of xAppQMIV::processAllFrames
// (xAppQMIV.cpp:761-827) on frames already in memory, with the reference
// thread pool of `threads` workers. Stages timed: unpack (4:2:0 -> 4:4:4,
// xSeq.cpp:231-301), preproc (exact-component check + margin extension),
// rearrange (planar -> interleaved), GCD, SCP (if IV-SSIM), IV-PSNR, IV-SSIM.
// metrics bit 0 = IV-PSNR, bit 1 = IV-SSIM. Returns total seconds.
// ---------------------------------------------------------------------------
// stage_s[7] (may be NULL) receives the seconds spent per stage, summed over the pairs, in the
// app's stage order (xAppQMIV.cpp:761-827): 0 unpack (readFrame's 4:2:0 upsample, no file I/O),
// 1 PREPROC (exact-component check, margins), 2 Rearrange, 3 GCD, 4 SCP, 5 IV-PSNR, 6 IV-SSIM.
// unsigned hardware_concurrency() is what the app's thread policy reads (xAppQMIV.cpp:517).
unsigned ref_hardware_concurrency() { return std::thread::hardware_concurrency(); }

double ref_bench_pairs(int32_t n_pairs, const uint16_t* const* p0_planes, const uint16_t* const* p1_planes,
                       int32_t w, int32_t h, int32_t bd, int32_t chroma420, int32_t erp, int32_t metrics,
                       int32_t threads, double* ivpsnr_out, double* ivssim_out, double* stage_s) {
  const int32 R = 2, m = margin_for(R);
  const int32_t cw[4] = {4, 1, 1, 0};
  xThreadPool* pool = nullptr;
  tThPI tpi;
  if (threads > 0) {
    pool = new xThreadPool;
    pool->create(threads, h + 1);
    tpi.init(pool, h + 1, h + 1);
  }
  xPicP P[2], SCP[2];
  xPicI I[2], SCI[2];
  for (int i = 0; i < 2; i++) {
    P[i].create(int32V2(w, h), bd, m);
    I[i].create(int32V2(w, h), bd, m);
    SCP[i].create(int32V2(w, h), bd, m);
    SCI[i].create(int32V2(w, h), bd, m);
  }
  xGlobClrDiffProc gcd;
  gcd.bindThrdPoolIntf(&tpi);
  xShftCompPicProc scp;
  scp.setSearchRange(R);
  scp.setCmpWeightsSearch(v4(cw));
  scp.bindThrdPoolIntf(&tpi);
  xIVPSNR ivp;
  ivp.setSearchRange(R);
  ivp.setCmpWeightsSearch(v4(cw));
  ivp.setCmpWeightsAverage(v4(cw));
  ivp.bindThrdPoolIntf(&tpi);
  ivp.initRowBuffers(h);
  if (erp) ivp.initWS(true, w, h, bd, 360, 180);
  xIVSSIM ivs;
  setup_ssim(ivs, &tpi, w, h, bd, R, false, 8, 4, cw, erp, 180);

  const int32 cs = chroma420 ? 1 : 0;
  double st[7] = {0, 0, 0, 0, 0, 0, 0};
  auto lap = std::chrono::steady_clock::now();
  auto mark = [&](int k) {
    const auto now = std::chrono::steady_clock::now();
    st[k] += std::chrono::duration<double>(now - lap).count();
    lap = now;
  };
  auto t0 = std::chrono::steady_clock::now();
  for (int32_t f = 0; f < n_pairs; f++) {
    lap = std::chrono::steady_clock::now();
    const uint16_t* const* src[2] = {p0_planes + 3 * f, p1_planes + 3 * f};
    for (int i = 0; i < 2; i++)
      tpi.storeTask([&, i](int32) {
        for (int c = 0; c < 3; c++) {
          uint16* dst = P[i].getAddr((eCmp)c);
          if (c == 0 || !cs) {
            for (int32 y = 0; y < h; y++) std::memcpy(dst + (size_t)y * P[i].getStride(), src[i][c] + (size_t)y * w, (size_t)w * 2);
          } else {
            xPixelOps::UpsampleHV(dst, src[i][c], P[i].getStride(), w >> 1, w, h);
          }
        }
      });
    tpi.executeStoredTasks();
    mark(0);
    bool exact[3];
    for (int c = 0; c < 3; c++) exact[c] = P[0].equalCmp(&P[1], (eCmp)c);
    (void)exact;
    for (int i = 0; i < 2; i++) tpi.storeTask([&, i](int32) { P[i].extend(); });
    tpi.executeStoredTasks();
    mark(1);
    for (int i = 0; i < 2; i++) tpi.storeTask([&, i](int32) { I[i].rearrangeFromPlanar(&P[i]); });
    tpi.executeStoredTasks();
    mark(2);
    int32V4 g = gcd.CalcGlobalColorDiff(&P[0], &P[1]);
    mark(3);
    if (metrics & 2) {
      scp.GenShftCompPics(&SCI[1], &SCI[0], &I[1], &I[0], g);
      for (int i = 0; i < 2; i++) tpi.storeTask([&, i](int32) { SCI[i].rearrangeToPlanar(&SCP[i]); });
      tpi.executeStoredTasks();
    }
    mark(4);
    if (metrics & 1) {
      double v = ivp.calcPicIVPSNR(&I[0], &I[1], g);
      if (ivpsnr_out) ivpsnr_out[f] = v;
    }
    mark(5);
    if (metrics & 2) {
      double v = ivs.calcPicIVSSIM(&P[0], &P[1], &SCP[0], &SCP[1]);
      if (ivssim_out) ivssim_out[f] = v;
    }
    mark(6);
  }
  auto t1 = std::chrono::steady_clock::now();
  if (stage_s)
    for (int k = 0; k < 7; k++) stage_s[k] = st[k];
  ivs.destroy();
  if (pool) {
    tpi.uninit();
    pool->destroy();
    delete pool;
  }
  return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
