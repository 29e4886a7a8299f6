// ivqm_dropin.hpp - header-only C++ classes with the call shape of QMIV's
// pmbb-ivqm metric processors, forwarding to the C ABI (ivqm_hip.h).
//
// They let a QMIV-style frame loop (apps/QMIV/src/xAppQMIV.cpp:927-1108) keep
// its calls while the work runs on the MI355X. Each processor owns one device
// context, created lazily from the first picture it sees; every call uploads
// the pictures it is given (as the reference reads them from host memory).
// The fused single-pass form is ivqm_frame() in ivqm_hip.h.
//
// Picture access goes through PicTraits<Pic>; the default expects
// getWidth(), getHeight(), getBitDepth(), getStride() and getAddr(int).
// For the reference's xPicP specialise it:
//   template <> struct ivqm::dropin::PicTraits<PMBB::xPicP> {
//     static const uint16_t* plane(const PMBB::xPicP* p, int c) { return p->getAddr((PMBB::eCmp)c); }
//     static uint16_t* plane_mut(PMBB::xPicP* p, int c) { return p->getAddr((PMBB::eCmp)c); }
//   };
#pragma once
#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>

#include "ivqm_hip.h"

namespace ivqm {
namespace dropin {

using int32V4 = std::array<int32_t, 4>;
using flt64V4 = std::array<double, 4>;
using flt32V4 = std::array<float, 4>;

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};
inline void check(int s) {
  if (s != IVQM_OK) throw Error(s, std::string("ivqm: ") + ivqm_strerror(s) + ": " + ivqm_last_error());
}

template <class Pic> struct PicTraits {
  static const uint16_t* plane(const Pic* p, int c) { return p->getAddr(c); }
  static uint16_t* plane_mut(Pic* p, int c) { return p->getAddr(c); }
};

// Parameter state shared by all processors (the reference's setters:
// xCorrespPixelShiftPrms.h:30-32, xGlobClrDiff.h:24, xWeightedSpherically.h:21,
// xSSIM.h:81, xPSNR.h:28).
class Processor {
 protected:
  ivqm_config m_Cfg;
  ivqm_ctx* m_Ctx = nullptr;

  template <class Pic> void xEnsure(const Pic* p) {
    if (m_Ctx && m_Cfg.width == p->getWidth() && m_Cfg.height == p->getHeight() && m_Cfg.bit_depth == p->getBitDepth())
      return;
    if (m_Ctx) ivqm_destroy(m_Ctx);
    m_Ctx = nullptr;
    m_Cfg.width = p->getWidth();
    m_Cfg.height = p->getHeight();
    m_Cfg.bit_depth = p->getBitDepth();
    check(ivqm_create(&m_Cfg, &m_Ctx));
  }
  template <class Pic> void xUpload(int slot, const Pic* p) {
    ivqm_picture v;
    for (int c = 0; c < 3; c++) {
      v.plane[c] = PicTraits<Pic>::plane(p, c);
      v.stride[c] = p->getStride();
    }
    v.chroma_420 = 0;  // xPicP holds 4:4:4 planes (xSeq.cpp:231-301 upsamples on read)
    v.bytes_per_sample = 2;
    check(ivqm_upload(m_Ctx, slot, &v));
  }
  template <class Pic> void xPair(const Pic* Tst, const Pic* Ref) {
    xEnsure(Tst);
    xUpload(0, Tst);
    xUpload(1, Ref);
  }
  void xReset() {
    if (m_Ctx) ivqm_destroy(m_Ctx);
    m_Ctx = nullptr;
  }

 public:
  Processor() { ivqm_config_init(&m_Cfg); m_Cfg.width = m_Cfg.height = 0; }
  Processor(const Processor&) = delete;
  Processor& operator=(const Processor&) = delete;
  virtual ~Processor() { xReset(); }

  void setSearchRange(int32_t R) { m_Cfg.search_range = R; xReset(); }
  void setCmpWeightsSearch(const int32V4& w) { for (int i = 0; i < 4; i++) m_Cfg.cmp_weights_search[i] = w[i]; xReset(); }
  void setCmpWeightsAverage(const int32V4& w) { for (int i = 0; i < 4; i++) m_Cfg.cmp_weights_average[i] = w[i]; xReset(); }
  void setUnntcbCoef(const flt32V4& c) { for (int i = 0; i < 4; i++) m_Cfg.unnoticeable_coef[i] = c[i]; xReset(); }
  void setFakeValsForExact(bool v) { m_Cfg.fake_vals_for_exact = v; xReset(); }
  // xWeightedSpherically::initWS(UseWS, W, H, BitDepth, Lon, Lat)
  void initWS(bool UseWS, int32_t, int32_t, int32_t, int32_t Lon = 360, int32_t Lat = 180) {
    m_Cfg.erp = UseWS;
    m_Cfg.lon_range_deg = Lon;
    m_Cfg.lat_range_deg = Lat;
    xReset();
  }
  void initRowBuffers(int32_t) {}           // device row buffers are sized per context
  bool bindThrdPoolIntf(void*) { return true; }  // the thread pool is replaced by the HIP stream
  ivqm_ctx* context() { return m_Ctx; }
};

// xGlobClrDiffProc (libs/pmbb-ivqm/src/xGlobClrDiff.h:41): Tst = P0, Ref = P1.
class xGlobClrDiffProc : public Processor {
 public:
  template <class Pic> int32V4 CalcGlobalColorDiff(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    int32V4 g;
    check(ivqm_gcd(m_Ctx, g.data()));
    return g;
  }
};

// xPSNR / xWSPSNR (libs/pmbb-ivqm/src/xPSNR.h:30-32, xWSPSNR.h:22).
class xWSPSNR : public Processor {
 public:
  template <class Pic> std::array<uint64_t, 4> calcPicSSD(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    std::array<uint64_t, 4> s;
    check(ivqm_ssd(m_Ctx, s.data()));
    return s;
  }
  template <class Pic> flt64V4 calcPicMSE(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    flt64V4 v;
    check(ivqm_mse(m_Ctx, v.data()));
    return v;
  }
  template <class Pic> flt64V4 calcPicPSNR(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    flt64V4 v;
    check(ivqm_psnr(m_Ctx, v.data()));
    return v;
  }
  template <class Pic> flt64V4 calcPicWSPSNR(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    flt64V4 v;
    check(ivqm_wspsnr(m_Ctx, v.data()));
    return v;
  }
};
using xPSNR = xWSPSNR;

// xIVPSNR (libs/pmbb-ivqm/src/xIVPSNR.h:33-36).
class xIVPSNR : public xWSPSNR {
 public:
  double LastR2T = 0, LastT2R = 0;  // what the reference's debug callback QAP receives
  template <class Pic> double calcPicIVPSNR(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    double v;
    check(ivqm_ivpsnr(m_Ctx, nullptr, &v, &LastR2T, &LastT2R));
    return v;
  }
  template <class Pic> double calcPicIVPSNR(const Pic* Tst, const Pic* Ref, const int32V4& GlobalColorDiffRef2Tst) {
    xPair(Tst, Ref);
    double v;
    check(ivqm_ivpsnr(m_Ctx, GlobalColorDiffRef2Tst.data(), &v, &LastR2T, &LastT2R));
    return v;
  }
};

// xShftCompPicProc (libs/pmbb-ivqm/src/xShftCompPic.h:44-47): as the app calls it,
// GenShftCompPics(SCP1, SCP0, P1, P0, GCD_R2T).
class xShftCompPicProc : public Processor {
 public:
  template <class Pic> void GenShftCompPics(Pic* DstRef, Pic* DstTst, const Pic* SrcRef, const Pic* SrcTst,
                                            const int32V4& GlobalColorDiffRef2Tst) {
    xPair(SrcTst, SrcRef);
    check(ivqm_gen_scp(m_Ctx, GlobalColorDiffRef2Tst.data()));
    xDownload(1, DstRef);
    xDownload(0, DstTst);
  }

 private:
  template <class Pic> void xDownload(int idx, Pic* dst) {
    const int w = dst->getWidth(), h = dst->getHeight(), s = dst->getStride();
    if (s == w) {
      uint16_t* planes[3] = {PicTraits<Pic>::plane_mut(dst, 0), PicTraits<Pic>::plane_mut(dst, 1),
                             PicTraits<Pic>::plane_mut(dst, 2)};
      check(ivqm_get_scp(m_Ctx, idx, planes));
      return;
    }
    std::string tmp((size_t)w * h * 3 * sizeof(uint16_t), '\0');
    uint16_t* t = (uint16_t*)&tmp[0];
    uint16_t* planes[3] = {t, t + (size_t)w * h, t + 2 * (size_t)w * h};
    check(ivqm_get_scp(m_Ctx, idx, planes));
    for (int c = 0; c < 3; c++)
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) PicTraits<Pic>::plane_mut(dst, c)[(size_t)y * s + x] = planes[c][(size_t)y * w + x];
  }
};

// xSSIM / xIVSSIM (libs/pmbb-ivqm/src/xSSIM.h:78-84, xIVSSIM.h:36-40).
class xIVSSIM : public Processor {
 public:
  double LastR2T = 0, LastT2R = 0;
  void create(std::array<int32_t, 2>, int32_t BitDepth, int32_t, bool) { m_Cfg.bit_depth = BitDepth; }
  void destroy() { xReset(); }
  // Mode = xSSIM::eMode (xSSIM.h:24-31), MarginMode = eMrgExt (xCommonDefCORE.h:322-331); the
  // reference's enum values pass through (the checks run when the context is created).
  template <class ModeT, class ExtT> void setStructSimParams(ModeT Mode, ExtT MarginMode, int32_t Window, int32_t Stride) {
    m_Cfg.ssim_mode = (int32_t)Mode;
    m_Cfg.ssim_border = (int32_t)MarginMode;
    m_Cfg.ssim_window = Window;
    m_Cfg.ssim_stride = Stride;
    xReset();
  }
  template <class Pic> flt64V4 calcPicSSIM(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    flt64V4 v;
    check(ivqm_ssim(m_Ctx, v.data()));
    return v;
  }
  template <class Pic> flt64V4 calcPicMSSSIM(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    flt64V4 v;
    check(ivqm_msssim(m_Ctx, v.data()));
    return v;
  }
  // calcPicIVSSIM(P0, P1, SCP0, SCP1) as xAppQMIV.cpp:1077 calls it.
  template <class Pic> double calcPicIVSSIM(const Pic* Tst, const Pic* Ref, const Pic* TstSCP, const Pic* RefSCP) {
    xWithScp(Tst, Ref, TstSCP, RefSCP);
    double v;
    check(ivqm_ivssim(m_Ctx, &v, &LastR2T, &LastT2R));
    return v;
  }
  template <class Pic> double calcPicIVMSSSIM(const Pic* Tst, const Pic* Ref, const Pic* TstSCP, const Pic* RefSCP) {
    xWithScp(Tst, Ref, TstSCP, RefSCP);
    double v;
    check(ivqm_ivmsssim(m_Ctx, &v, &LastR2T, &LastT2R));
    return v;
  }

 private:
  template <class Pic> void xWithScp(const Pic* Tst, const Pic* Ref, const Pic* TstSCP, const Pic* RefSCP) {
    if (!TstSCP || !RefSCP) throw Error(IVQM_EINVAL, "IV-SSIM needs both shift-compensated pictures");
    xPair(Tst, Ref);
    xScp(0, TstSCP);
    xScp(1, RefSCP);
  }
  template <class Pic> void xScp(int idx, const Pic* p) {
    ivqm_picture v;
    for (int c = 0; c < 3; c++) {
      v.plane[c] = PicTraits<Pic>::plane(p, c);
      v.stride[c] = p->getStride();
    }
    v.chroma_420 = 0;
    v.bytes_per_sample = 2;
    check(ivqm_upload_scp(m_Ctx, idx, &v));
  }
};
using xSSIM = xIVSSIM;

}  // namespace dropin
}  // namespace ivqm
