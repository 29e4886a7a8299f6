// ivqm_dropin.hpp - header-only C++ classes with the names, methods and
// call semantics of QMIV's pmbb-ivqm metric processors, forwarding to the C
// ABI (ivqm_hip.h) so that the work runs on the MI355X.
//
// Covered: every processor call of the QMIV app's set-up and frame loop
// (apps/QMIV/src/xAppQMIV.cpp:680-744, :927-1108), mask mode included:
//   xGlobClrDiffProc  CalcGlobalColorDiff / CalcGlobalColorDiffM    (xGlobClrDiff.h:38-43)
//   xShftCompPicProc  GenShftCompPics (planar and interleaved) /
//                     GenShftCompPicsM                              (xShftCompPic.h:36-51)
//   xPSNR             calcPicSSD/MSE/PSNR, calcPicSSDM/MSEM/PSNRM,
//                     getFakePSNR, setDebugCallbackMSK              (xPSNR.h:15-56)
//   xWSPSNR           calcPicWSPSNR / calcPicWSPSNRM, initWS        (xWSPSNR.h:15-28)
//   xIVPSNR           calcPicIVPSNR (3 overloads), callbacks
//                     GCS / QAP                                     (xIVPSNR.h:17-42)
//   xIVPSNRM          calcPicIVPSNRM (2 overloads)                  (xIVPSNR.h:46-60)
//   xSSIM / xIVSSIM   create, setStructSimParams, calcPicSSIM,
//                     calcPicMSSSIM, calcPicIVSSIM, calcPicIVMSSSIM,
//                     callbacks, mode helpers                       (xSSIM.h, xIVSSIM.h)
// plus the parameter setters (xCorrespPixelShiftPrms.h:30-32,
// xGlobClrDiff.h:24, xWeightedSpherically.h:21, xMetricCommon.h,
// xCommonDefIVQM.h:42-55).
//
// Pictures. PicTraits<Pic> reads a picture. A type with getAddr() taking no
// argument is interleaved (the reference's xPicI: 4 uint16 samples per pixel,
// Y Cb Cr and an unused 4th, stride in pixels); any other type is planar and is
// read with getAddr(int) (specialise PicTraits for a planar type whose
// getAddr takes an enum, as the reference's xPicP does with eCmp). Every
// picture also needs getWidth(), getHeight(), getBitDepth() and getStride().
// Pictures are 4:4:4 (the reference's xPicP holds upsampled planes).
//
// Vectors. Setters take any type indexable [0..3] (the reference's xVec4,
// std::array, a C array); 4-wide results are std::array, which the
// reference's xVec4 converts from implicitly (xVec.h:267).
//
// Each processor owns one device context, created on the first picture it sees
// and again when the size, bit depth or a parameter changes; every call uploads
// the pictures it is given, as the reference reads them from host memory.
// Failures throw ivqm::dropin::Error (the reference asserts).
//
// Mask mode: the mask is a planar picture whose luma plane holds the weights
// and whose bit depth is BitDepthM. The device counts the non-zero weights
// itself; a NumNonMasked argument of NOT_VALID (-1) means "count it", any
// other value must equal that count (the app passes the count,
// xAppQMIV.cpp:904), else Error. Pixels with weight 0 keep their contents in
// GenShftCompPicsM's destinations, as in the reference.
//
// Not provided (Error when called): structural similarity with a mask
// (calcPicSSIMM, calcPicIVSSIMM; the app rejects the combination,
// xAppQMIV.cpp:360) and the debug dumps visualizeSSIM / visualizeIVSSIM
// (xSSIM.h:87, xIVSSIM.h:46; -dd only).
#pragma once
#include <array>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "ivqm_hip.h"

namespace ivqm {
namespace dropin {

using I32x4 = std::array<int32_t, 4>;
using F32x4 = std::array<float, 4>;
using F64x4 = std::array<double, 4>;
using U64x4 = std::array<uint64_t, 4>;
constexpr int32_t kNotValid = -1;  // the reference's NOT_VALID

struct Error : std::runtime_error {
  int status;
  Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};
inline void check(int s) {
  if (s != IVQM_OK) throw Error(s, std::string("ivqm: ") + ivqm_strerror(s) + ": " + ivqm_last_error());
}

namespace detail {
template <class V> I32x4 i32x4(const V& v) { return {(int32_t)v[0], (int32_t)v[1], (int32_t)v[2], (int32_t)v[3]}; }
template <class V> F32x4 f32x4(const V& v) { return {(float)v[0], (float)v[1], (float)v[2], (float)v[3]}; }
template <class P, class = void> struct has_noarg_addr : std::false_type {};
template <class P>
struct has_noarg_addr<P, std::void_t<decltype(std::declval<const P&>().getAddr())>> : std::true_type {};
}  // namespace detail

template <class Pic> struct PicTraits {
  static constexpr bool interleaved = detail::has_noarg_addr<Pic>::value;
  // planar: component c (0 Y, 1 Cb, 2 Cr) at the picture origin, getStride() samples per row
  static const uint16_t* plane(const Pic* p, int c) { return p->getAddr(c); }
  static uint16_t* plane_mut(Pic* p, int c) { return p->getAddr(c); }
  // interleaved: the first sample of the pixel at the origin, 4 * getStride() samples per row
  static const uint16_t* samples(const Pic* p) { return reinterpret_cast<const uint16_t*>(p->getAddr()); }
  static uint16_t* samples_mut(Pic* p) { return reinterpret_cast<uint16_t*>(p->getAddr()); }
};

// ---------------------------------------------------------------------------
// State shared by every processor: the parameters (the reference's Prms
// classes), the device context and the mask binding.
class Processor {
 protected:
  ivqm_config m_Cfg;
  ivqm_ctx* m_Ctx = nullptr;
  bool m_MaskBound = false;
  std::vector<uint16_t> m_Stage[2], m_Out;

  void xReset() {
    if (m_Ctx) ivqm_destroy(m_Ctx);
    m_Ctx = nullptr;
    m_MaskBound = false;
  }
  template <class Pic> void xEnsure(const Pic* p) {
    if (!p) throw Error(IVQM_EINVAL, "ivqm: NULL picture");
    if (m_Ctx && m_Cfg.width == p->getWidth() && m_Cfg.height == p->getHeight() && m_Cfg.bit_depth == p->getBitDepth())
      return;
    xReset();
    m_Cfg.width = p->getWidth();
    m_Cfg.height = p->getHeight();
    m_Cfg.bit_depth = p->getBitDepth();
    check(ivqm_create(&m_Cfg, &m_Ctx));
  }
  // Planar view of a picture: its own planes, or (interleaved) a planar copy in m_Stage[k].
  template <class Pic> ivqm_picture xView(const Pic* p, int k) {
    ivqm_picture v;
    v.chroma_420 = 0;
    v.bytes_per_sample = 2;
    const int w = p->getWidth(), h = p->getHeight(), s = p->getStride();
    if constexpr (PicTraits<Pic>::interleaved) {
      std::vector<uint16_t>& b = m_Stage[k];
      b.resize((size_t)3 * w * h);
      const uint16_t* src = PicTraits<Pic>::samples(p);
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++)
          for (int c = 0; c < 3; c++) b[((size_t)c * h + y) * w + x] = src[((size_t)y * s + x) * 4 + c];
      for (int c = 0; c < 3; c++) {
        v.plane[c] = b.data() + (size_t)c * w * h;
        v.stride[c] = w;
      }
    } else {
      for (int c = 0; c < 3; c++) {
        v.plane[c] = PicTraits<Pic>::plane(p, c);
        v.stride[c] = s;
      }
    }
    return v;
  }
  template <class PicT, class PicR> void xPair(const PicT* Tst, const PicR* Ref) {
    if (!Tst || !Ref) throw Error(IVQM_EINVAL, "ivqm: NULL picture");
    if (Tst->getWidth() != Ref->getWidth() || Tst->getHeight() != Ref->getHeight() ||
        Tst->getBitDepth() != Ref->getBitDepth())
      throw Error(IVQM_EINVAL, "ivqm: pictures differ in size or bit depth");
    xEnsure(Tst);
    ivqm_picture a = xView(Tst, 0), b = xView(Ref, 1);
    check(ivqm_upload(m_Ctx, 0, &a));
    check(ivqm_upload(m_Ctx, 1, &b));
  }
  void xNoMask() {
    if (!m_MaskBound) return;
    check(ivqm_bind_mask(m_Ctx, nullptr));
    m_MaskBound = false;
  }
  // Binds the mask (after xPair) and returns its non-zero count; NumNonMasked other than NOT_VALID
  // must equal it.
  template <class MskPic> int32_t xMask(const MskPic* Msk, int32_t NumNonMasked) {
    if (!Msk) throw Error(IVQM_EINVAL, "ivqm: NULL mask");
    if (Msk->getWidth() != m_Cfg.width || Msk->getHeight() != m_Cfg.height)
      throw Error(IVQM_EINVAL, "ivqm: mask size differs from the pictures'");
    const uint16_t* m = PicTraits<MskPic>::plane(Msk, 0);
    const int s = Msk->getStride();
    int64_t n = 0;
    for (int y = 0; y < m_Cfg.height; y++)
      for (int x = 0; x < m_Cfg.width; x++) n += m[(size_t)y * s + x] != 0;
    if (NumNonMasked != kNotValid && NumNonMasked != n)
      throw Error(IVQM_EINVAL, "ivqm: NumNonMasked differs from the mask's non-zero count (the device counts it)");
    ivqm_mask k;
    k.plane = m;
    k.stride = s;
    k.bytes_per_sample = 2;
    k.bit_depth = Msk->getBitDepth();
    k.reserved = 0;
    check(ivqm_upload_mask(m_Ctx, &k));
    m_MaskBound = true;
    return (int32_t)n;
  }

 public:
  Processor() {
    ivqm_config_init(&m_Cfg);
    m_Cfg.width = m_Cfg.height = 0;
  }
  Processor(const Processor&) = delete;
  Processor& operator=(const Processor&) = delete;
  virtual ~Processor() { xReset(); }

  // xCorrespPixelShiftPrms.h:30-32
  void setSearchRange(int32_t R) { m_Cfg.search_range = R; xReset(); }
  template <class V> void setCmpWeightsSearch(const V& w) {
    const I32x4 a = detail::i32x4(w);
    for (int i = 0; i < 4; i++) m_Cfg.cmp_weights_search[i] = a[i];
    xReset();
  }
  template <class V> void setCmpWeightsAverage(const V& w) {
    const I32x4 a = detail::i32x4(w);
    for (int i = 0; i < 4; i++) m_Cfg.cmp_weights_average[i] = a[i];
    xReset();
  }
  // xGlobClrDiff.h:24
  template <class V> void setUnntcbCoef(const V& c) {
    const F32x4 a = detail::f32x4(c);
    for (int i = 0; i < 4; i++) m_Cfg.unnoticeable_coef[i] = a[i];
    xReset();
  }
  // xPSNR.h:28
  void setFakeValsForExact(bool v) { m_Cfg.fake_vals_for_exact = v; xReset(); }
  // xWeightedSpherically.h:21 (the size and bit depth come from the pictures)
  void initWS(bool UseWS, int32_t /*Width*/, int32_t /*Height*/, int32_t /*BitDepth*/, int32_t LonRangeDeg = 360,
              int32_t LatRangeDeg = 180) {
    m_Cfg.erp = UseWS;
    m_Cfg.lon_range_deg = LonRangeDeg;
    m_Cfg.lat_range_deg = LatRangeDeg;
    xReset();
  }
  // xMetricCommon.h: device row buffers are sized per context
  void initRowBuffers(int32_t /*Height*/) {}
  // xCommonDefIVQM.h:42-55: the device stream replaces the thread pool
  template <class ThPI> bool bindThrdPoolIntf(ThPI*) { return true; }
  void* unbindThrdPoolIntf() { return nullptr; }
  ivqm_ctx* context() { return m_Ctx; }
};

// xGlobClrDiffProc (xGlobClrDiff.h:38-43). CalcGlobalColorDiff(Tst, Ref) with the app's
// (P0, P1) gives GCD_R2T (xAppQMIV.cpp:930-931).
class xGlobClrDiffProc : public Processor {
 public:
  template <class Pic> I32x4 CalcGlobalColorDiff(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    xNoMask();
    I32x4 g;
    check(ivqm_gcd(m_Ctx, g.data()));
    return g;
  }
  template <class Pic, class MskPic>
  I32x4 CalcGlobalColorDiffM(const Pic* Tst, const Pic* Ref, const MskPic* Msk, int32_t NumNonMasked) {
    xPair(Tst, Ref);
    xMask(Msk, NumNonMasked);
    I32x4 g;
    check(ivqm_gcd(m_Ctx, g.data()));
    return g;
  }
};

// xShftCompPicProc (xShftCompPic.h:36-51), as the app calls it (xAppQMIV.cpp:934-948):
// GenShftCompPics(SCP1, SCP0, P1, P0, GCD_R2T), on planar or interleaved pictures.
class xShftCompPicProc : public Processor {
 public:
  template <class Pic>
  void GenShftCompPics(Pic* DstRef, Pic* DstTst, const Pic* SrcRef, const Pic* SrcTst, const I32x4& GCD) {
    xGen<false>(DstRef, DstTst, SrcRef, SrcTst, GCD, SrcRef);
  }
  template <class Pic, class V>
  void GenShftCompPics(Pic* DstRef, Pic* DstTst, const Pic* SrcRef, const Pic* SrcTst, const V& GCD) {
    GenShftCompPics(DstRef, DstTst, SrcRef, SrcTst, detail::i32x4(GCD));
  }
  // Pixels whose mask weight is 0 keep the destinations' contents (xCorrespPixelShiftSTD.cpp:283).
  template <class Pic, class MskPic, class V>
  void GenShftCompPicsM(Pic* DstRef, Pic* DstTst, const Pic* SrcRef, const Pic* SrcTst, const MskPic* Msk,
                        const V& GCD) {
    xGen<true>(DstRef, DstTst, SrcRef, SrcTst, detail::i32x4(GCD), Msk);
  }

 private:
  template <bool Masked, class Pic, class MskPic>
  void xGen(Pic* DstRef, Pic* DstTst, const Pic* SrcRef, const Pic* SrcTst, const I32x4& GCD, const MskPic* Msk) {
    if (!DstRef || !DstTst) throw Error(IVQM_EINVAL, "ivqm: NULL destination picture");
    xPair(SrcTst, SrcRef);
    const uint16_t* m = nullptr;
    int ms = 0;
    if constexpr (Masked) {
      xMask(Msk, kNotValid);
      m = PicTraits<MskPic>::plane(Msk, 0);
      ms = Msk->getStride();
    } else {
      xNoMask();
    }
    check(ivqm_gen_scp(m_Ctx, GCD.data()));
    xDownload(1, DstRef, m, ms);
    xDownload(0, DstTst, m, ms);
  }
  template <class Pic> void xDownload(int idx, Pic* dst, const uint16_t* msk, int ms) {
    const int w = m_Cfg.width, h = m_Cfg.height, s = dst->getStride();
    if (dst->getWidth() != w || dst->getHeight() != h) throw Error(IVQM_EINVAL, "ivqm: destination size differs");
    if constexpr (!PicTraits<Pic>::interleaved) {
      if (!msk && s == w) {
        uint16_t* planes[3] = {PicTraits<Pic>::plane_mut(dst, 0), PicTraits<Pic>::plane_mut(dst, 1),
                               PicTraits<Pic>::plane_mut(dst, 2)};
        check(ivqm_get_scp(m_Ctx, idx, planes));
        return;
      }
    }
    m_Out.resize((size_t)3 * w * h);
    uint16_t* planes[3] = {m_Out.data(), m_Out.data() + (size_t)w * h, m_Out.data() + 2 * (size_t)w * h};
    check(ivqm_get_scp(m_Ctx, idx, planes));
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        if (msk && msk[(size_t)y * ms + x] == 0) continue;
        const size_t o = (size_t)y * w + x;
        if constexpr (PicTraits<Pic>::interleaved) {
          uint16_t* d = PicTraits<Pic>::samples_mut(dst) + ((size_t)y * s + x) * 4;
          d[0] = planes[0][o];
          d[1] = planes[1][o];
          d[2] = planes[2][o];
          d[3] = 0;  // clip(0 - 0): the unused 4th sample of the interleaved inputs is 0
        } else {
          for (int c = 0; c < 3; c++) PicTraits<Pic>::plane_mut(dst, c)[(size_t)y * s + x] = planes[c][o];
        }
      }
  }
};

// xPSNR (xPSNR.h:15-56).
class xPSNR : public Processor {
 public:
  using tDCfMSK = std::function<void(int32_t)>;

 protected:
  tDCfMSK m_DebugCallbackMSK;
  template <class Pic> U64x4 xSSD(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    xNoMask();
    U64x4 s;
    check(ivqm_ssd(m_Ctx, s.data()));
    return s;
  }

 public:
  void setDebugCallbackMSK(tDCfMSK f) { m_DebugCallbackMSK = std::move(f); }

  template <class Pic> U64x4 calcPicSSD(const Pic* Tst, const Pic* Ref) { return xSSD(Tst, Ref); }
  template <class Pic> F64x4 calcPicMSE(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    xNoMask();
    F64x4 v;
    check(ivqm_mse(m_Ctx, v.data()));
    return v;
  }
  template <class Pic> F64x4 calcPicPSNR(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    xNoMask();
    F64x4 v;
    check(ivqm_psnr(m_Ctx, v.data()));
    return v;
  }
  // xPSNR.cpp:36-66: SSDM; MSEM (no callback); PSNRM (callback MSK with NumNonMasked).
  template <class Pic, class MskPic> U64x4 calcPicSSDM(const Pic* Tst, const Pic* Ref, const MskPic* Msk) {
    xPair(Tst, Ref);
    xMask(Msk, kNotValid);
    U64x4 s;
    check(ivqm_ssd(m_Ctx, s.data()));
    return s;
  }
  template <class Pic, class MskPic>
  F64x4 calcPicMSEM(const Pic* Tst, const Pic* Ref, const MskPic* Msk, int32_t NumNonMasked = kNotValid) {
    xPair(Tst, Ref);
    xMask(Msk, NumNonMasked);
    F64x4 v;
    check(ivqm_mse(m_Ctx, v.data()));
    return v;
  }
  template <class Pic, class MskPic>
  F64x4 calcPicPSNRM(const Pic* Tst, const Pic* Ref, const MskPic* Msk, int32_t NumNonMasked = kNotValid) {
    xPair(Tst, Ref);
    const int32_t n = xMask(Msk, NumNonMasked);
    F64x4 v;
    check(ivqm_psnr(m_Ctx, v.data()));
    if (m_DebugCallbackMSK) m_DebugCallbackMSK(n);
    return v;
  }
  // xPSNR.h:55: PSNR of an SSD of 1, for components identical in both pictures
  static double getFakePSNR(int32_t Area, int32_t BitDepth) {
    const uint64_t maxv = (uint64_t)((1 << BitDepth) - 1);
    return 10.0 * std::log10((double)((uint64_t)Area * (maxv * maxv)) / 1.0);
  }
};

// xWSPSNR (xWSPSNR.h:15-28).
class xWSPSNR : public xPSNR {
 public:
  template <class Pic> F64x4 calcPicWSPSNR(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    xNoMask();
    F64x4 v;
    check(ivqm_wspsnr(m_Ctx, v.data()));
    return v;
  }
  // xWSPSNR.cpp:51-77: without ERP the value is calcPicPSNRM's (which reports MSK itself); MSK
  // is reported again at the end.
  template <class Pic, class MskPic>
  F64x4 calcPicWSPSNRM(const Pic* Tst, const Pic* Ref, const MskPic* Msk, int32_t NumNonMasked = kNotValid) {
    xPair(Tst, Ref);
    const int32_t n = xMask(Msk, NumNonMasked);
    F64x4 v;
    check(ivqm_wspsnr(m_Ctx, v.data()));
    if (m_DebugCallbackMSK) {
      if (!m_Cfg.erp) m_DebugCallbackMSK(n);
      m_DebugCallbackMSK(n);
    }
    return v;
  }
};

// xIVPSNR (xIVPSNR.h:17-42).
class xIVPSNR : public xWSPSNR {
 public:
  using tDCfGCS = std::function<void(const I32x4&)>;
  using tDCfQAP = std::function<void(double, double)>;
  static constexpr int32_t c_DefaultSearchRange = 2;  // xCorrespPixelShiftPrms.h:20

 protected:
  tDCfGCS m_DebugCallbackGCS;
  tDCfQAP m_DebugCallbackQAP;
  double xIV(const I32x4* GCD) {
    double v, r2t, t2r;
    check(ivqm_ivpsnr(m_Ctx, GCD ? GCD->data() : nullptr, &v, &r2t, &t2r));
    if (m_DebugCallbackQAP) m_DebugCallbackQAP(r2t, t2r);
    return v;
  }

 public:
  // Any callable taking the reference's int32V4 (or std::array) / (flt64, flt64).
  template <class F> void setDebugCallbackGCS(F f) { m_DebugCallbackGCS = tDCfGCS(std::move(f)); }
  template <class F> void setDebugCallbackQAP(F f) { m_DebugCallbackQAP = tDCfQAP(std::move(f)); }

  // xIVPSNR.cpp:18-29: GCD first (callback GCS), then IV-PSNR on the interleaved pictures when given.
  template <class Pic, class PicI = Pic>
  double calcPicIVPSNR(const Pic* Tst, const Pic* Ref, const PicI* TstI = nullptr, const PicI* RefI = nullptr) {
    xPair(Tst, Ref);
    xNoMask();
    I32x4 g;
    check(ivqm_gcd(m_Ctx, g.data()));
    if (m_DebugCallbackGCS) m_DebugCallbackGCS(g);
    if (TstI && RefI) xPair(TstI, RefI);
    return xIV(&g);
  }
  // xIVPSNR.cpp:31-57, planar or interleaved pictures.
  template <class Pic> double calcPicIVPSNR(const Pic* Tst, const Pic* Ref, const I32x4& GCD) {
    xPair(Tst, Ref);
    xNoMask();
    return xIV(&GCD);
  }
  template <class Pic, class V,
            class = std::enable_if_t<!std::is_pointer<V>::value && !std::is_same<V, I32x4>::value>>
  double calcPicIVPSNR(const Pic* Tst, const Pic* Ref, const V& GCD) {
    return calcPicIVPSNR(Tst, Ref, detail::i32x4(GCD));
  }
};

// xIVPSNRM (xIVPSNR.h:46-60): R2T is the direction that searches P1 - GCD in P0
// (xIVPSNR.cpp:161-162).
class xIVPSNRM : public xIVPSNR {
 public:
  // xIVPSNR.cpp:152-159 + :163: NumNonMasked (callback MSK), masked GCD (callback GCS), IV-PSNR.
  template <class Pic, class MskPic, class PicI>
  double calcPicIVPSNRM(const Pic* Tst, const Pic* Ref, const MskPic* Msk, const PicI* TstI, const PicI* RefI) {
    xPair(Tst, Ref);
    const int32_t n = xMask(Msk, kNotValid);
    if (m_DebugCallbackMSK) m_DebugCallbackMSK(n);
    I32x4 g;
    check(ivqm_gcd(m_Ctx, g.data()));
    if (m_DebugCallbackGCS) m_DebugCallbackGCS(g);
    if (TstI && RefI) {
      xPair(TstI, RefI);  // same size: the context and its mask stay
      if (!m_MaskBound) xMask(Msk, n);
    }
    return xIV(&g);
  }
  // xIVPSNR.cpp:161-176, as xAppQMIV.cpp:1026 calls it (interleaved pictures, planar mask).
  template <class Pic, class MskPic, class V>
  double calcPicIVPSNRM(const Pic* Tst, const Pic* Ref, const MskPic* Msk, int32_t NumNonMasked, const V& GCD) {
    xPair(Tst, Ref);
    xMask(Msk, NumNonMasked);
    const I32x4 g = detail::i32x4(GCD);
    return xIV(&g);
  }
};

// xSSIM (xSSIM.h:20-95) and xIVSSIM (xIVSSIM.h:15-50).
class xSSIM : public Processor {
 public:
  enum class eMode {  // xSSIM.h:24-31
    INVALID = kNotValid,
    RegularGaussianFlt = 0,
    RegularGaussianInt = 1,
    RegularAveraged = 2,
    BlockGaussianInt = 3,
    BlockAveraged = 4,
  };
  static constexpr eMode c_DefaultStructSimMode = eMode::BlockAveraged;
  static constexpr int32_t c_DefaultStructSimStride = 4;
  static constexpr int32_t c_DefaultStructSimWindow = 8;

  // Mode names, compared case-insensitively (xSSIM.cpp:17-35).
  static eMode xStrToMode(const std::string& Mode) {
    std::string s = Mode;
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    static const char* const names[] = {"regulargaussianflt", "regulargaussianint", "regularaveraged",
                                        "blockgaussianint", "blockaveraged"};
    for (int i = 0; i < 5; i++)
      if (s == names[i]) return (eMode)i;
    return eMode::INVALID;
  }
  static std::string xModeToStr(eMode Mode) {
    static const char* const names[] = {"RegularGaussianFlt", "RegularGaussianInt", "RegularAveraged",
                                        "BlockGaussianInt", "BlockAveraged"};
    const int i = (int)Mode;
    return i >= 0 && i < 5 ? names[i] : "INVALID";
  }
  static bool isRegularMode(eMode Mode) { return (int)Mode >= 0 && (int)Mode <= 2; }
  // regular modes use an 11 x 11 window (xSSIM.h:91)
  static int32_t determineWindowSize(eMode Mode, int32_t BlockSize) { return isRegularMode(Mode) ? 11 : BlockSize; }

  // xSSIM.h:79-80: the device sizes its buffers per context
  template <class V2> void create(const V2& /*PicSize*/, int32_t /*BitDepth*/, int32_t /*Margin*/, bool /*EnableMS*/) {}
  void destroy() { xReset(); }
  // xSSIM.h:82: Mode = eMode (or its int value), MarginMode = the reference's eMrgExt
  // (xCommonDefCORE.h:322-331, passed through by value).
  template <class ModeT, class ExtT> void setStructSimParams(ModeT Mode, ExtT MarginMode, int32_t Window, int32_t Stride) {
    m_Cfg.ssim_mode = (int32_t)Mode;
    m_Cfg.ssim_border = (int32_t)MarginMode;
    m_Cfg.ssim_window = Window;
    m_Cfg.ssim_stride = Stride;
    xReset();
  }
  template <class Pic> F64x4 calcPicSSIM(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    xNoMask();
    F64x4 v;
    check(ivqm_ssim(m_Ctx, v.data()));
    return v;
  }
  template <class Pic> F64x4 calcPicMSSSIM(const Pic* Tst, const Pic* Ref) {
    xPair(Tst, Ref);
    xNoMask();
    F64x4 v;
    check(ivqm_msssim(m_Ctx, v.data()));
    return v;
  }
  template <class Pic, class MskPic>
  F64x4 calcPicSSIMM(const Pic*, const Pic*, const MskPic*, int32_t = kNotValid) {
    throw Error(IVQM_EINVAL, "ivqm: structural similarity cannot be combined with a mask");
  }
  template <class Plane, class Pic, class CmpT> void visualizeSSIM(Plane*, const Pic*, const Pic*, CmpT) {
    throw Error(IVQM_EINVAL, "ivqm: the SSIM debug dump is not provided");
  }
};

class xIVSSIM : public xSSIM {
 public:
  using tDCfGCS = std::function<void(const I32x4&)>;
  using tDCfQAP = std::function<void(double, double)>;

 protected:
  tDCfGCS m_DebugCallbackGCS;  // reported by the reference only when it generates the SCP itself
  tDCfQAP m_DebugCallbackQAP;
  template <class Pic> void xWithScp(const Pic* Tst, const Pic* Ref, const Pic* TstSCP, const Pic* RefSCP) {
    if (!TstSCP || !RefSCP) throw Error(IVQM_EINVAL, "ivqm: IV-SSIM needs both shift-compensated pictures");
    xPair(Tst, Ref);
    xNoMask();
    ivqm_picture a = xView(TstSCP, 0), b = xView(RefSCP, 1);
    check(ivqm_upload_scp(m_Ctx, 0, &a));
    check(ivqm_upload_scp(m_Ctx, 1, &b));
  }

 public:
  template <class F> void setDebugCallbackGCS(F f) { m_DebugCallbackGCS = tDCfGCS(std::move(f)); }
  template <class F> void setDebugCallbackQAP(F f) { m_DebugCallbackQAP = tDCfQAP(std::move(f)); }

  // calcPicIVSSIM(P0, P1, SCP0, SCP1) as xAppQMIV.cpp:1077 calls it; callback QAP(R2T, T2R) (xIVSSIM.cpp:58).
  template <class Pic> double calcPicIVSSIM(const Pic* Tst, const Pic* Ref, const Pic* TstSCP, const Pic* RefSCP) {
    xWithScp(Tst, Ref, TstSCP, RefSCP);
    double v, r2t, t2r;
    check(ivqm_ivssim(m_Ctx, &v, &r2t, &t2r));
    if (m_DebugCallbackQAP) m_DebugCallbackQAP(r2t, t2r);
    return v;
  }
  template <class Pic> double calcPicIVMSSSIM(const Pic* Tst, const Pic* Ref, const Pic* TstSCP, const Pic* RefSCP) {
    xWithScp(Tst, Ref, TstSCP, RefSCP);
    double v, r2t, t2r;
    check(ivqm_ivmsssim(m_Ctx, &v, &r2t, &t2r));
    if (m_DebugCallbackQAP) m_DebugCallbackQAP(r2t, t2r);
    return v;
  }
  template <class Pic, class MskPic>
  double calcPicIVSSIMM(const Pic*, const Pic*, const Pic*, const Pic*, const MskPic*, int32_t = kNotValid) {
    throw Error(IVQM_EINVAL, "ivqm: structural similarity cannot be combined with a mask");
  }
  template <class Plane, class Pic, class CmpT>
  void visualizeIVSSIM(Plane*, const Pic*, const Pic*, const Pic*, const Pic*, CmpT) {
    throw Error(IVQM_EINVAL, "ivqm: the IV-SSIM debug dump is not provided");
  }
};

}  // namespace dropin
}  // namespace ivqm
