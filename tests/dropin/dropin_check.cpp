// Exercises include/ivqm_dropin.hpp the way the QMIV app drives the reference
// processors: processor set-up as in xAppQMIV::createProcessors
// (apps/QMIV/src/xAppQMIV.cpp:680-744) and the per-frame calls of :927-1108,
// with picture and vector types shaped like the reference's (a planar picture
// whose getAddr takes a component enum, an interleaved picture of 4-sample
// pixels with its stride in pixels, a 4-vector built implicitly from
// std::array). Without a mask it runs GCD, SCP (interleaved), MSE / PSNR /
// WS-PSNR, IV-PSNR (planar, interleaved and the form that computes the GCD
// itself), SSIM, MS-SSIM, IV-SSIM and IV-MS-SSIM; with a mask the masked
// calls of the app's mask mode (GCD, SCP, MSE, PSNR, WS-PSNR, IV-PSNR). The
// debug callbacks (QAP, GCS, MSK) record what they receive. Results go to
// stdout as JSON (repr-exact doubles) and the SCP pictures to out_dir as raw
// 4:4:4 uint16 files, for tests/test_gpu_dropin.py.
//
// usage: ivqm_dropin_check w h bd p0.raw p1.raw erp out_dir [mask.raw bdm]
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ivqm_dropin.hpp"

namespace app {

// The reference's xVec4 in the respects the processors rely on: operator[] and an implicit
// constructor from std::array (libs/pmbb-core/src/xVec.h:262-283).
template <class T> struct Vec4 {
  std::array<T, 4> v{};
  Vec4() = default;
  Vec4(T a, T b, T c, T d) : v{a, b, c, d} {}
  Vec4(const std::array<T, 4>& a) : v(a) {}  // NOLINT: implicit, as xVec4's
  T& operator[](int i) { return v[i]; }
  const T& operator[](int i) const { return v[i]; }
};
using int32V4 = Vec4<int32_t>;
using flt32V4 = Vec4<float>;
using flt64V4 = Vec4<double>;
using uint16V4 = Vec4<uint16_t>;

enum class eCmp : int32_t { LM = 0, C0 = 0, C1 = 1, C2 = 2 };

// Planar picture with a margin (the reference's xPicP accessors).
struct PicP {
  int w = 0, h = 0, bd = 0, margin = 0, stride = 0;
  std::vector<uint16_t> buf[3];
  size_t org = 0;
  void create(int w_, int h_, int bd_, int margin_, uint16_t fill = 0) {
    w = w_, h = h_, bd = bd_, margin = margin_, stride = w_ + 2 * margin_;
    for (auto& b : buf) b.assign((size_t)stride * (h + 2 * margin), fill);
    org = (size_t)margin * stride + margin;
  }
  int getWidth() const { return w; }
  int getHeight() const { return h; }
  int getArea() const { return w * h; }
  int getBitDepth() const { return bd; }
  int getStride() const { return stride; }
  const uint16_t* getAddr(eCmp c) const { return buf[(int)c].data() + org; }
  uint16_t* getAddr(eCmp c) { return buf[(int)c].data() + org; }
};

// Interleaved picture: 4 samples per pixel, stride in pixels (the reference's xPicI accessors).
struct PicI {
  int w = 0, h = 0, bd = 0, margin = 0, stride = 0;
  std::vector<uint16V4> buf;
  size_t org = 0;
  void create(int w_, int h_, int bd_, int margin_, uint16_t fill = 0) {
    w = w_, h = h_, bd = bd_, margin = margin_, stride = w_ + 2 * margin_;
    buf.assign((size_t)stride * (h + 2 * margin), uint16V4(fill, fill, fill, 0));
    org = (size_t)margin * stride + margin;
  }
  void rearrangeFromPlanar(const PicP& p) {
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++) {
        uint16V4& d = buf[org + (size_t)y * stride + x];
        for (int c = 0; c < 3; c++) d[c] = p.getAddr((eCmp)c)[(size_t)y * p.stride + x];
        d[3] = 0;
      }
  }
  void rearrangeToPlanar(PicP& p) const {
    for (int y = 0; y < h; y++)
      for (int x = 0; x < w; x++)
        for (int c = 0; c < 3; c++) p.getAddr((eCmp)c)[(size_t)y * p.stride + x] = buf[org + (size_t)y * stride + x][c];
  }
  int getWidth() const { return w; }
  int getHeight() const { return h; }
  int getBitDepth() const { return bd; }
  int getStride() const { return stride; }
  const uint16V4* getAddr() const { return buf.data() + org; }
  uint16V4* getAddr() { return buf.data() + org; }
};

struct ThreadPoolIntf {};

}  // namespace app

// The planar picture's getAddr takes the component enum.
template <> struct ivqm::dropin::PicTraits<app::PicP> {
  static constexpr bool interleaved = false;
  static const uint16_t* plane(const app::PicP* p, int c) { return p->getAddr((app::eCmp)c); }
  static uint16_t* plane_mut(app::PicP* p, int c) { return p->getAddr((app::eCmp)c); }
};

namespace app {
using namespace ivqm::dropin;

static void load(PicP& p, const char* path, int ncmp) {
  FILE* f = std::fopen(path, "rb");
  if (!f) { std::perror(path); std::exit(2); }
  for (int c = 0; c < ncmp; c++)
    for (int y = 0; y < p.h; y++)
      if (std::fread(p.getAddr((eCmp)c) + (size_t)y * p.stride, 2, p.w, f) != (size_t)p.w) std::exit(3);
  std::fclose(f);
  for (int c = ncmp; c < 3; c++)  // a mask read as luma only (the app's mask is a picture too)
    for (int y = 0; y < p.h; y++)
      for (int x = 0; x < p.w; x++) p.getAddr((eCmp)c)[(size_t)y * p.stride + x] = p.getAddr(eCmp::LM)[(size_t)y * p.stride + x];
}
static void store(const PicP& p, const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  for (int c = 0; c < 3; c++)
    for (int y = 0; y < p.h; y++) std::fwrite(p.getAddr((eCmp)c) + (size_t)y * p.stride, 2, p.w, f);
  std::fclose(f);
}
static std::string v3(const flt64V4& v) {
  char b[128];
  std::snprintf(b, sizeof b, "[%.17g, %.17g, %.17g]", v[0], v[1], v[2]);
  return b;
}
static std::string i4(const int32V4& v) {
  char b[96];
  std::snprintf(b, sizeof b, "[%d, %d, %d, %d]", v[0], v[1], v[2], v[3]);
  return b;
}
static std::string d(double x) {
  char b[40];
  std::snprintf(b, sizeof b, "%.17g", x);
  return b;
}

// The app's state, processors and set-up (xAppQMIV.h:125-145, xAppQMIV.cpp:680-744).
struct App {
  int W, H, BitDepth, PicMargin = 4;
  bool UseMask, IsEquirectangular, CalcMSs = true;
  int32_t SearchRange = xIVPSNR::c_DefaultSearchRange;
  int32V4 CmpWeightsSearch{4, 1, 1, 0}, CmpWeightsAverage{4, 1, 1, 0};
  flt32V4 UnnoticeableCoef{0.01f, 0.01f, 0.01f, 0.0f};
  xSSIM::eMode StructSimMode = xSSIM::c_DefaultStructSimMode;
  int StructSimBrdExt = 0;
  int32_t StructSimWindow = xSSIM::determineWindowSize(xSSIM::c_DefaultStructSimMode, xSSIM::c_DefaultStructSimWindow);
  int32_t StructSimStride = xSSIM::c_DefaultStructSimStride;
  ThreadPoolIntf TPI;

  PicP PicInP[3], PicSCP[2];
  PicI PicInI[2], PicSCI[2];
  xGlobClrDiffProc ProcGCD;
  xShftCompPicProc ProcSCP;
  xIVPSNRM ProcPSNR;
  xIVSSIM ProcSSIM;
  int32V4 GCD_R2T;
  int32_t NumNonMasked = 0;
  double LastR2T = 0, LastT2R = 0;
  std::vector<int32_t> MskReports;
  std::vector<int32V4> GcsReports;

  void createProcessors() {
    ProcGCD.setUnntcbCoef(UnnoticeableCoef);
    ProcGCD.bindThrdPoolIntf(&TPI);
    ProcSCP.setSearchRange(SearchRange);
    ProcSCP.setCmpWeightsSearch(CmpWeightsSearch);
    ProcSCP.setCmpWeightsAverage(CmpWeightsAverage);
    ProcSCP.bindThrdPoolIntf(&TPI);
    ProcPSNR.setSearchRange(SearchRange);
    ProcPSNR.setCmpWeightsSearch(CmpWeightsSearch);
    ProcPSNR.setCmpWeightsAverage(CmpWeightsAverage);
    ProcPSNR.setUnntcbCoef(UnnoticeableCoef);
    ProcPSNR.bindThrdPoolIntf(&TPI);
    ProcPSNR.initRowBuffers(H);
    if (IsEquirectangular) ProcPSNR.initWS(true, W, H, BitDepth, 360, 180);
    if (!UseMask) {
      ProcSSIM.create(std::array<int32_t, 2>{W, H}, BitDepth, PicMargin, CalcMSs);
      ProcSSIM.setSearchRange(SearchRange);
      ProcSSIM.setCmpWeightsSearch(CmpWeightsSearch);
      ProcSSIM.setCmpWeightsAverage(CmpWeightsAverage);
      ProcSSIM.setUnntcbCoef(UnnoticeableCoef);
      ProcSSIM.setStructSimParams(StructSimMode, StructSimBrdExt, StructSimWindow, StructSimStride);
      ProcSSIM.bindThrdPoolIntf(&TPI);
      ProcSSIM.initRowBuffers(H);
      if (IsEquirectangular) ProcSSIM.initWS(true, W, H, BitDepth, 360, 180);
    }
    ProcPSNR.setDebugCallbackQAP([this](double R2T, double T2R) { LastR2T = R2T; LastT2R = T2R; });
    ProcSSIM.setDebugCallbackQAP([this](double R2T, double T2R) { LastR2T = R2T; LastT2R = T2R; });
    ProcPSNR.setDebugCallbackMSK([this](int32_t n) { MskReports.push_back(n); });
    ProcPSNR.setDebugCallbackGCS([this](const int32V4& g) { GcsReports.push_back(g); });
  }
};

}  // namespace app

int main(int argc, char** argv) {
  using namespace app;
  if (argc < 8) {
    std::fprintf(stderr, "usage: w h bd p0.raw p1.raw erp out_dir [mask.raw bdm]\n");
    return 1;
  }
  App A;
  A.W = std::atoi(argv[1]);
  A.H = std::atoi(argv[2]);
  A.BitDepth = std::atoi(argv[3]);
  A.IsEquirectangular = std::atoi(argv[6]) != 0;
  const std::string out = argv[7];
  A.UseMask = argc > 9;
  const int W = A.W, H = A.H, bd = A.BitDepth;
  for (int i = 0; i < 2; i++) {
    A.PicInP[i].create(W, H, bd, A.PicMargin);
    A.PicSCP[i].create(W, H, bd, A.PicMargin);
    A.PicInI[i].create(W, H, bd, A.PicMargin);
    A.PicSCI[i].create(W, H, bd, A.PicMargin, 7);  // masked pixels keep 7
  }
  load(A.PicInP[0], argv[4], 3);
  load(A.PicInP[1], argv[5], 3);
  if (A.UseMask) {
    A.PicInP[2].create(W, H, std::atoi(argv[9]), A.PicMargin);
    load(A.PicInP[2], argv[8], 1);
    for (int y = 0; y < H; y++)
      for (int x = 0; x < W; x++) A.NumNonMasked += A.PicInP[2].getAddr(eCmp::LM)[(size_t)y * A.PicInP[2].stride + x] != 0;
  }
  for (int i = 0; i < 2; i++) A.PicInI[i].rearrangeFromPlanar(A.PicInP[i]);
  try {
    A.createProcessors();
    std::string js = "{";
    auto put = [&](const std::string& k, const std::string& v) { js += (js.size() > 1 ? ", \"" : "\"") + k + "\": " + v; };
    PicP* P = A.PicInP;
    if (A.UseMask) {
      // xAppQMIV.cpp:930, :939, :953, :971, :999, :1026 (mask mode)
      A.GCD_R2T = A.ProcGCD.CalcGlobalColorDiffM(&P[0], &P[1], &P[2], A.NumNonMasked);
      A.ProcSCP.GenShftCompPicsM(&A.PicSCI[1], &A.PicSCI[0], &A.PicInI[1], &A.PicInI[0], &P[2], A.GCD_R2T);
      for (int i = 0; i < 2; i++) A.PicSCI[i].rearrangeToPlanar(A.PicSCP[i]);
      flt64V4 MSE = A.ProcPSNR.calcPicMSEM(&P[0], &P[1], &P[2], A.NumNonMasked);
      flt64V4 PSNR = A.ProcPSNR.calcPicPSNRM(&P[0], &P[1], &P[2], A.NumNonMasked);
      flt64V4 WSPSNR = A.ProcPSNR.calcPicWSPSNRM(&P[0], &P[1], &P[2], A.NumNonMasked);
      double IVPSNR = A.ProcPSNR.calcPicIVPSNRM(&A.PicInI[0], &A.PicInI[1], &P[2], A.NumNonMasked, A.GCD_R2T);
      put("gcd", i4(A.GCD_R2T));
      put("mse", v3(MSE));
      put("psnr", v3(PSNR));
      put("wspsnr", v3(WSPSNR));
      put("ivpsnr", d(IVPSNR));
      put("r2t", d(A.LastR2T));
      put("t2r", d(A.LastT2R));
      // the form that counts the mask and computes the masked GCD itself (xIVPSNR.cpp:152-159)
      A.MskReports.push_back(-1000);  // separator
      double IVPSNR2 = A.ProcPSNR.calcPicIVPSNRM(&P[0], &P[1], &P[2], &A.PicInI[0], &A.PicInI[1]);
      put("ivpsnr_auto", d(IVPSNR2));
      put("nnm", std::to_string(A.NumNonMasked));
      // a wrong NumNonMasked is refused
      bool refused = false;
      try {
        A.ProcPSNR.calcPicPSNRM(&P[0], &P[1], &P[2], A.NumNonMasked + 1);
      } catch (const Error&) {
        refused = true;
      }
      put("wrong_nnm_refused", refused ? "true" : "false");
    } else {
      A.GCD_R2T = A.ProcGCD.CalcGlobalColorDiff(&P[0], &P[1]);
      A.ProcSCP.GenShftCompPics(&A.PicSCI[1], &A.PicSCI[0], &A.PicInI[1], &A.PicInI[0], A.GCD_R2T);
      for (int i = 0; i < 2; i++) A.PicSCI[i].rearrangeToPlanar(A.PicSCP[i]);
      // the planar SCP form as well (xAppQMIV.cpp:946), into separate pictures
      PicP S0, S1;
      S0.create(W, H, bd, 2);
      S1.create(W, H, bd, 2);
      A.ProcSCP.GenShftCompPics(&S1, &S0, &P[1], &P[0], A.GCD_R2T);
      bool same = true;
      for (int c = 0; c < 3; c++)
        for (int y = 0; y < H; y++)
          for (int x = 0; x < W; x++) {
            same &= S0.getAddr((eCmp)c)[(size_t)y * S0.stride + x] == A.PicSCP[0].getAddr((eCmp)c)[(size_t)y * A.PicSCP[0].stride + x];
            same &= S1.getAddr((eCmp)c)[(size_t)y * S1.stride + x] == A.PicSCP[1].getAddr((eCmp)c)[(size_t)y * A.PicSCP[1].stride + x];
          }
      put("scp_planar_equals_interleaved", same ? "true" : "false");
      flt64V4 MSE = A.ProcPSNR.calcPicMSE(&P[0], &P[1]);
      flt64V4 PSNR = A.ProcPSNR.calcPicPSNR(&P[0], &P[1]);
      flt64V4 WSPSNR = A.ProcPSNR.calcPicWSPSNR(&P[0], &P[1]);
      double IV_I = A.ProcPSNR.calcPicIVPSNR(&A.PicInI[0], &A.PicInI[1], A.GCD_R2T);
      const double qr = A.LastR2T, qt = A.LastT2R;
      double IV_P = A.ProcPSNR.calcPicIVPSNR(&P[0], &P[1], A.GCD_R2T);
      double IV_auto = A.ProcPSNR.calcPicIVPSNR(&P[0], &P[1], &A.PicInI[0], &A.PicInI[1]);
      put("gcd", i4(A.GCD_R2T));
      put("mse", v3(MSE));
      put("psnr", v3(PSNR));
      put("wspsnr", v3(WSPSNR));
      put("ivpsnr", d(IV_I));
      put("r2t", d(qr));
      put("t2r", d(qt));
      put("ivpsnr_planar", d(IV_P));
      put("ivpsnr_auto", d(IV_auto));
      put("fake_psnr", d(xPSNR::getFakePSNR(P[0].getArea(), P[0].getBitDepth())));
      flt64V4 SSIM = A.ProcSSIM.calcPicSSIM(&P[0], &P[1]);
      flt64V4 MSSSIM = A.ProcSSIM.calcPicMSSSIM(&P[0], &P[1]);
      double IVSSIM = A.ProcSSIM.calcPicIVSSIM(&P[0], &P[1], &A.PicSCP[0], &A.PicSCP[1]);
      const double sr = A.LastR2T, st = A.LastT2R;
      double IVMSSSIM = A.ProcSSIM.calcPicIVMSSSIM(&P[0], &P[1], &A.PicSCP[0], &A.PicSCP[1]);
      put("ssim", v3(SSIM));
      put("msssim", v3(MSSSIM));
      put("ivssim", d(IVSSIM));
      put("ivssim_r2t", d(sr));
      put("ivssim_t2r", d(st));
      put("ivmsssim", d(IVMSSSIM));
      put("ivmsssim_r2t", d(A.LastR2T));
      put("ivmsssim_t2r", d(A.LastT2R));
      bool refused = false;
      try {
        A.ProcSSIM.calcPicIVSSIMM(&P[0], &P[1], &A.PicSCP[0], &A.PicSCP[1], &P[0], 1);
      } catch (const Error&) {
        refused = true;
      }
      put("ssim_mask_refused", refused ? "true" : "false");
      put("mode_names", std::string("[\"") + xSSIM::xModeToStr(xSSIM::xStrToMode("regularGAUSSIANint")) + "\", \"" +
                            xSSIM::xModeToStr(xSSIM::xStrToMode("nonsense")) + "\"]");
    }
    std::string msk = "[";
    for (size_t i = 0; i < A.MskReports.size(); i++) msk += (i ? ", " : "") + std::to_string(A.MskReports[i]);
    put("msk_reports", msk + "]");
    std::string gcs = "[";
    for (size_t i = 0; i < A.GcsReports.size(); i++) gcs += (i ? ", " : "") + i4(A.GcsReports[i]);
    put("gcs_reports", gcs + "]");
    std::printf("%s}\n", js.c_str());
    store(A.PicSCP[0], out + "/scp0.raw");
    store(A.PicSCP[1], out + "/scp1.raw");
  } catch (const Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 4;
  }
  return 0;
}
