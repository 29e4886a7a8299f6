// Exercises include/ivqm_dropin.hpp the way the QMIV frame loop calls the
// reference processors (apps/QMIV/src/xAppQMIV.cpp:927-1108): reads two raw
// 4:4:4 uint16 pictures, runs GCD -> SCP -> PSNR / WS-PSNR / IV-PSNR -> SSIM /
// MS-SSIM / IV-SSIM / IV-MS-SSIM through the drop-in classes and prints the
// results as JSON (repr-exact doubles) for tests/test_gpu_dropin.py.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ivqm_dropin.hpp"

// A minimal planar picture with the accessor names of the reference's xPicP.
struct Pic {
  int w, h, bd, stride;
  std::vector<uint16_t> buf[3];
  Pic(int w_, int h_, int bd_, int margin) : w(w_), h(h_), bd(bd_), stride(w_ + 2 * margin) {
    for (auto& b : buf) b.assign((size_t)stride * (h + 2 * margin), 0);
    org = margin * stride + margin;
  }
  size_t org;
  int getWidth() const { return w; }
  int getHeight() const { return h; }
  int getBitDepth() const { return bd; }
  int getStride() const { return stride; }
  const uint16_t* getAddr(int c) const { return buf[c].data() + org; }
  uint16_t* getAddr(int c) { return buf[c].data() + org; }
};

static void load(Pic& p, const char* path) {
  FILE* f = std::fopen(path, "rb");
  if (!f) { std::perror(path); std::exit(2); }
  for (int c = 0; c < 3; c++)
    for (int y = 0; y < p.h; y++)
      if (std::fread(p.getAddr(c) + (size_t)y * p.stride, 2, p.w, f) != (size_t)p.w) { std::exit(3); }
  std::fclose(f);
}

int main(int argc, char** argv) {
  if (argc < 6) { std::fprintf(stderr, "usage: w h bd p0.raw p1.raw [erp]\n"); return 1; }
  const int w = std::atoi(argv[1]), h = std::atoi(argv[2]), bd = std::atoi(argv[3]);
  const bool erp = argc > 6 && std::atoi(argv[6]);
  Pic P0(w, h, bd, 4), P1(w, h, bd, 4), SCP0(w, h, bd, 4), SCP1(w, h, bd, 4);
  load(P0, argv[4]);
  load(P1, argv[5]);
  using namespace ivqm::dropin;
  try {
    xGlobClrDiffProc gcd;
    int32V4 g = gcd.CalcGlobalColorDiff(&P0, &P1);
    xShftCompPicProc scp;
    scp.GenShftCompPics(&SCP1, &SCP0, &P1, &P0, g);
    xIVPSNR psnr;
    psnr.initRowBuffers(h);
    if (erp) psnr.initWS(true, w, h, bd, 360, 180);
    flt64V4 ps = psnr.calcPicPSNR(&P0, &P1);
    flt64V4 ws = psnr.calcPicWSPSNR(&P0, &P1);
    double iv = psnr.calcPicIVPSNR(&P0, &P1, g);
    xIVSSIM ssim;
    ssim.create({w, h}, bd, 4, true);
    ssim.setStructSimParams(4, 0, 8, 4);
    if (erp) ssim.initWS(true, w, h, bd, 360, 180);
    flt64V4 ss = ssim.calcPicSSIM(&P0, &P1);
    flt64V4 ms = ssim.calcPicMSSSIM(&P0, &P1);
    double ivs = ssim.calcPicIVSSIM(&P0, &P1, &SCP0, &SCP1);
    double ivms = ssim.calcPicIVMSSSIM(&P0, &P1, &SCP0, &SCP1);
    unsigned long long scp_sum = 0;
    for (int c = 0; c < 3; c++)
      for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) scp_sum += SCP0.getAddr(c)[(size_t)y * SCP0.stride + x] * 3ull + SCP1.getAddr(c)[(size_t)y * SCP1.stride + x];
    std::printf("{\"gcd\": [%d, %d, %d, %d], \"psnr\": [%.17g, %.17g, %.17g], \"wspsnr\": [%.17g, %.17g, %.17g], "
                "\"ivpsnr\": %.17g, \"r2t\": %.17g, \"t2r\": %.17g, \"ssim\": [%.17g, %.17g, %.17g], "
                "\"msssim\": [%.17g, %.17g, %.17g], \"ivssim\": %.17g, \"ivmsssim\": %.17g, \"scp_sum\": %llu}\n",
                g[0], g[1], g[2], g[3], ps[0], ps[1], ps[2], ws[0], ws[1], ws[2], iv, psnr.LastR2T, psnr.LastT2R, ss[0],
                ss[1], ss[2], ms[0], ms[1], ms[2], ivs, ivms, scp_sum);
  } catch (const Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 4;
  }
  return 0;
}
