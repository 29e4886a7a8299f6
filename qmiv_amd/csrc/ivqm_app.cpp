// ivqm_qmiv: the QMIV command-line front end over the batch ABI.
//
// Reads two raw planar YUV files, computes the requested metrics of every
// frame pair on the GPU and prints the per-frame lines and the averages the
// reference app prints:
//   options            apps/QMIV/src/xAppQMIV.cpp:40-122 (help), :160-200, :217-340
//   frame loop         apps/QMIV/src/xAppQMIV.cpp:756-832
//   per-frame lines    apps/QMIV/src/xAppQMIV.cpp:950-1110
//   averages, results  apps/QMIV/src/xAppQMIV.cpp:1118-1160, xMetricQMIV.cpp:44-160
//
// Pipeline: reader threads fill a ring of pinned host slots with frame pairs
// read from the files; each slot is copied to a device slot on a copy stream;
// groups of up to 8 pairs run as one batch (ivqm_batch_enqueue) on the compute
// stream once their copies are done, and a group's results are collected after
// the next group is enqueued, so reading, copying and computing overlap.
//
// Also built: PNG / BMP picture lists (-ff), mask mode (-im -bdm -cfm), colour
// spaces (-csi -csm), every SSIM mode, window, stride and border extension
// (-ssm -ssw -sss -ssb), 4:2:0, 4:2:2 and 4:4:4 input (-cf, -pf). Not supported
// (the reference supports it): config files (-c); rejected with a message.
#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/ivqm_batch.h"
#include "ivqm_epilogue.h"
#include "ivqm_images.h"

namespace {

enum Metric { kMSE, kPSNR, kWSPSNR, kIVPSNR, kSSIM, kMSSSIM, kIVSSIM, kIVMSSSIM, kNumMetrics };
// apps/QMIV/src/xMetricQMIV.h:57-66: name, default, per-component, per-picture, normalized, unit
struct MetricInfo {
  const char* name;
  bool is_default, per_cmp, per_pic, normalized;
  const char* unit;
  uint32_t bit;
};
const MetricInfo kInfo[kNumMetrics] = {
    {"MSE", false, true, false, false, "  ", IVQM_M_MSE},
    {"PSNR", true, true, false, false, "dB", IVQM_M_PSNR},
    {"WSPSNR", false, true, false, false, "dB", IVQM_M_WSPSNR},
    {"IVPSNR", true, false, true, false, "dB", IVQM_M_IVPSNR},
    {"SSIM", false, true, false, true, "  ", IVQM_M_SSIM},
    {"MSSSIM", false, true, false, true, "  ", IVQM_M_MSSSIM},
    {"IVSSIM", true, false, true, true, "  ", IVQM_M_IVSSIM},
    {"IVMSSSIM", false, false, true, true, "  ", IVQM_M_IVMSSSIM},
};
// metric name suffixes (apps/QMIV/src/xMetricQMIV.cpp:61-82; mask mode adds "-M")
const char* kSuffixCmp = " Y:Cb:Cr  ";
const char* kSuffixPicPerPic = "          ";
const char* kSuffixPicPerCmp = "-YCbCr    ";
void set_suffixes(bool mask, bool rgb) {
  kSuffixCmp = mask ? (rgb ? "-M R:G:B  " : "-M Y:Cb:Cr") : (rgb ? " R:G:B    " : " Y:Cb:Cr  ");
  kSuffixPicPerPic = mask ? (rgb ? "-M-RGB    " : "-M        ") : (rgb ? "-RGB      " : "          ");
  kSuffixPicPerCmp = mask ? (rgb ? "-M-RGB    " : "-M-YCbCr  ") : (rgb ? "-RGB      " : "-YCbCr    ");
}

// Colour spaces of the app (apps/QMIV/src/xUtilsQMIV.h:12-27, xUtilsQMIV.cpp:13-25).
enum ClrSpc { kRGB = 2, kBGR = 3, kGBR = 4, kYCbCr = 16, kBT601 = 17, kSMPTE170M = 18, kBT709 = 19, kSMPTE240M = 20,
              kBT2020 = 21, kInvalidClrSpc = -1 };
int clrspc_from(const std::string& s) {
  static const std::pair<const char*, int> names[] = {
      {"RGB", kRGB}, {"BGR", kBGR}, {"GBR", kGBR}, {"YCbCr", kYCbCr}, {"YCbCr_BT601", kBT601},
      {"YCbCr_SMPTE170M", kSMPTE170M}, {"YCbCr_BT709", kBT709}, {"YCbCr_SMPTE240M", kSMPTE240M},
      {"YCbCr_BT2020", kBT2020}};
  for (const auto& n : names)
    if (s == n.first) return n.second;
  return kInvalidClrSpc;
}
bool is_rgb(int c) { return c == kRGB || c == kBGR || c == kGBR; }
bool is_ycbcr(int c) { return c >= kYCbCr; }
bool is_defined_ycbcr(int c) { return c > kYCbCr; }
// eClrSpcLC of a defined YCbCr (xUtilsQMIV.cpp:42-53): 0 BT.601 / SMPTE170M, 1 BT.709, 2 SMPTE240M, 3 BT.2020
int clrspc_lc(int c) { return c == kBT601 || c == kSMPTE170M ? 0 : c == kBT709 ? 1 : c == kSMPTE240M ? 2 : 3; }

struct Options {
  std::string in0, in1, result_file;
  std::string inm;               // mask file (-im), apps/QMIV/src/xAppQMIV.cpp:69-71
  std::string ff = "RAW";        // file format (-ff RAW | PNG | BMP), xAppQMIV.cpp:222-224
  int bdm = -1, cfm = -1;        // mask bit depth / chroma format (default: the pictures')
  int width = -1, height = -1, bit_depth = 8, chroma = 420;
  int start0 = 0, start1 = 0, num_frames = -1;
  bool calc[kNumMetrics] = {};
  bool erp = false;
  int lon = 360, lat = 180;
  int search_range = 2;
  int cws[4] = {4, 1, 1, 0}, cwa[4] = {4, 1, 1, 0};
  float unc[4] = {0.01f, 0.01f, 0.01f, 0.0f};
  int ss_stride = 4, ss_window = -1;  // window default: 11 for regular modes, else 8 (xAppQMIV.cpp:330)
  std::string ss_mode = "BlockAveraged", ss_border = "None";
  int ss_mode_id = 4, ss_border_id = 0;
  std::string csi_s, csm_s;      // -csi / -csm (default YCbCr, then ColorSpaceInput)
  int csi = kYCbCr, csm = kYCbCr;
  std::string ipa = "STOP";
  int verbose = 1;
};

std::string upper(std::string s) {
  for (char& c : s) c = (char)toupper((unsigned char)c);
  return s;
}
std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\""), b = s.find_last_not_of(" \t\"");
  return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}
bool parse_int(const std::string& s, int& v) {
  char* e = nullptr;
  const long x = strtol(s.c_str(), &e, 10);
  if (s.empty() || *e) return false;
  v = (int)x;
  return true;
}
template <typename T, typename F>
bool parse_vec4(const std::string& s, T out[4], F conv) {
  std::vector<std::string> parts;
  size_t p = 0;
  const std::string t = trim(s);
  while (true) {
    const size_t q = t.find(':', p);
    parts.push_back(t.substr(p, q == std::string::npos ? std::string::npos : q - p));
    if (q == std::string::npos) break;
    p = q + 1;
  }
  if (parts.size() != 4) return false;
  for (int k = 0; k < 4; k++)
    if (!conv(parts[k], out[k])) return false;
  return true;
}

void usage() {
  printf(
      "ivqm_qmiv: QMIV metrics of two raw YUV sequences on the GPU\n"
      " -i0 InputFile0  -i1 InputFile1  -ps WxH | -pw W -ph H  -bd BitDepth (8..16)  -cf 420|422|444\n"
      " -pf pix_fmt (yuv420p, yuv420p10le, yuv444p16le, ...)  -s0 / -s1 StartFrame  -nf NumberOfFrames\n"
      " -r ResultFile  -ml \"PSNR, WSPSNR, IVPSNR, SSIM, MSSSIM, IVSSIM, IVMSSSIM, MSE\" | All\n"
      " -erp  -lor LonRangeDeg  -lar LatRangeDeg  -sr SearchRange (1..4)  -cws / -cwa \"4:1:1:0\"\n"
      " -unc \"0.01:0.01:0.01:0\"  -ipa SKIP|WARN|STOP|CNCL  -v VerboseLevel\n"
      " -ssm RegularGaussianFlt|RegularGaussianInt|RegularAveraged|BlockGaussianInt|BlockAveraged\n"
      " -ssb None|Nearest|Reflect|Mirror|Constant|Zero (regular modes)  -sss Stride  -ssw Window (block modes)\n"
      " -im MaskFile  -bdm BitDepthM (8..14)  -cfm 400|420|422|444 (mask mode: MSE, PSNR, WSPSNR, IVPSNR)\n"
      " -csi / -csm RGB|BGR|GBR|YCbCr|YCbCr_BT601|YCbCr_SMPTE170M|YCbCr_BT709|YCbCr_SMPTE240M|YCbCr_BT2020\n"
      " -ff RAW|PNG|BMP (PNG / BMP: a file name pattern with an index field, e.g. frame_{:04d}.png; 8-bit RGB)\n");
}

// Returns false (after printing the problems) when the options are not usable.
bool parse_options(int argc, char** argv, Options& o) {
  std::string errs;
  bool metric_list = false;
  auto need = [&](int& i) -> std::string {
    if (i + 1 >= argc) {
      errs += std::string("!  missing value for ") + argv[i] + "\n";
      return std::string();
    }
    return argv[++i];
  };
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    auto int_opt = [&](int& v) {
      const std::string s = need(i);
      if (!parse_int(trim(s), v)) errs += "!  invalid value for " + a + ": " + s + "\n";
    };
    if (a == "-i0") o.in0 = need(i);
    else if (a == "-i1") o.in1 = need(i);
    else if (a == "-r") o.result_file = need(i);
    else if (a == "-ps") {
      const std::string s = need(i);
      const size_t x = s.find('x');
      if (x == std::string::npos || !parse_int(s.substr(0, x), o.width) || !parse_int(s.substr(x + 1), o.height))
        errs += "!  invalid PictureSize " + s + "\n";
    } else if (a == "-pw") int_opt(o.width);
    else if (a == "-ph") int_opt(o.height);
    else if (a == "-bd") int_opt(o.bit_depth);
    else if (a == "-cf") int_opt(o.chroma);
    else if (a == "-pf") {
      // FFmpeg pix_fmt: yuv4CCp[BB][le]
      const std::string s = need(i);
      int cf = 0, bd = 8;
      if (s.size() >= 7 && s.compare(0, 3, "yuv") == 0 && s[6] == 'p' && parse_int(s.substr(3, 3), cf)) {
        const std::string rest = s.substr(7);
        if (!rest.empty() && !(rest.size() > 2 && rest.compare(rest.size() - 2, 2, "le") == 0 &&
                               parse_int(rest.substr(0, rest.size() - 2), bd)))
          cf = -1;
      } else {
        cf = -1;
      }
      if (cf < 0) errs += "!  invalid or unsupported PictureFormat " + s + "\n";
      else { o.chroma = cf; o.bit_depth = bd; }
    } else if (a == "-s0") int_opt(o.start0);
    else if (a == "-s1") int_opt(o.start1);
    else if (a == "-nf") int_opt(o.num_frames);
    else if (a == "-ml") {
      metric_list = true;
      const std::string s = need(i);
      size_t p = 0;
      while (p <= s.size()) {
        const size_t q = s.find(',', p);
        const std::string tok = upper(trim(s.substr(p, q == std::string::npos ? std::string::npos : q - p)));
        if (tok == "ALL") {
          for (bool& c : o.calc) c = true;
        } else if (!tok.empty()) {
          int m = 0;
          while (m < kNumMetrics && tok != kInfo[m].name) m++;
          if (m == kNumMetrics) errs += "!  MetricList contains not valid entry. Token \"" + tok + "\" in not a metric.\n";
          else o.calc[m] = true;
        }
        if (q == std::string::npos) break;
        p = q + 1;
      }
    } else if (a == "-erp") o.erp = true;
    else if (a == "-lor") int_opt(o.lon);
    else if (a == "-lar") int_opt(o.lat);
    else if (a == "-sr") int_opt(o.search_range);
    else if (a == "-cws" || a == "-cwa") {
      const std::string s = need(i);
      if (!parse_vec4(s, a == "-cws" ? o.cws : o.cwa, [](const std::string& t, int& v) { return parse_int(trim(t), v); }))
        errs += "!  invalid component weights " + s + "\n";
    } else if (a == "-unc") {
      const std::string s = need(i);
      if (!parse_vec4(s, o.unc, [](const std::string& t, float& v) {
            char* e = nullptr;
            v = strtof(trim(t).c_str(), &e);
            return !trim(t).empty() && !*e;
          }))
        errs += "!  invalid UnnoticeableCoef " + s + "\n";
    } else if (a == "-ssm") o.ss_mode = need(i);
    else if (a == "-ssb") o.ss_border = need(i);
    else if (a == "-sss") int_opt(o.ss_stride);
    else if (a == "-ssw") int_opt(o.ss_window);
    else if (a == "-ipa") o.ipa = upper(need(i));
    else if (a == "-v") int_opt(o.verbose);
    else if (a == "-nth") (void)need(i);  // the GPU path has no worker-thread count
    else if (a == "-im") o.inm = need(i);
    else if (a == "-bdm") int_opt(o.bdm);
    else if (a == "-cfm") int_opt(o.cfm);
    else if (a == "-csi") o.csi_s = need(i);
    else if (a == "-csm") o.csm_s = need(i);
    else if (a == "-ff") o.ff = upper(trim(need(i)));
    else if (a == "-c" || a == "-nma") {
      const std::string v = need(i);
      if (!(a == "-ff" && upper(v) == "RAW") && !(a == "-nma")) errs += "!  unsupported option " + a + " " + v + "\n";
    } else errs += "!  unknown option " + a + "\n";
  }
  if (!metric_list)
    for (int m = 0; m < kNumMetrics; m++) o.calc[m] = kInfo[m].is_default;
  if (o.in0.empty()) errs += "!  InputFile0 is empty\n";
  if (o.in1.empty()) errs += "!  InputFile1 is empty\n";
  if (o.width <= 0 || o.height <= 0) errs += "!  Invalid or missing PictureSize\n";
  if (o.bit_depth < 8 || o.bit_depth > 16) errs += "!  Invalid or unsuported BitDepth value\n";
  if (o.chroma != 420 && o.chroma != 422 && o.chroma != 444)
    errs += "!  Invalid or unsuported ChromaFormat value (420, 422 or 444)\n";
  if (o.chroma == 420 && ((o.width | o.height) & 1)) errs += "!  4:2:0 needs even picture sizes\n";
  if (o.chroma == 422 && (o.width & 1)) errs += "!  4:2:2 needs an even picture width\n";
  if (o.search_range < 1 || o.search_range > 4) errs += "!  SearchRange must be 1..4\n";
  // xSSIM::xStrToMode (xSSIM.cpp:17-26, case-insensitive), xStr2MrgExt (xMiscUtilsCORE.cpp:152-164)
  const char* kModes[] = {"REGULARGAUSSIANFLT", "REGULARGAUSSIANINT", "REGULARAVERAGED", "BLOCKGAUSSIANINT",
                          "BLOCKAVERAGED"};
  const char* kBorders[] = {"None", "Nearest", "Reflect", "Mirror", "Constant", "Zero"};
  o.ss_mode_id = -1;
  for (int k = 0; k < 5; k++)
    if (upper(trim(o.ss_mode)) == kModes[k]) o.ss_mode_id = k;
  o.ss_border_id = -1;
  for (int k = 0; k < 6; k++)
    if (trim(o.ss_border) == kBorders[k]) o.ss_border_id = k;
  if (o.ss_mode_id < 0) errs += "!  StructSimMode value is not valid\n";
  if (o.ss_border_id < 0) errs += "!  StructSimBrdExt value is not valid\n";
  const bool regular = o.ss_mode_id >= 0 && o.ss_mode_id <= 2;
  if (o.ss_window < 0) o.ss_window = regular ? 11 : 8;
  if (o.ss_stride < 1 || o.ss_stride > o.ss_window) errs += "! StructSimStride must be in range 1-StructSimWindow\n";
  if (regular && o.ss_window != 11) errs += "! In regular struct sim mode only StructSimWindow==11 is allowed\n";
  if (o.ss_mode_id == 3 && o.ss_window != 8 && o.ss_window != 16)
    errs += "!  BlockGaussianInt supports StructSimWindow 8 and 16\n";
  if (o.ss_mode_id == 4 && (o.ss_window < 1 || o.ss_window > 64)) errs += "!  StructSimWindow must be 1..64\n";
  if (o.ipa != "SKIP" && o.ipa != "WARN" && o.ipa != "STOP" && o.ipa != "CNCL") errs += "!  invalid InvalidPelActn\n";
  // file format: PNG / BMP lists hold 8-bit RGB pictures (xAppQMIV.cpp:222-224, :297, :361-362)
  const bool images = o.ff == "PNG" || o.ff == "BMP";
  if (o.ff != "RAW" && !images) errs += "!  FileFormat is invalid\n";
  // colour spaces: xAppQMIV.cpp:297-311
  o.csi = o.csi_s.empty() ? (images ? kRGB : kYCbCr) : clrspc_from(trim(o.csi_s));
  o.csm = o.csm_s.empty() ? o.csi : clrspc_from(trim(o.csm_s));
  if (o.csi == kInvalidClrSpc) errs += "!  ColorSpaceInput value is not valid\n";
  if (o.csm == kInvalidClrSpc) errs += "!  ColorSpaceMetric value is not valid\n";
  if (o.csi != kInvalidClrSpc && o.csm != kInvalidClrSpc && o.csi != o.csm) {
    if (is_ycbcr(o.csi) && is_ycbcr(o.csm)) errs += "!  YCbCr to YCbCr conversion is not supported.\n";
    if (o.csi == kYCbCr || o.csm == kYCbCr) errs += "!  Generic YCbCr cannot be used for colorspace conversion.\n";
  }
  if (images && o.csi != kRGB)
    errs += "! Input FileFormat=" + o.ff + " contains data in RGB color space whitch conflicts with defined ColorSpaceInput=" +
            o.csi_s + "\n";
  if (images && o.bit_depth != 8)
    errs += "! Input FileFormat=" + o.ff + " contains 8-bit per pixel data whitch conflicts with defined BitDepth=" +
            std::to_string(o.bit_depth) + "\n";
  if (images) o.chroma = 444;  // R, G, B planes at full resolution
  if (images && !o.inm.empty() && o.bdm != 8) errs += "!  PNG / BMP masks are 8-bit (BitDepthM must be 8)\n";
  // mask mode: xAppQMIV.cpp:284-288, :360
  if (o.bdm < 0) o.bdm = o.bit_depth;
  if (o.cfm < 0) o.cfm = o.chroma;
  if (!o.inm.empty()) {
    if (o.bdm < 8 || o.bdm > 14) errs += "!  Invalid or unsuported BitDepthM value\n";
    if (o.cfm != 400 && o.cfm != 420 && o.cfm != 422 && o.cfm != 444) errs += "!  Invalid or unsuported ChromaFormatM value\n";
    if (o.calc[kSSIM] || o.calc[kMSSSIM] || o.calc[kIVSSIM] || o.calc[kIVMSSSIM])
      errs += "! Structural Similarity metrics cannot be combined with Mask mode\n";
  }
  if (o.start0 < 0 || o.start1 < 0) errs += "!  invalid StartFrame\n";
  if (!errs.empty()) {
    fputs(errs.c_str(), stderr);
    return false;
  }
  return true;
}

// ---- per-metric statistics: apps/QMIV/src/xMetricQMIV.cpp:44-160
struct MetricStat {
  std::vector<double> cmp[3], pic;
};

std::string fmt_per_cmp(int m, const double v[3]) {
  char b[256];
  const char* f = kInfo[m].normalized ? "%8.6f " : "%8.4f ";
  int n = snprintf(b, sizeof b, "%8s%s ", kInfo[m].name, kSuffixCmp);
  for (int c = 0; c < 3; c++) n += snprintf(b + n, sizeof b - n, f, v[c]);
  return b;
}
std::string fmt_per_pic(int m, double v) {
  char b[256];
  int n = snprintf(b, sizeof b, "%8s%s ", kInfo[m].name, kInfo[m].per_pic ? kSuffixPicPerPic : kSuffixPicPerCmp);
  n += snprintf(b + n, sizeof b - n, kInfo[m].normalized ? "%8.6f " : "%8.4f ", v);
  snprintf(b + n, sizeof b - n, "                  ");
  return b;
}
std::string fmt_avg(int m, const MetricStat& st, int nframes, const char* header) {
  const bool hdr = header[0] != 0;
  const char* single = kInfo[m].normalized ? "%10.8f %s  " : "%10.6f %s  ";
  std::string out;
  char b[512];
  auto avg = [&](const std::vector<double>& v) {
    ivqm::host::Kbns k;
    for (double x : v) k.acc(x);
    return k.result() / (double)nframes;
  };
  if (kInfo[m].per_cmp) {
    int n = snprintf(b, sizeof b, hdr ? "%s%8s%s " : "%s%-8s%s ", header, kInfo[m].name, kSuffixCmp);
    for (int c = 0; c < 3; c++) n += snprintf(b + n, sizeof b - n, single, avg(st.cmp[c]), kInfo[m].unit);
    out += std::string(b) + "\n";
  }
  int n = snprintf(b, sizeof b, hdr ? "%s%8s%s " : "%s%-8s%s ", header, kInfo[m].name,
                   kInfo[m].per_pic ? kSuffixPicPerPic : kSuffixPicPerCmp);
  n += snprintf(b + n, sizeof b - n, single, avg(st.pic), kInfo[m].unit);
  snprintf(b + n, sizeof b - n, "                  ");
  return out + b;
}

#define HIPX(x)                                                                            \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) {                                                                \
      fprintf(stderr, "ERROR --> HIP %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
      return 2;                                                                            \
    }                                                                                      \
  } while (0)
#define IVQMX(x)                                                                          \
  do {                                                                                    \
    int s_ = (x);                                                                         \
    if (s_ != IVQM_OK) {                                                                  \
      fprintf(stderr, "ERROR --> %s (%s)\n", ivqm_strerror(s_), ivqm_last_error());       \
      return 2;                                                                           \
    }                                                                                     \
  } while (0)

constexpr int kGroup = 8;            // frame pairs per batch
constexpr int kSlots = 2 * kGroup;   // ring of pinned host / device slots

struct Ring {
  std::mutex mu;
  std::condition_variable cv;
  std::vector<int> state;  // per slot: frame index loaded into it, or -1 (free)
  std::vector<int> error;  // per slot: 0, or 1 = read error, 2 = invalid sample value
  bool stop = false;
};

}  // namespace

size_t chroma_of(int W, int H, int cs) { return (size_t)(W >> cs) * (H >> cs); }

int main(int argc, char** argv) {
  Options o;
  for (int i = 1; i < argc; i++)
    if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "--help")) {
      usage();
      return 0;
    }
  if (!parse_options(argc, argv, o)) {
    usage();
    return 1;
  }
  // chroma planes: 4:2:0 (W/2) x (H/2), 4:2:2 (W/2) x H (xSeq.cpp:231-301); cf is the ABI's format code
  const int W = o.width, H = o.height, bps = o.bit_depth > 8 ? 2 : 1;
  const int csx = o.chroma == 444 ? 0 : 1, csy = o.chroma == 420 ? 1 : 0, cf = o.chroma == 420 ? 1 : (o.chroma == 422 ? 2 : 0);
  const size_t luma = (size_t)W * H, chroma = (size_t)(W >> csx) * (H >> csy);
  const size_t frame_bytes = (luma + 2 * chroma) * bps;
  // mask frames: luma plane of InputFileM (BitDepthM, ChromaFormatM); only the luma plane is used
  const bool use_mask = !o.inm.empty();
  const int bpsm = o.bdm > 8 ? 2 : 1;
  const size_t chroma_m = o.cfm == 400 ? 0 : o.cfm == 420 ? 2 * chroma_of(W, H, 1) : o.cfm == 422 ? (size_t)W * H : 2 * luma;
  const size_t frame_bytes_m = (luma + chroma_m) * bpsm, mask_bytes = use_mask ? luma * bpsm : 0;
  const size_t slot_bytes = 2 * frame_bytes + mask_bytes;
  // device slots keep rows at a pitch of a multiple of 4 samples (the batch ABI's alignment for
  // full-resolution planes); the upload converts the pitch with 2-D copies
  const int Wc = W >> csx, Hc = H >> csy, Wp = (W + 3) & ~3, Wcp = (Wc + 3) & ~3;
  const size_t dframe = ((size_t)Wp * H + 2 * (size_t)Wcp * Hc) * bps;
  const size_t dslot = 2 * dframe + mask_bytes;
  set_suffixes(use_mask, is_rgb(o.csm));
  // colour processing of the app's preprocessFrames (xAppQMIV.cpp:308-311, :853-892): conversions run
  // on the device into 4:4:4 planes; BGR / GBR reorders (metrics in RGB) are plane swaps
  const bool cvt_y2r = is_defined_ycbcr(o.csi) && is_rgb(o.csm), cvt_r2y = is_rgb(o.csi) && is_defined_ycbcr(o.csm);
  const bool reorder = is_rgb(o.csi) && o.csi != kRGB && o.csm == kRGB;
  const bool colour = cvt_y2r || cvt_r2y || reorder;
  const int space = clrspc_lc(cvt_y2r ? o.csi : o.csm);

  const bool images = o.ff != "RAW";
  int first_idx[3] = {0, 0, 0};  // image lists: index of each sequence's first file (0 or 1)
  long long nfiles[3] = {0, 0, 0};
  if (images) {
    const std::string* in[3] = {&o.in0, &o.in1, &o.inm};
    for (int i = 0; i < (o.inm.empty() ? 2 : 3); i++) {
      nfiles[i] = ivqm::images::count_frames(*in[i], &first_idx[i]);
      if (nfiles[i] <= 0) {
        fprintf(stderr, "ERROR --> InputFile opening failure (%s) First Idx=(0 or 1) not found\n", in[i]->c_str());
        return 1;
      }
    }
  }
  FILE* f0 = images ? nullptr : fopen(o.in0.c_str(), "rb");
  FILE* f1 = images ? nullptr : fopen(o.in1.c_str(), "rb");
  if (!images && (!f0 || !f1)) {
    fprintf(stderr, "ERROR --> InputFile%d does not exist (%s)\n", f0 ? 1 : 0, (f0 ? o.in1 : o.in0).c_str());
    return 1;
  }
  auto frames_in = [&](FILE* f, int i) {
    if (images) return nfiles[i];
    fseeko(f, 0, SEEK_END);
    const long long n = (long long)(ftello(f) / (off_t)frame_bytes);
    return n;
  };
  const long long n0 = frames_in(f0, 0) - o.start0, n1 = frames_in(f1, 1) - o.start1;
  long long nf = std::min(n0, n1);
  if (o.num_frames >= 0) nf = std::min<long long>(nf, o.num_frames);
  if (nf <= 0) {
    fprintf(stderr, "ERROR --> no frames to process (%lld and %lld frames after the start frames)\n", n0, n1);
    return 1;
  }
  const int N = (int)nf;
  if (use_mask && images && N > nfiles[2]) {
    fprintf(stderr, "ERROR --> FramesToProcess > NumOfFramesM\n");
    return 1;
  }
  if (use_mask && !images) {  // the mask sequence is read from its first frame (xAppQMIV.cpp:622-627)
    FILE* fm = fopen(o.inm.c_str(), "rb");
    if (!fm) {
      fprintf(stderr, "ERROR --> InputFileM does not exist (%s)\n", o.inm.c_str());
      return 1;
    }
    fseeko(fm, 0, SEEK_END);
    const long long nm = (long long)(ftello(fm) / (off_t)frame_bytes_m);
    fclose(fm);
    if (N > nm) {
      fprintf(stderr, "ERROR --> FramesToProcess > NumOfFramesM\n");
      return 1;
    }
  }

  uint32_t metrics = 0;
  for (int m = 0; m < kNumMetrics; m++)
    if (o.calc[m]) metrics |= kInfo[m].bit;

  ivqm_config cfg;
  IVQMX(ivqm_config_init(&cfg));
  cfg.width = W;
  cfg.height = H;
  cfg.bit_depth = o.bit_depth;
  cfg.search_range = o.search_range;
  for (int k = 0; k < 4; k++) {
    cfg.cmp_weights_search[k] = o.cws[k];
    cfg.cmp_weights_average[k] = o.cwa[k];
    cfg.unnoticeable_coef[k] = o.unc[k];
  }
  cfg.erp = o.erp;
  cfg.lon_range_deg = o.lon;
  cfg.lat_range_deg = o.lat;
  cfg.fake_vals_for_exact = 1;  // apps/QMIV/src/xAppQMIV.cpp:974-981, :1002-1009
  cfg.ssim_mode = o.ss_mode_id;
  cfg.ssim_border = o.ss_border_id;
  cfg.ssim_window = o.ss_window;
  cfg.ssim_stride = o.ss_stride;
  ivqm_batch* batch = nullptr;
  IVQMX(ivqm_batch_create(&cfg, &batch));

  hipStream_t cstream, xstream;
  HIPX(hipStreamCreateWithFlags(&cstream, hipStreamNonBlocking));
  HIPX(hipStreamCreateWithFlags(&xstream, hipStreamNonBlocking));
  IVQMX(ivqm_batch_set_stream(batch, (void*)cstream));
  IVQMX(ivqm_batch_set_tuning(batch, 0, 0, 1));
  std::vector<uint8_t*> host(kSlots), dev(kSlots);
  std::vector<uint16_t*> conv(kSlots, nullptr);  // converted 4:4:4 pictures (colour processing)
  std::vector<hipEvent_t> copied(kSlots);
  for (int s = 0; s < kSlots; s++) {
    HIPX(hipHostMalloc((void**)&host[s], slot_bytes, hipHostMallocDefault));
    HIPX(hipMalloc((void**)&dev[s], dslot));
    if (colour) HIPX(hipMalloc((void**)&conv[s], 2 * 3 * (size_t)Wp * H * sizeof(uint16_t)));
    HIPX(hipEventCreateWithFlags(&copied[s], hipEventDisableTiming));
  }

  // ---- reader: frame f goes to slot f % kSlots once that slot is free
  Ring ring;
  ring.state.assign(kSlots, -1);
  ring.error.assign(kSlots, 0);
  const uint32_t maxv = (1u << o.bit_depth) - 1;
  const bool check = o.ipa != "SKIP";
  // kReaders threads read frames f = t, t + kReaders, ... (file reads and the
  // sample check run in parallel; the slot order is kept by frame index)
  const int kReaders = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency() / 2));
  auto read_frames = [&](int t) {
    FILE* fs[3] = {images ? nullptr : fopen(o.in0.c_str(), "rb"), images ? nullptr : fopen(o.in1.c_str(), "rb"),
                   use_mask && !images ? fopen(o.inm.c_str(), "rb") : nullptr};
    std::vector<uint8_t> mrgb(images && use_mask ? 3 * luma : 0);
    const int st[2] = {o.start0, o.start1};
    for (int f = t; f < N; f += kReaders) {
      const int s = f % kSlots;
      {
        std::unique_lock<std::mutex> lk(ring.mu);
        ring.cv.wait(lk, [&] { return ring.stop || ring.state[s] == -1; });
        if (ring.stop) break;
      }
      int err = 0;
      uint8_t* dst = host[s];
      for (int i = 0; i < 2 && !err && images; i++) {  // one image file per frame (xSeqLST.cpp:189-241, :337-415)
        const std::string name = ivqm::images::format_name(i ? o.in1 : o.in0, first_idx[i] + st[i] + f);
        const std::string e = o.ff == "PNG" ? ivqm::images::read_png(name, W, H, dst + i * frame_bytes)
                                            : ivqm::images::read_bmp(name, W, H, dst + i * frame_bytes);
        if (!e.empty()) {
          fprintf(stderr, "Frame %08d ERROR --> %s\n", f, e.c_str());
          err = 1;
        }
      }
      if (!err && images && use_mask) {  // the mask is the first (R) plane of its image
        const std::string name = ivqm::images::format_name(o.inm, first_idx[2] + f);
        const std::string e = o.ff == "PNG" ? ivqm::images::read_png(name, W, H, mrgb.data())
                                            : ivqm::images::read_bmp(name, W, H, mrgb.data());
        if (!e.empty()) {
          fprintf(stderr, "Frame %08d ERROR --> %s\n", f, e.c_str());
          err = 1;
        } else {
          memcpy(dst + 2 * frame_bytes, mrgb.data(), luma);
        }
      }
      for (int i = 0; i < 2 && !err && !images; i++) {
        if (!fs[i] || fseeko(fs[i], (off_t)(st[i] + f) * (off_t)frame_bytes, SEEK_SET) != 0 ||
            fread(dst + i * frame_bytes, 1, frame_bytes, fs[i]) != frame_bytes)
          err = 1;
      }
      if (!err && use_mask && !images &&
          (!fs[2] || fseeko(fs[2], (off_t)f * (off_t)frame_bytes_m, SEEK_SET) != 0 ||
           fread(dst + 2 * frame_bytes, 1, mask_bytes, fs[2]) != mask_bytes))
        err = 1;
      // samples above the bit depth's maximum are invalid (InvalidPelActn,
      // apps/QMIV/src/xAppQMIV.cpp:770-775): a vectorised max first, the
      // per-sample action only when one is found
      if (!err && check && bps == 2) {
        uint16_t* p = (uint16_t*)dst;
        const size_t ns = 2 * frame_bytes / 2;
        uint16_t mx = 0;
        for (size_t k = 0; k < ns; k++) mx = std::max(mx, p[k]);
        if (mx > maxv) {
          if (o.ipa == "CNCL")
            for (size_t k = 0; k < ns; k++) p[k] = (uint16_t)std::min<uint32_t>(p[k], maxv);
          else if (o.ipa == "STOP")
            err = 2;
        }
      }
      std::lock_guard<std::mutex> lk(ring.mu);
      ring.state[s] = f;
      ring.error[s] = err;
      ring.cv.notify_all();
    }
    for (FILE* f : fs)
      if (f) fclose(f);
  };
  std::vector<std::thread> readers;
  for (int t = 0; t < kReaders; t++) readers.emplace_back(read_frames, t);
  auto fail = [&](int rc) {
    {
      std::lock_guard<std::mutex> lk(ring.mu);
      ring.stop = true;
      ring.cv.notify_all();
    }
    for (std::thread& r : readers) r.join();
    return rc;
  };

  std::vector<ivqm_frame_result> res(N);
  auto picture = [&](uint8_t* base) {  // a picture of a device slot (pitched rows)
    ivqm_picture p;
    p.plane[0] = base;
    p.plane[1] = base + (size_t)Wp * H * bps;
    p.plane[2] = base + ((size_t)Wp * H + (size_t)Wcp * Hc) * bps;
    p.stride[0] = Wp;
    p.stride[1] = p.stride[2] = Wcp;
    p.chroma_420 = cf;
    p.bytes_per_sample = bps;
    return p;
  };
  // host frame (tight rows) -> device slot (pitched rows), on the copy stream
  auto upload = [&](uint8_t* dst, const uint8_t* src) -> bool {
    const int pw[3] = {W, Wc, Wc}, ph[3] = {H, Hc, Hc}, pp[3] = {Wp, Wcp, Wcp};
    size_t so = 0, doff = 0;
    for (int k = 0; k < 3; k++) {
      if (hipMemcpy2DAsync(dst + doff, (size_t)pp[k] * bps, src + so, (size_t)pw[k] * bps, (size_t)pw[k] * bps, ph[k],
                           hipMemcpyHostToDevice, xstream) != hipSuccess)
        return false;
      so += (size_t)pw[k] * ph[k] * bps;
      doff += (size_t)pp[k] * ph[k] * bps;
    }
    return true;
  };
  const int ngroups = (N + kGroup - 1) / kGroup;
  const auto t_begin = std::chrono::steady_clock::now();
  auto release = [&](int g) {  // the group's slots may be refilled
    std::lock_guard<std::mutex> lk(ring.mu);
    for (int f = g * kGroup; f < std::min(N, (g + 1) * kGroup); f++) ring.state[f % kSlots] = -1;
    ring.cv.notify_all();
  };
  for (int g = 0; g <= ngroups; g++) {
    if (g < ngroups) {
      const int fb = g * kGroup, fe = std::min(N, fb + kGroup);
      std::vector<ivqm_pair> pairs;
      std::vector<ivqm_mask> masks;
      for (int f = fb; f < fe; f++) {
        const int s = f % kSlots;
        int err;
        {
          std::unique_lock<std::mutex> lk(ring.mu);
          ring.cv.wait(lk, [&] { return ring.state[s] == f; });
          err = ring.error[s];
        }
        if (err == 1) {
          fprintf(stderr, "Frame %08d ERROR --> InputFile read error\n", f);
          return fail(1);
        }
        if (err == 2) {
          fprintf(stderr, "Frame %08d ERROR --> sample value above %u (InvalidPelActn=STOP)\n", f, maxv);
          return fail(1);
        }
        if (!upload(dev[s], host[s]) || !upload(dev[s] + dframe, host[s] + frame_bytes) ||
            (use_mask && hipMemcpyAsync(dev[s] + 2 * dframe, host[s] + 2 * frame_bytes, mask_bytes,
                                        hipMemcpyHostToDevice, xstream) != hipSuccess))
          return fail(2);
        ivqm_pair pr;
        pr.p0 = picture(dev[s]);
        pr.p1 = picture(dev[s] + dframe);
        if (colour) {
          ivqm_picture* pp[2] = {&pr.p0, &pr.p1};
          for (int i = 0; i < 2; i++) {
            uint16_t* d = conv[s] + (size_t)i * 3 * Wp * H;
            const int rc = reorder ? ivqm_upsample420(pp[i], W, H, d, Wp, (void*)xstream)
                                   : ivqm_convert_colour(pp[i], W, H, o.bit_depth, cvt_y2r ? 0 : 1, space, d, Wp,
                                                         (void*)xstream);
            if (rc != IVQM_OK) {
              fprintf(stderr, "ERROR --> %s (%s)\n", ivqm_strerror(rc), ivqm_last_error());
              return fail(2);
            }
            uint16_t* pl[3] = {d, d + (size_t)Wp * H, d + 2 * (size_t)Wp * H};
            if (reorder && o.csi == kBGR) std::swap(pl[0], pl[2]);  // BGR -> RGB
            if (reorder && o.csi == kGBR) {                         // GBR -> BGR -> RGB
              std::swap(pl[0], pl[1]);
              std::swap(pl[0], pl[2]);
            }
            ivqm_picture q;
            for (int k = 0; k < 3; k++) {
              q.plane[k] = pl[k];
              q.stride[k] = Wp;
            }
            q.chroma_420 = 0;
            q.bytes_per_sample = 2;
            *pp[i] = q;
          }
        }
        if (hipEventRecord(copied[s], xstream) != hipSuccess) return fail(2);
        pr.width = W;
        pr.height = H;
        pr.bit_depth = o.bit_depth;
        pairs.push_back(pr);
        masks.push_back(ivqm_mask{use_mask ? dev[s] + 2 * dframe : nullptr, W, bpsm, o.bdm, 0});
      }
      if (hipStreamWaitEvent(cstream, copied[(fe - 1) % kSlots], 0) != hipSuccess) return fail(2);
      const int rc = use_mask ? ivqm_batch_enqueue_masked(batch, pairs.data(), masks.data(), (int)pairs.size(), metrics,
                                                          g % 2)
                              : ivqm_batch_enqueue(batch, pairs.data(), (int)pairs.size(), metrics, g % 2);
      if (rc != IVQM_OK) {
        fprintf(stderr, "ERROR --> %s (%s)\n", ivqm_strerror(rc), ivqm_last_error());
        return fail(2);
      }
    }
    if (g > 0) {
      const int pg = g - 1;
      const int rc = ivqm_batch_collect(batch, pg % 2, res.data() + pg * kGroup);
      if (rc != IVQM_OK) {
        fprintf(stderr, "ERROR --> %s (%s)\n", ivqm_strerror(rc), ivqm_last_error());
        return fail(2);
      }
      release(pg);
    }
  }
  for (std::thread& r : readers) r.join();
  const double loop_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_begin).count();

  // ---- per-frame lines (apps/QMIV/src/xAppQMIV.cpp:950-1110) and statistics
  MetricStat st[kNumMetrics];
  const int32_t* cw = cfg.cmp_weights_average;
  for (int f = 0; f < N; f++) {
    const ivqm_frame_result& r = res[f];
    const bool print = o.verbose >= 2;
    std::string exact;
    if (r.exact[0]) exact += " ExactY";
    if (r.exact[1]) exact += " ExactU";
    if (r.exact[2]) exact += " ExactV";
    auto per_cmp = [&](int m, const double* v, double pic, bool with_exact, bool always) {
      for (int c = 0; c < 3; c++) st[m].cmp[c].push_back(v[c]);
      st[m].pic.push_back(pic);
      if (print || always) {
        printf("Frame %08d %s%s\n", f, fmt_per_cmp(m, v).c_str(), with_exact ? exact.c_str() : "");
        printf("Frame %08d %s\n", f, fmt_per_pic(m, pic).c_str());
      }
    };
    auto per_pic = [&](int m, double v) {
      st[m].pic.push_back(v);
      if (print) printf("Frame %08d %s\n", f, fmt_per_pic(m, v).c_str());
    };
    if (o.calc[kMSE]) per_cmp(kMSE, r.mse, ivqm::host::pic_value(r.mse, cw), true, false);
    if (o.calc[kPSNR]) per_cmp(kPSNR, r.psnr, r.psnr_pic, true, false);
    if (o.calc[kWSPSNR]) per_cmp(kWSPSNR, r.wspsnr, r.wspsnr_pic, true, false);
    if (o.calc[kIVPSNR]) per_pic(kIVPSNR, r.ivpsnr);
    if (o.calc[kSSIM]) per_cmp(kSSIM, r.ssim, r.ssim_pic, false, false);
    if (o.calc[kMSSSIM]) per_cmp(kMSSSIM, r.msssim, r.msssim_pic, false, true);  // printed at every level
    if (o.calc[kIVSSIM]) per_pic(kIVSSIM, r.ivssim);
    if (o.calc[kIVMSSSIM]) per_pic(kIVMSSSIM, r.ivmsssim);
  }

  // ---- averages (apps/QMIV/src/xAppQMIV.cpp:1131-1160)
  for (int m = 0; m < kNumMetrics; m++)
    if (o.calc[m]) printf("%s\n", fmt_avg(m, st[m], N, "Average      ").c_str());
  if (o.verbose >= 3) {
    // wall time of the frame loop: file reads, host-to-device copies and every metric, overlapped
    printf("\nAvgTime         TOTAL %9.2f ms   (%d frame pairs, %.1f Mpixel-pairs/s)\n", loop_ms / N, N,
           (double)W * H * N / (loop_ms * 1e3));
  }
  if (!o.result_file.empty()) {
    FILE* rf = fopen(o.result_file.c_str(), "w");
    if (!rf) {
      fprintf(stderr, "ERROR --> cannot write ResultFile %s\n", o.result_file.c_str());
      return 1;
    }
    char tb[64];
    const time_t now = time(nullptr);
    strftime(tb, sizeof tb, "%Y-%m-%d  %H:%M:%S", localtime(&now));
    fprintf(rf, "FILE0  \"%s\"\nFILE1  \"%s\"\n", o.in0.c_str(), o.in1.c_str());
    if (use_mask) fprintf(rf, "FILEM  \"%s\"\n", o.inm.c_str());  // xAppQMIV.cpp:1140
    fprintf(rf, "TIME   %s\n", tb);
    for (int m = 0; m < kNumMetrics; m++)
      if (o.calc[m]) fprintf(rf, "%s\n", fmt_avg(m, st[m], N, "").c_str());
    fclose(rf);
  }
  for (int s = 0; s < kSlots; s++) {
    if (conv[s]) (void)hipFree(conv[s]);
    (void)hipHostFree(host[s]);
    (void)hipFree(dev[s]);
    (void)hipEventDestroy(copied[s]);
  }
  ivqm_batch_destroy(batch);
  (void)hipStreamDestroy(cstream);
  (void)hipStreamDestroy(xstream);
  if (f0) fclose(f0);  // image lists open their files per frame
  if (f1) fclose(f1);
  return 0;
}
