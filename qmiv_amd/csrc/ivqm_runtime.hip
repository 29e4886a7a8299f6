// Host runtime and C ABI (include/ivqm_hip.h) of the gfx950 metric hot path.
//
// One context = one picture size / bit depth / parameter set (the reference
// configures one set of processor objects per sequence,
// apps/QMIV/src/xAppQMIV.cpp:680-744). A frame pass enqueues, on the
// context's stream: stats -> GCD -> search (both directions, fused SCP) ->
// pyramid -> SSIM jobs -> FP64 totals -> one device-to-host copy of the small
// integer / FP64 results into a pinned ring slot. Collecting a slot runs the
// host FP64 epilogue (ivqm_epilogue.h). Nothing on the device path falls back
// to the CPU: a missing or failing HIP runtime is reported as an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ivqm_hip.h"
#include "ivqm_device.h"
#include "ivqm_epilogue.h"

using namespace ivqm;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return IVQM_EHIP - (int)e;
}
#define HIPCHK(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_fail(_e, #expr);  \
  } while (0)

constexpr int kRing = 8;
constexpr int kPitchAlign = 64;  // samples

int round_up(int v, int a) { return (v + a - 1) / a * a; }

int key_bits(int R) {
  const int n = (2 * R + 1) * (2 * R + 1);
  return n <= 16 ? 4 : n <= 32 ? 5 : n <= 64 ? 6 : 7;
}

int ms_num_scales(int w, int h) {  // libs/pmbb-ivqm/src/xSSIM.cpp:405-417
  for (int i = 0; i < 5; i++) {
    const int lim = 32 << i;
    if (w < lim || h < lim) return i;
  }
  return 5;
}
int num_blocks(int len, int wnd, int stride) {  // libs/pmbb-ivqm/src/xSSIM.cpp:381-404
  int cnt = 0;
  for (int i = 0; i < len - wnd + 1; i += stride) cnt++;
  return cnt;
}
const double kMsExp[5][5] = {{1.0000, 0.0000, 0.0000, 0.0000, 0.0000}, {0.1356, 0.8644, 0.0000, 0.0000, 0.0000},
                             {0.0711, 0.4530, 0.4760, 0.0000, 0.0000}, {0.0517, 0.3295, 0.3462, 0.2726, 0.0000},
                             {0.0448, 0.2856, 0.3001, 0.2363, 0.1333}};  // xStructSimConsts.h:42-49

// Host-visible part of one ring slot (pinned).
struct RingHost {
  int32_t gcd[4];
  int64_t sd[3];
  uint64_t ssd[3];
  double ssim_out[kMaxSsimJobs * 2 * 3 * 2];  // [group][job][3][2]
};

struct SsimPlan {
  SsimJobTable tab[2];
  int ngroups = 0;
  int max_nchunk[2] = {0, 0}, max_bands[2] = {0, 0};
  // job indices (group * kMaxSsimJobs + j) of each evaluation, -1 = absent
  int ssim = -1;
  int ms[5] = {-1, -1, -1, -1, -1};
  int iv_t2r = -1, iv_r2t = -1;
  int ivms_t2r[5] = {-1, -1, -1, -1, -1}, ivms_r2t[5] = {-1, -1, -1, -1, -1};
  int nux[kMaxSsimJobs * 2], nuy[kMaxSsimJobs * 2];
};

}  // namespace

struct ivqm_ctx {
  ivqm_config cfg;
  int W = 0, H = 0, bd = 8, maxv = 255, R = 2, pitch = 0;
  int32_t maxdiff[3] = {0, 0, 0};
  bool narrow = false;
  int tile_py = 8;
  int ns = 1;  // MS-SSIM scales
  hipStream_t stream = nullptr;
  bool own_stream = false;

  // bound pictures
  PicDev pic[2];
  bool bound[2] = {false, false};
  void* own[2] = {nullptr, nullptr};
  size_t own_bytes[2] = {0, 0};

  // device work buffers
  int64_t* d_row_sd = nullptr;
  uint64_t* d_row_ssd = nullptr;  // [3][H]
  uint32_t* d_rows = nullptr;     // [2][3][H]
  int32_t* d_gcd = nullptr;
  int64_t* d_sd = nullptr;
  uint64_t* d_ssd = nullptr;
  uint16_t* d_scp = nullptr;  // [2][3][H][pitch]
  uint8_t* d_best = nullptr;  // [2][H][W]
  bool best_enabled = false;
  uint16_t* d_pyr = nullptr;  // [4 pictures][levels 1..4][3][h_l][pitch_l]
  size_t pyr_off[4][5];
  int pyr_pitch[5];
  double* d_partial = nullptr;
  size_t partial_cap = 0;
  double* d_ssim_out = nullptr;  // [2 groups][kMaxSsimJobs][3][2]
  double* d_erp = nullptr;
  std::vector<double> erp_w;
  double erp_corr = 1.0;

  // ring
  RingHost* h_ring = nullptr;       // pinned, kRing entries
  uint64_t* h_row_ssd = nullptr;    // pinned [kRing][3][H]
  uint32_t* h_rows = nullptr;       // pinned [kRing][2][3][H]
  hipEvent_t ev[kRing];
  uint32_t ring_metrics[kRing];
  SsimPlan ring_plan[kRing];
  bool ring_busy[kRing];
  bool scp_valid = false;
};

namespace {

PicDev scp_pic(const ivqm_ctx* c, int idx) {
  PicDev p;
  const size_t plane = (size_t)c->pitch * c->H;
  for (int k = 0; k < 3; k++) {
    p.p[k] = c->d_scp + ((size_t)idx * 3 + k) * plane;
    p.stride[k] = c->pitch;
  }
  p.cs = 0;
  p.bps = 2;
  return p;
}
// Pyramid picture: 0 = P0, 1 = P1, 2 = SCP0, 3 = SCP1; level >= 1.
PicDev pyr_pic(const ivqm_ctx* c, int which, int level) {
  PicDev p;
  const int h = c->H >> level;
  const size_t plane = (size_t)c->pyr_pitch[level] * h;
  for (int k = 0; k < 3; k++) {
    p.p[k] = c->d_pyr + c->pyr_off[which][level] + (size_t)k * plane;
    p.stride[k] = c->pyr_pitch[level];
  }
  p.cs = 0;
  p.bps = 2;
  return p;
}
PicDev level_pic(const ivqm_ctx* c, int which, int level) {
  if (level == 0) return which < 2 ? c->pic[which] : scp_pic(c, which - 2);
  return pyr_pic(c, which, level);
}

bool narrow_ok(const ivqm_ctx* c, const int32_t* gcd_abs_bound) {
  const int32_t* w = c->cfg.cmp_weights_search;
  for (int k = 0; k < 3; k++)
    if (w[k] < 0) return false;
  int64_t gmax = 0;
  for (int k = 0; k < 3; k++) gmax = std::max<int64_t>(gmax, gcd_abs_bound[k]);
  const int64_t A = (int64_t)c->maxv + gmax;
  const int64_t S = 1 << key_bits(c->R);
  const int64_t wsum = (int64_t)w[0] + w[1] + w[2];
  const int64_t wmax = std::max<int64_t>(w[0], std::max<int64_t>(w[1], w[2]));
  if (3 * S * wsum * A * A >= (1ll << 31)) return false;
  if (2 * S * wmax * A >= (1ll << 23)) return false;
  return true;
}

int ensure_partial(ivqm_ctx* c, size_t n) {
  if (n <= c->partial_cap) return IVQM_OK;
  if (c->d_partial) (void)hipFree(c->d_partial);
  c->d_partial = nullptr;
  c->partial_cap = 0;
  HIPCHK(hipMalloc(&c->d_partial, n * sizeof(double)));
  c->partial_cap = n;
  return IVQM_OK;
}

int ensure_pyramid(ivqm_ctx* c) {
  if (c->d_pyr || c->ns <= 1) return IVQM_OK;
  size_t off = 0;
  for (int l = 1; l < c->ns; l++) c->pyr_pitch[l] = round_up(c->W >> l, kPitchAlign);
  for (int which = 0; which < 4; which++)
    for (int l = 1; l < c->ns; l++) {
      c->pyr_off[which][l] = off;
      off += (size_t)3 * c->pyr_pitch[l] * (c->H >> l);
    }
  HIPCHK(hipMalloc(&c->d_pyr, off * sizeof(uint16_t)));
  return IVQM_OK;
}

int validate_picture(const ivqm_ctx* c, const ivqm_picture* p) {
  if (!p) return fail(IVQM_EINVAL, "picture is NULL");
  if (p->bytes_per_sample != 1 && p->bytes_per_sample != 2) return fail(IVQM_EINVAL, "bytes_per_sample must be 1 or 2");
  if (p->bytes_per_sample == 1 && c->bd != 8) return fail(IVQM_EINVAL, "8-bit samples require bit_depth 8");
  if (p->chroma_420 && ((c->W & 1) || (c->H & 1))) return fail(IVQM_EINVAL, "4:2:0 requires even width and height");
  const int cs = p->chroma_420 ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    if (!p->plane[k]) return fail(IVQM_EINVAL, "plane pointer is NULL");
    const int wp = k ? (c->W >> cs) : c->W;
    if (p->stride[k] < wp) return fail(IVQM_EINVAL, "stride smaller than plane width");
  }
  return IVQM_OK;
}

// Copies a picture (host or device) into the slot's own pitched buffer.
int copy_into_slot(ivqm_ctx* c, int slot, const ivqm_picture* p, hipMemcpyKind kind) {
  const int cs = p->chroma_420 ? 1 : 0, bps = p->bytes_per_sample;
  const int pitch_c = round_up(c->W >> cs, kPitchAlign);
  const size_t bytes = ((size_t)c->pitch * c->H + 2 * (size_t)pitch_c * (c->H >> cs)) * bps;
  if (c->own_bytes[slot] < bytes) {
    if (c->own[slot]) (void)hipFree(c->own[slot]);
    c->own[slot] = nullptr;
    c->own_bytes[slot] = 0;
    HIPCHK(hipMalloc(&c->own[slot], bytes));
    c->own_bytes[slot] = bytes;
  }
  PicDev d;
  char* base = (char*)c->own[slot];
  size_t off = 0;
  for (int k = 0; k < 3; k++) {
    const int wp = k ? (c->W >> cs) : c->W, hp = k ? (c->H >> cs) : c->H;
    const int pp = k ? pitch_c : c->pitch;
    d.p[k] = base + off;
    d.stride[k] = pp;
    HIPCHK(hipMemcpy2DAsync(base + off, (size_t)pp * bps, p->plane[k], (size_t)p->stride[k] * bps, (size_t)wp * bps,
                            hp, kind, c->stream));
    off += (size_t)pp * hp * bps;
  }
  d.cs = cs;
  d.bps = bps;
  c->pic[slot] = d;
  c->bound[slot] = true;
  c->scp_valid = false;
  return IVQM_OK;
}

bool aligned_for_kernels(const ivqm_picture* p) {
  const int bps = p->bytes_per_sample;
  for (int k = 0; k < 3; k++) {
    const bool full = (k == 0) || !p->chroma_420;
    if (full) {
      if (((uintptr_t)p->plane[k]) % (size_t)(4 * bps)) return false;
      if (p->stride[k] % 4) return false;
    }
  }
  return true;
}

int check_pair(const ivqm_ctx* c) {
  if (!c->bound[0] || !c->bound[1]) return fail(IVQM_ESTATE, "both picture slots must be bound");
  if (c->pic[0].cs != c->pic[1].cs || c->pic[0].bps != c->pic[1].bps)
    return fail(IVQM_EINVAL, "both pictures must share chroma format and sample size");
  return IVQM_OK;
}

void add_job(SsimPlan& pl, int* slot_out, const PicDev& a, const PicDev& b, int w, int h, int calc_l, int erp,
             const ivqm_config& cfg, size_t& part_off) {
  int g = pl.ngroups ? pl.ngroups - 1 : 0;
  if (pl.ngroups == 0) { pl.ngroups = 1; pl.tab[0].njobs = 0; }
  if (pl.tab[g].njobs == kMaxSsimJobs) {
    g = pl.ngroups++;
    pl.tab[g].njobs = 0;
  }
  SsimJob& J = pl.tab[g].job[pl.tab[g].njobs];
  J.a = a;
  J.b = b;
  J.w = w;
  J.h = h;
  J.nux = num_blocks(w, cfg.ssim_window, cfg.ssim_stride);
  J.nuy = num_blocks(h, cfg.ssim_window, cfg.ssim_stride);
  J.nchunk = (J.nux + 62) / 63;
  J.calc_l = calc_l;
  J.erp = erp;
  J.part_off = (int32_t)part_off;
  part_off += (size_t)3 * J.nuy * J.nchunk;
  const int idx = g * kMaxSsimJobs + pl.tab[g].njobs;
  pl.nux[idx] = J.nux;
  pl.nuy[idx] = J.nuy;
  pl.max_nchunk[g] = std::max(pl.max_nchunk[g], J.nchunk);
  pl.max_bands[g] = std::max(pl.max_bands[g], (J.nuy + 7) / 8);
  pl.tab[g].njobs++;
  *slot_out = idx;
}

// Builds the SSIM job list of a metric set.
SsimPlan plan_ssim(const ivqm_ctx* c, uint32_t m, size_t& part_total) {
  SsimPlan pl;
  size_t off = 0;
  const ivqm_config& cfg = c->cfg;
  const bool ms_any = (m & (IVQM_M_MSSSIM | IVQM_M_IVMSSSIM)) != 0;
  if (m & IVQM_M_SSIM) add_job(pl, &pl.ssim, c->pic[0], c->pic[1], c->W, c->H, 1, 0, cfg, off);
  if (m & IVQM_M_IVSSIM) {
    add_job(pl, &pl.iv_t2r, c->pic[0], scp_pic(c, 1), c->W, c->H, 1, cfg.erp, cfg, off);
    add_job(pl, &pl.iv_r2t, c->pic[1], scp_pic(c, 0), c->W, c->H, 1, cfg.erp, cfg, off);
  }
  if (ms_any) {
    for (int l = 0; l < c->ns; l++) {
      const int w = c->W >> l, h = c->H >> l, cl = (l == c->ns - 1) ? 1 : 0;
      if (m & IVQM_M_MSSSIM) add_job(pl, &pl.ms[l], level_pic(c, 0, l), level_pic(c, 1, l), w, h, cl, 0, cfg, off);
      if (m & IVQM_M_IVMSSSIM) {
        add_job(pl, &pl.ivms_t2r[l], level_pic(c, 0, l), level_pic(c, 3, l), w, h, cl, 0, cfg, off);
        add_job(pl, &pl.ivms_r2t[l], level_pic(c, 1, l), level_pic(c, 2, l), w, h, cl, 0, cfg, off);
      }
    }
  }
  part_total = off;
  return pl;
}

// Enqueues the device work of one frame pass into ring slot `ring`.
// gcd_given: use these GCD values instead of computing them.
constexpr uint32_t kReuseScp = 1u << 31;  // IV-SSIM metrics use the stored SCP pictures, no new search

int enqueue(ivqm_ctx* c, uint32_t m, int ring, const int32_t* gcd_given, bool search_needed_for_scp_only) {
  const bool reuse_scp = (m & kReuseScp) != 0;
  m &= 0xFFu;
  int rc = check_pair(c);
  if (rc) return rc;
  if (ring < 0 || ring >= kRing) return fail(IVQM_EINVAL, "ring slot out of range");
  if (c->ring_busy[ring]) {
    HIPCHK(hipEventSynchronize(c->ev[ring]));
    c->ring_busy[ring] = false;
  }
  const hipStream_t st = c->stream;
  const int W = c->W, H = c->H;
  const uint32_t iv_ssim_bits = IVQM_M_IVSSIM | IVQM_M_IVMSSSIM;
  const bool need_search =
      (m & IVQM_M_IVPSNR) || (!reuse_scp && (m & iv_ssim_bits)) || search_needed_for_scp_only;
  const bool need_scp = (!reuse_scp && (m & iv_ssim_bits)) || search_needed_for_scp_only;
  const bool need_stats = (m & (IVQM_M_PSNR | IVQM_M_MSE | IVQM_M_WSPSNR)) || (need_search && !gcd_given);
  const bool ms_any = (m & (IVQM_M_MSSSIM | IVQM_M_IVMSSSIM)) != 0;

  if (need_stats) {
    HIPCHK(launch_stats(c->pic[0], c->pic[1], W, H, c->d_row_sd, c->d_row_ssd, c->d_rows, 6 * H, st));
    HIPCHK(launch_gcd_finalize(c->d_row_sd, c->d_row_ssd, W, H, c->pic[0].cs, c->bd, c->maxdiff, c->cfg.gcd_sd_wrap32,
                               c->d_gcd, c->d_sd, c->d_ssd, st));
  }
  if (need_search) {
    bool narrow = c->narrow;
    if (gcd_given) {
      int32_t g4[4] = {gcd_given[0], gcd_given[1], gcd_given[2], 0};
      HIPCHK(hipMemcpyAsync(c->d_gcd, g4, sizeof(g4), hipMemcpyHostToDevice, st));
      int32_t ab[3];
      for (int k = 0; k < 3; k++) ab[k] = std::abs(gcd_given[k]);
      narrow = narrow_ok(c, ab);
      if (!need_stats) HIPCHK(launch_zero_u32(c->d_rows, 6 * H, st));
    }
    HIPCHK(launch_search(c->pic[0], c->pic[1], W, H, c->bd, c->d_gcd, c->cfg.cmp_weights_search, c->R, narrow,
                         c->d_rows, need_scp ? c->d_scp : nullptr, c->pitch, c->best_enabled ? c->d_best : nullptr,
                         c->tile_py, st));
    if (need_scp) c->scp_valid = true;
  }
  if ((m & (IVQM_M_IVSSIM | IVQM_M_IVMSSSIM)) && !c->scp_valid)
    return fail(IVQM_ESTATE, "IV-SSIM needs shift-compensated pictures (ivqm_gen_scp)");

  if (ms_any && c->ns > 1) {
    rc = ensure_pyramid(c);
    if (rc) return rc;
    const bool iv = (m & IVQM_M_IVMSSSIM) != 0;
    for (int l = 1; l < c->ns; l++) {
      const int w = c->W >> (l - 1), h = c->H >> (l - 1);
      for (int which = 0; which < 4; which++) {
        if (which >= 2 && !iv) continue;
        const PicDev src = level_pic(c, which, l - 1);
        uint16_t* dst = c->d_pyr + c->pyr_off[which][l];
        HIPCHK(launch_downsample(src, w, h, dst, c->pyr_pitch[l], st));
      }
    }
  }

  size_t part_total = 0;
  SsimPlan pl = plan_ssim(c, m, part_total);
  if (pl.ngroups) {
    rc = ensure_partial(c, part_total);
    if (rc) return rc;
    const double maxv = (double)c->maxv;
    const double c1 = (0.01 * maxv) * (0.01 * maxv), c2 = (0.03 * maxv) * (0.03 * maxv);  // xSSIM.cpp:50-52
    const float n2 = (float)(c->cfg.ssim_window * c->cfg.ssim_window) * (float)(c->cfg.ssim_window * c->cfg.ssim_window);
    for (int g = 0; g < pl.ngroups; g++) {
      HIPCHK(launch_ssim(pl.tab[g], pl.max_nchunk[g], pl.max_bands[g], (float)(c1 * n2), (float)(c2 * n2), c->bd > 14,
                         c->d_partial, st));
      HIPCHK(launch_ssim_finalize(pl.tab[g], c->d_partial, c->d_erp, c->cfg.ssim_window, c->cfg.ssim_stride,
                                  c->d_ssim_out + (size_t)g * kMaxSsimJobs * 6, st));
    }
  }

  // one copy of everything the epilogue needs
  RingHost* rh = c->h_ring + ring;
  if (need_stats || need_search) HIPCHK(hipMemcpyAsync(rh->gcd, c->d_gcd, sizeof(rh->gcd), hipMemcpyDeviceToHost, st));
  if (need_stats) {
    HIPCHK(hipMemcpyAsync(rh->sd, c->d_sd, sizeof(rh->sd), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(rh->ssd, c->d_ssd, sizeof(rh->ssd), hipMemcpyDeviceToHost, st));
    if (m & IVQM_M_WSPSNR)
      HIPCHK(hipMemcpyAsync(c->h_row_ssd + (size_t)ring * 3 * H, c->d_row_ssd, sizeof(uint64_t) * 3 * H,
                            hipMemcpyDeviceToHost, st));
  }
  if (need_search)
    HIPCHK(hipMemcpyAsync(c->h_rows + (size_t)ring * 6 * H, c->d_rows, sizeof(uint32_t) * 6 * H, hipMemcpyDeviceToHost, st));
  if (pl.ngroups)
    HIPCHK(hipMemcpyAsync(rh->ssim_out, c->d_ssim_out, sizeof(double) * pl.ngroups * kMaxSsimJobs * 6,
                          hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(c->ev[ring], st));
  c->ring_busy[ring] = true;
  c->ring_metrics[ring] = m;
  c->ring_plan[ring] = pl;
  return IVQM_OK;
}

double ssim_value(const ivqm_ctx* c, const SsimPlan& pl, const RingHost* rh, int idx, int comp) {
  const double sum = rh->ssim_out[(size_t)idx * 6 + comp * 2 + 0];
  const double wsum = rh->ssim_out[(size_t)idx * 6 + comp * 2 + 1];
  const int g = idx / kMaxSsimJobs, j = idx % kMaxSsimJobs;
  const SsimJob& J = pl.tab[g].job[j];
  const double active = (double)((int64_t)J.nuy * J.nux);
  if (J.erp) return sum / (active * (wsum / J.nuy));  // xIVSSIM.cpp:158-163
  return sum / active;                                 // xSSIM.cpp:233-236
}

void ms_combine(const ivqm_ctx* c, const SsimPlan& pl, const RingHost* rh, const int* idx, double out[4]) {
  // libs/pmbb-ivqm/src/xSSIM.cpp:132-140: product of ReLU(score)^exponent
  for (int comp = 0; comp < 3; comp++) {
    double v = 1.0;
    for (int l = 0; l < c->ns; l++) {
      const double s = ssim_value(c, pl, rh, idx[l], comp);
      v *= pow(s > 0 ? s : 0.0, kMsExp[c->ns - 1][l]);
    }
    out[comp] = v;
  }
  // lane 3 of the reference vector starts at 1 and is multiplied by pow(ReLU(0), w_i)
  double v3 = 1.0;
  for (int l = 0; l < c->ns; l++) v3 *= pow(0.0, kMsExp[c->ns - 1][l]);
  out[3] = v3;
}

int collect(ivqm_ctx* c, int ring, ivqm_frame_result* out) {
  if (ring < 0 || ring >= kRing) return fail(IVQM_EINVAL, "ring slot out of range");
  if (!c->ring_busy[ring]) return fail(IVQM_ESTATE, "ring slot has no pending frame");
  HIPCHK(hipEventSynchronize(c->ev[ring]));
  c->ring_busy[ring] = false;
  const uint32_t m = c->ring_metrics[ring];
  const SsimPlan& pl = c->ring_plan[ring];
  const RingHost* rh = c->h_ring + ring;
  const int W = c->W, H = c->H;
  const int64_t area = (int64_t)W * H;
  const int32_t* cwa = c->cfg.cmp_weights_average;
  ivqm_frame_result r;
  std::memset(&r, 0, sizeof(r));
  r.metrics = m;
  for (int k = 0; k < 4; k++) r.gcd_r2t[k] = rh->gcd[k];
  for (int k = 0; k < 3; k++) {
    r.ssd[k] = rh->ssd[k];
    r.exact[k] = rh->ssd[k] == 0;
  }
  const double fake = host::psnr_from_ssd(1.0, area, c->bd);
  if (m & (IVQM_M_PSNR | IVQM_M_MSE)) {
    // libs/pmbb-ivqm/src/xPSNR.cpp:131-151; exact components -> fake PSNR (xAppQMIV.cpp:974-981)
    for (int k = 0; k < 4; k++) { r.mse[k] = DBL_MAX; r.psnr[k] = DBL_MAX; }
    for (int k = 0; k < 3; k++) {
      const double ssd = (double)rh->ssd[k];
      r.mse[k] = ssd / (double)area;
      r.psnr[k] = host::psnr_from_ssd(ssd, area, c->bd);
      if (c->cfg.fake_vals_for_exact && r.exact[k]) r.psnr[k] = fake;
    }
    r.psnr_pic = host::pic_value(r.psnr, cwa);
  }
  if (m & IVQM_M_WSPSNR) {
    for (int k = 0; k < 4; k++) r.wspsnr[k] = DBL_MAX;
    if (!c->cfg.erp) {
      for (int k = 0; k < 3; k++) r.wspsnr[k] = host::psnr_from_ssd((double)rh->ssd[k], area, c->bd);
    } else {
      // libs/pmbb-ivqm/src/xWSPSNR.cpp:81-107; a 4:2:0 chroma plane row counts twice per luma row
      const uint64_t* rs = c->h_row_ssd + (size_t)ring * 3 * H;
      const int cs = c->pic[0].cs;
      for (int k = 0; k < 3; k++) {
        host::Kbns kb;
        for (int y = 0; y < H; y++) {
          const uint64_t row = (k && cs) ? 2 * rs[(size_t)k * H + (y >> 1)] : rs[(size_t)k * H + y];
          kb.acc((double)row * c->erp_w[y]);
        }
        const double err = kb.result() * c->erp_corr;
        r.wspsnr[k] = host::psnr_from_ssd(err, area, c->bd);
        if (c->cfg.fake_vals_for_exact && err == 0) r.wspsnr[k] = fake;
      }
    }
    for (int k = 0; k < 3; k++)
      if (c->cfg.fake_vals_for_exact && r.exact[k]) r.wspsnr[k] = fake;
    r.wspsnr_pic = host::pic_value(r.wspsnr, cwa);
  }
  if (m & IVQM_M_IVPSNR) {
    const uint32_t* rows = c->h_rows + (size_t)ring * 6 * H;
    const double* ew = c->cfg.erp ? c->erp_w.data() : nullptr;
    r.ivpsnr_r2t = host::ivpsnr_quality(rows, W, H, c->bd, ew, cwa);
    r.ivpsnr_t2r = host::ivpsnr_quality(rows + (size_t)3 * H, W, H, c->bd, ew, cwa);
    r.ivpsnr = std::min(r.ivpsnr_r2t, r.ivpsnr_t2r);  // xIVPSNR.cpp:55 (xMin = std::min)
  }
  if (m & IVQM_M_SSIM) {
    for (int k = 0; k < 3; k++) r.ssim[k] = ssim_value(c, pl, rh, pl.ssim, k);
    r.ssim[3] = 0.0;
    r.ssim_pic = host::pic_value(r.ssim, cwa);
  }
  if (m & IVQM_M_MSSSIM) {
    ms_combine(c, pl, rh, pl.ms, r.msssim);
    r.msssim_pic = host::pic_value(r.msssim, cwa);
  }
  if (m & IVQM_M_IVSSIM) {
    double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
    for (int k = 0; k < 3; k++) {
      a[k] = ssim_value(c, pl, rh, pl.iv_t2r, k);
      b[k] = ssim_value(c, pl, rh, pl.iv_r2t, k);
    }
    r.ivssim_t2r = host::weighted_mean4(a, cwa);
    r.ivssim_r2t = host::weighted_mean4(b, cwa);
    r.ivssim = std::min(r.ivssim_t2r, r.ivssim_r2t);  // xIVSSIM.cpp:56
  }
  if (m & IVQM_M_IVMSSSIM) {
    double a[4], b[4];
    ms_combine(c, pl, rh, pl.ivms_t2r, a);
    ms_combine(c, pl, rh, pl.ivms_r2t, b);
    a[3] = 0.0;
    b[3] = 0.0;
    r.ivmsssim_t2r = host::weighted_mean4(a, cwa);
    r.ivmsssim_r2t = host::weighted_mean4(b, cwa);
    r.ivmsssim = std::min(r.ivmsssim_t2r, r.ivmsssim_r2t);  // xIVSSIM.cpp:87
  }
  if (out) *out = r;
  return IVQM_OK;
}

int run_sync(ivqm_ctx* c, uint32_t m, const int32_t* gcd_given, bool scp_only, ivqm_frame_result* out) {
  int rc = enqueue(c, m, 0, gcd_given, scp_only);
  if (rc) return rc;
  return collect(c, 0, out);
}

}  // namespace

// ================================ C ABI ====================================
extern "C" {

const char* ivqm_strerror(int s) {
  switch (s) {
    case IVQM_OK: return "ok";
    case IVQM_EINVAL: return "invalid argument";
    case IVQM_ENOMEM: return "out of memory";
    case IVQM_ESTATE: return "call out of order";
    case IVQM_EBUSY: return "busy";
    default: return s <= IVQM_EHIP ? "HIP runtime error" : "unknown error";
  }
}
const char* ivqm_last_error(void) { return g_last_error.c_str(); }

int ivqm_config_init(ivqm_config* cfg) {
  if (!cfg) return fail(IVQM_EINVAL, "cfg is NULL");
  std::memset(cfg, 0, sizeof(*cfg));
  cfg->bit_depth = 10;
  cfg->search_range = 2;
  const int32_t w[4] = {4, 1, 1, 0};
  std::memcpy(cfg->cmp_weights_search, w, sizeof(w));
  std::memcpy(cfg->cmp_weights_average, w, sizeof(w));
  cfg->unnoticeable_coef[0] = cfg->unnoticeable_coef[1] = cfg->unnoticeable_coef[2] = 0.01f;
  cfg->lon_range_deg = 360;
  cfg->lat_range_deg = 180;
  cfg->ssim_window = 8;
  cfg->ssim_stride = 4;
  cfg->gcd_sd_wrap32 = 0;
  cfg->search_tile_rows = 0;
  return IVQM_OK;
}

int ivqm_create(const ivqm_config* cfg, ivqm_ctx** out) {
  if (!cfg || !out) return fail(IVQM_EINVAL, "NULL argument");
  *out = nullptr;
  if (cfg->width < 8 || cfg->height < 8) return fail(IVQM_EINVAL, "picture must be at least 8x8");
  if (cfg->bit_depth < 8 || cfg->bit_depth > 16) return fail(IVQM_EINVAL, "bit_depth must be 8..16");
  if (cfg->search_range < 1 || cfg->search_range > 4) return fail(IVQM_EINVAL, "search_range must be 1..4");
  if (cfg->ssim_window != 8 || cfg->ssim_stride != 4)
    return fail(IVQM_EINVAL, "only the default block-averaged SSIM (window 8, stride 4) is implemented");
  if (cfg->erp && (cfg->lat_range_deg <= 0 || cfg->lat_range_deg > 180)) return fail(IVQM_EINVAL, "lat_range_deg");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return fail(IVQM_EHIP - (int)(e ? e : hipErrorNoDevice), "no HIP device available");
  ivqm_ctx* c = new ivqm_ctx();
  c->cfg = *cfg;
  c->W = cfg->width;
  c->H = cfg->height;
  c->bd = cfg->bit_depth;
  c->maxv = (1 << c->bd) - 1;
  c->R = cfg->search_range;
  c->pitch = round_up(c->W, kPitchAlign);
  for (int k = 0; k < 3; k++) c->maxdiff[k] = host::round_f32(cfg->unnoticeable_coef[k] * (float)c->maxv);  // xGlobClrDiff.cpp:29
  int32_t ab[3] = {std::abs(c->maxdiff[0]), std::abs(c->maxdiff[1]), std::abs(c->maxdiff[2])};
  c->narrow = narrow_ok(c, ab);
  c->tile_py = cfg->search_tile_rows == 4 ? 4 : 8;
  c->ns = ms_num_scales(c->W, c->H);
  for (int i = 0; i < kRing; i++) { c->ring_busy[i] = false; c->ring_metrics[i] = 0; }
  auto cleanup = [&](int rc) { ivqm_destroy(c); return rc; };
#define HIPCHK_C(expr)                                         \
  do {                                                         \
    hipError_t _e = (expr);                                    \
    if (_e != hipSuccess) return cleanup(hip_fail(_e, #expr)); \
  } while (0)
  HIPCHK_C(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  c->own_stream = true;
  const size_t H = c->H;
  HIPCHK_C(hipMalloc(&c->d_row_sd, sizeof(int64_t) * 3 * H));
  HIPCHK_C(hipMalloc(&c->d_row_ssd, sizeof(uint64_t) * 3 * H));
  HIPCHK_C(hipMalloc(&c->d_rows, sizeof(uint32_t) * 6 * H));
  HIPCHK_C(hipMalloc(&c->d_gcd, sizeof(int32_t) * 4));
  HIPCHK_C(hipMalloc(&c->d_sd, sizeof(int64_t) * 3));
  HIPCHK_C(hipMalloc(&c->d_ssd, sizeof(uint64_t) * 3));
  HIPCHK_C(hipMalloc(&c->d_scp, sizeof(uint16_t) * 6 * (size_t)c->pitch * H));
  HIPCHK_C(hipMalloc(&c->d_ssim_out, sizeof(double) * 2 * kMaxSsimJobs * 6));
  HIPCHK_C(hipMemset(c->d_gcd, 0, sizeof(int32_t) * 4));
  HIPCHK_C(hipHostMalloc((void**)&c->h_ring, sizeof(RingHost) * kRing, hipHostMallocDefault));
  HIPCHK_C(hipHostMalloc((void**)&c->h_row_ssd, sizeof(uint64_t) * 3 * H * kRing, hipHostMallocDefault));
  HIPCHK_C(hipHostMalloc((void**)&c->h_rows, sizeof(uint32_t) * 6 * H * kRing, hipHostMallocDefault));
  std::memset(c->h_ring, 0, sizeof(RingHost) * kRing);
  for (int i = 0; i < kRing; i++) HIPCHK_C(hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming));
  host::ws_weights(c->H, cfg->erp ? cfg->lat_range_deg : 180, c->erp_w, c->erp_corr);
  if (cfg->erp) {
    HIPCHK_C(hipMalloc(&c->d_erp, sizeof(double) * H));
    HIPCHK_C(hipMemcpy(c->d_erp, c->erp_w.data(), sizeof(double) * H, hipMemcpyHostToDevice));
  }
#undef HIPCHK_C
  *out = c;
  return IVQM_OK;
}

int ivqm_destroy(ivqm_ctx* c) {
  if (!c) return IVQM_OK;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (int s = 0; s < 2; s++)
    if (c->own[s]) (void)hipFree(c->own[s]);
  void* bufs[] = {c->d_row_sd, c->d_row_ssd, c->d_rows, c->d_gcd, c->d_sd, c->d_ssd, c->d_scp, c->d_best,
                  c->d_pyr, c->d_partial, c->d_ssim_out, c->d_erp};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->h_ring) (void)hipHostFree(c->h_ring);
  if (c->h_row_ssd) (void)hipHostFree(c->h_row_ssd);
  if (c->h_rows) (void)hipHostFree(c->h_rows);
  for (int i = 0; i < kRing; i++)
    if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
  if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return IVQM_OK;
}

int ivqm_set_stream(ivqm_ctx* c, void* s) {
  if (!c) return fail(IVQM_EINVAL, "ctx is NULL");
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->own_stream) (void)hipStreamDestroy(c->stream);
  c->own_stream = false;
  c->stream = (hipStream_t)s;
  if (!s) {
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
  }
  return IVQM_OK;
}

int ivqm_upload(ivqm_ctx* c, int32_t slot, const ivqm_picture* p) {
  if (!c || slot < 0 || slot > 1) return fail(IVQM_EINVAL, "bad context or slot");
  int rc = validate_picture(c, p);
  if (rc) return rc;
  rc = copy_into_slot(c, slot, p, hipMemcpyHostToDevice);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return IVQM_OK;
}

int ivqm_bind(ivqm_ctx* c, int32_t slot, const ivqm_picture* p) {
  if (!c || slot < 0 || slot > 1) return fail(IVQM_EINVAL, "bad context or slot");
  int rc = validate_picture(c, p);
  if (rc) return rc;
  if (!aligned_for_kernels(p)) return copy_into_slot(c, slot, p, hipMemcpyDeviceToDevice);
  PicDev d;
  for (int k = 0; k < 3; k++) {
    d.p[k] = p->plane[k];
    d.stride[k] = p->stride[k];
  }
  d.cs = p->chroma_420 ? 1 : 0;
  d.bps = p->bytes_per_sample;
  c->pic[slot] = d;
  c->bound[slot] = true;
  c->scp_valid = false;
  return IVQM_OK;
}

int ivqm_gcd(ivqm_ctx* c, int32_t g[4]) {
  if (!c || !g) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_MSE, nullptr, false, &r);
  if (rc) return rc;
  std::memcpy(g, r.gcd_r2t, sizeof(r.gcd_r2t));
  return IVQM_OK;
}
int ivqm_ssd(ivqm_ctx* c, uint64_t s[4]) {
  if (!c || !s) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_MSE, nullptr, false, &r);
  if (rc) return rc;
  for (int k = 0; k < 3; k++) s[k] = r.ssd[k];
  s[3] = 0;
  return IVQM_OK;
}
int ivqm_mse(ivqm_ctx* c, double v[4]) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_MSE, nullptr, false, &r);
  if (rc) return rc;
  std::memcpy(v, r.mse, sizeof(r.mse));
  return IVQM_OK;
}
int ivqm_psnr(ivqm_ctx* c, double v[4]) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_PSNR, nullptr, false, &r);
  if (rc) return rc;
  std::memcpy(v, r.psnr, sizeof(r.psnr));
  return IVQM_OK;
}
int ivqm_wspsnr(ivqm_ctx* c, double v[4]) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_WSPSNR, nullptr, false, &r);
  if (rc) return rc;
  std::memcpy(v, r.wspsnr, sizeof(r.wspsnr));
  return IVQM_OK;
}
int ivqm_ivpsnr(ivqm_ctx* c, const int32_t* gcd, double* v, double* r2t, double* t2r) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_IVPSNR, gcd, false, &r);
  if (rc) return rc;
  *v = r.ivpsnr;
  if (r2t) *r2t = r.ivpsnr_r2t;
  if (t2r) *t2r = r.ivpsnr_t2r;
  return IVQM_OK;
}
int ivqm_gen_scp(ivqm_ctx* c, const int32_t* gcd) {
  if (!c) return fail(IVQM_EINVAL, "NULL argument");
  return run_sync(c, 0, gcd, true, nullptr);
}
int ivqm_ssim(ivqm_ctx* c, double v[4]) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_SSIM, nullptr, false, &r);
  if (rc) return rc;
  std::memcpy(v, r.ssim, sizeof(r.ssim));
  return IVQM_OK;
}
int ivqm_msssim(ivqm_ctx* c, double v[4]) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_MSSSIM, nullptr, false, &r);
  if (rc) return rc;
  std::memcpy(v, r.msssim, sizeof(r.msssim));
  return IVQM_OK;
}
int ivqm_ivssim(ivqm_ctx* c, double* v, double* r2t, double* t2r) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_IVSSIM | kReuseScp, nullptr, false, &r);
  if (rc) return rc;
  *v = r.ivssim;
  if (r2t) *r2t = r.ivssim_r2t;
  if (t2r) *t2r = r.ivssim_t2r;
  return IVQM_OK;
}
int ivqm_ivmsssim(ivqm_ctx* c, double* v, double* r2t, double* t2r) {
  if (!c || !v) return fail(IVQM_EINVAL, "NULL argument");
  ivqm_frame_result r;
  int rc = run_sync(c, IVQM_M_IVMSSSIM | kReuseScp, nullptr, false, &r);
  if (rc) return rc;
  *v = r.ivmsssim;
  if (r2t) *r2t = r.ivmsssim_r2t;
  if (t2r) *t2r = r.ivmsssim_t2r;
  return IVQM_OK;
}

int ivqm_frame(ivqm_ctx* c, uint32_t m, ivqm_frame_result* out) {
  if (!c || !out) return fail(IVQM_EINVAL, "NULL argument");
  return run_sync(c, m & 0xFFu, nullptr, false, out);
}
int ivqm_frame_enqueue(ivqm_ctx* c, uint32_t m, int32_t ring) {
  if (!c) return fail(IVQM_EINVAL, "NULL argument");
  return enqueue(c, m & 0xFFu, ring, nullptr, false);
}
int ivqm_frame_collect(ivqm_ctx* c, int32_t ring, ivqm_frame_result* out) {
  if (!c) return fail(IVQM_EINVAL, "NULL argument");
  return collect(c, ring, out);
}

int ivqm_get_rows(ivqm_ctx* c, int32_t dir, uint64_t* rows) {
  if (!c || !rows || dir < 0 || dir > 1) return fail(IVQM_EINVAL, "bad argument");
  std::vector<uint32_t> h((size_t)3 * c->H);
  HIPCHK(hipMemcpyAsync(h.data(), c->d_rows + (size_t)dir * 3 * c->H, sizeof(uint32_t) * 3 * c->H,
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int y = 0; y < c->H; y++) {
    for (int k = 0; k < 3; k++) rows[(size_t)y * 4 + k] = (uint64_t)(int64_t)(int32_t)h[(size_t)k * c->H + y];
    rows[(size_t)y * 4 + 3] = 0;
  }
  return IVQM_OK;
}
int ivqm_get_scp(ivqm_ctx* c, int32_t idx, uint16_t* const planes[3]) {
  if (!c || !planes || idx < 0 || idx > 1) return fail(IVQM_EINVAL, "bad argument");
  if (!c->scp_valid) return fail(IVQM_ESTATE, "no shift-compensated pictures computed");
  const size_t plane = (size_t)c->pitch * c->H;
  for (int k = 0; k < 3; k++)
    HIPCHK(hipMemcpy2DAsync(planes[k], sizeof(uint16_t) * c->W, c->d_scp + ((size_t)idx * 3 + k) * plane,
                            sizeof(uint16_t) * c->pitch, sizeof(uint16_t) * c->W, c->H, hipMemcpyDeviceToHost,
                            c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return IVQM_OK;
}
int ivqm_enable_best_map(ivqm_ctx* c, int32_t en) {
  if (!c) return fail(IVQM_EINVAL, "NULL argument");
  if (en && !c->d_best) HIPCHK(hipMalloc(&c->d_best, (size_t)2 * c->W * c->H));
  c->best_enabled = en != 0;
  return IVQM_OK;
}
int ivqm_get_best_map(ivqm_ctx* c, int32_t dir, uint8_t* map) {
  if (!c || !map || dir < 0 || dir > 1) return fail(IVQM_EINVAL, "bad argument");
  if (!c->d_best) return fail(IVQM_ESTATE, "best map not enabled");
  HIPCHK(hipMemcpyAsync(map, c->d_best + (size_t)dir * c->W * c->H, (size_t)c->W * c->H, hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return IVQM_OK;
}

static int pic_from(const ivqm_picture* p, PicDev& d) {
  if (!p || (p->bytes_per_sample != 1 && p->bytes_per_sample != 2)) return fail(IVQM_EINVAL, "bad picture");
  for (int k = 0; k < 3; k++) {
    if (!p->plane[k]) return fail(IVQM_EINVAL, "plane pointer is NULL");
    d.p[k] = p->plane[k];
    d.stride[k] = p->stride[k];
  }
  d.cs = p->chroma_420 ? 1 : 0;
  d.bps = p->bytes_per_sample;
  return IVQM_OK;
}

int ivqm_upsample420(const ivqm_picture* src, int32_t w, int32_t h, uint16_t* dst, int32_t dst_stride, void* stream) {
  PicDev d;
  int rc = pic_from(src, d);
  if (rc) return rc;
  if (!dst || w <= 0 || h <= 0 || dst_stride < w || (d.cs && ((w & 1) || (h & 1)))) return fail(IVQM_EINVAL, "bad size");
  HIPCHK(launch_upsample420(d, w, h, dst, dst_stride, (hipStream_t)stream));
  return IVQM_OK;
}
int ivqm_downsample(const ivqm_picture* src, int32_t w, int32_t h, uint16_t* dst, void* stream) {
  PicDev d;
  int rc = pic_from(src, d);
  if (rc) return rc;
  if (!dst || w < 2 || h < 2) return fail(IVQM_EINVAL, "bad size");
  HIPCHK(launch_downsample(d, w, h, dst, w >> 1, (hipStream_t)stream));
  return IVQM_OK;
}

}  // extern "C"

// ============================ profiling hook ================================
// Times `iters` back-to-back launches of one kernel family of a frame pass on
// the context's stream with HIP events (diagnostics; not a reference interface).
extern "C" int ivqm_profile(ivqm_ctx* c, int32_t kernel, int32_t iters, double* avg_ms) {
  if (!c || !avg_ms || iters <= 0) return fail(IVQM_EINVAL, "bad argument");
  int rc = check_pair(c);
  if (rc) return rc;
  // make sure GCD / SCP exist for the dependent kernels
  rc = run_sync(c, IVQM_M_IVPSNR | IVQM_M_IVSSIM, nullptr, false, nullptr);
  if (rc) return rc;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  size_t part_total = 0;
  SsimPlan pl = plan_ssim(c, IVQM_M_IVSSIM, part_total);
  const double maxv = (double)c->maxv;
  const double c1 = (0.01 * maxv) * (0.01 * maxv), c2 = (0.03 * maxv) * (0.03 * maxv);
  const float n2 = 4096.0f;
  HIPCHK(hipEventRecord(e0, c->stream));
  for (int i = 0; i < iters; i++) {
    switch (kernel) {
      case 0:
        HIPCHK(launch_stats(c->pic[0], c->pic[1], c->W, c->H, c->d_row_sd, c->d_row_ssd, c->d_rows, 6 * c->H, c->stream));
        break;
      case 1:
        HIPCHK(launch_search(c->pic[0], c->pic[1], c->W, c->H, c->bd, c->d_gcd, c->cfg.cmp_weights_search, c->R,
                             c->narrow, c->d_rows, c->d_scp, c->pitch, nullptr, c->tile_py, c->stream));
        break;
      case 2:
        HIPCHK(launch_ssim(pl.tab[0], pl.max_nchunk[0], pl.max_bands[0], (float)(c1 * n2), (float)(c2 * n2), c->bd > 14,
                           c->d_partial, c->stream));
        break;
      default:
        return fail(IVQM_EINVAL, "unknown kernel id");
    }
  }
  HIPCHK(hipEventRecord(e1, c->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / iters;
  return IVQM_OK;
}
