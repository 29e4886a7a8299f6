// C ABI (include/ivqm_hip.h, ivqm_batch.h, ivqm_hip_profile.h) over the batch
// engine (ivqm_engine.h).
//
// A context = one picture size / bit depth / parameter set with two picture
// slots, as the reference configures one set of processor objects per
// sequence (apps/QMIV/src/xAppQMIV.cpp:680-744); each per-frame call is a
// batch of one pair. A batch object takes many pairs of mixed sizes per call.
// Nothing on the device path falls back to the CPU: a missing or failing HIP
// runtime is reported as an error status.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/ivqm_batch.h"
#include "../../include/ivqm_hip.h"
#include "../../include/ivqm_hip_profile.h"
#include "ivqm_abi_util.h"
#include "ivqm_engine.h"
#include "ivqm_epilogue.h"

using namespace ivqm;

namespace ivqm {
const char* last_error_message();
}

#define HIPCHK(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_error(_e, #expr); \
  } while (0)
#define RCCHK(expr)      \
  do {                   \
    int _rc = (expr);    \
    if (_rc) return _rc; \
  } while (0)

struct ivqm_ctx {
  ivqm_config cfg;
  std::unique_ptr<Engine> eng;
  PicDev pic[2];
  bool bound[2] = {false, false};
  void* own[2] = {nullptr, nullptr};
  size_t own_bytes[2] = {0, 0};
  bool best_enabled = false;
  // mask mode
  const void* mask = nullptr;
  int mask_stride = 0, mask_bps = 0, mask_bd = 0;
  void* own_mask = nullptr;
  size_t own_mask_bytes = 0;
};

// Two engines, each with its own stream, device buffers and captured graphs:
// even ring slots run on the first, odd ones on the second, so consecutive
// batches enqueued on slots k and k + 1 overlap on the device (one batch's
// HBM-bound phases beside the other's VALU-bound search).
struct ivqm_batch {
  ivqm_config cfg;
  std::unique_ptr<Engine> eng, eng2;
  bool use_graph = false;
  Engine& lane(int ring) { return (ring & 1) ? *eng2 : *eng; }
};

namespace {

// Copies a picture (host or device) into the slot's own pitched buffer.
int copy_into_slot(ivqm_ctx* c, int slot, const ivqm_picture* p, hipMemcpyKind kind) {
  const int W = c->cfg.width, H = c->cfg.height;
  const PicDev f = to_dev(p);  // chroma shifts
  const int cs = f.cs, csy = f.csy, bps = p->bytes_per_sample;
  const int pitch = Engine::scp_pitch(W), pitch_c = Engine::scp_pitch(W >> cs);
  const size_t bytes = ((size_t)pitch * H + 2 * (size_t)pitch_c * (H >> csy)) * bps;
  RCCHK(c->eng->sync());
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
    const int wp = k ? (W >> cs) : W, hp = k ? (H >> csy) : H;
    const int pp = k ? pitch_c : pitch;
    d.p[k] = base + off;
    d.stride[k] = pp;
    HIPCHK(hipMemcpy2DAsync(base + off, (size_t)pp * bps, p->plane[k], (size_t)p->stride[k] * bps, (size_t)wp * bps,
                            hp, kind, c->eng->stream()));
    off += (size_t)pp * hp * bps;
  }
  d.cs = cs;
  d.csy = csy;
  d.bps = bps;
  c->pic[slot] = d;
  c->bound[slot] = true;
  c->eng->set_scp_valid(false);
  return IVQM_OK;
}

int validate_mask(int W, int H, const ivqm_mask* m) {
  if (!m || !m->plane) return set_error(IVQM_EINVAL, "mask is NULL");
  if (m->bit_depth < 8 || m->bit_depth > 14) return set_error(IVQM_EINVAL, "mask bit_depth must be 8..14 (BitDepthM)");
  if (m->bytes_per_sample != 1 && m->bytes_per_sample != 2) return set_error(IVQM_EINVAL, "mask bytes_per_sample");
  if (m->bytes_per_sample == 1 && m->bit_depth != 8) return set_error(IVQM_EINVAL, "8-bit mask samples require bit_depth 8");
  if (m->stride < W || W <= 0 || H <= 0) return set_error(IVQM_EINVAL, "mask stride smaller than the width");
  return IVQM_OK;
}

void mask_into_spec(const ivqm_mask& m, PairSpec& s) {
  s.msk = m.plane;
  s.msk_stride = m.stride;
  s.msk_bps = m.bytes_per_sample;
  s.msk_bd = m.bit_depth;
}

int ctx_pair(ivqm_ctx* c, PairSpec& s, const int32_t* gcd) {
  if (!c->bound[0] || !c->bound[1]) return set_error(IVQM_ESTATE, "both picture slots must be bound");
  if (chroma_format(c->pic[0]) != chroma_format(c->pic[1]) || c->pic[0].bps != c->pic[1].bps)
    return set_error(IVQM_EINVAL, "both pictures must share chroma format and sample size");
  s.p0 = c->pic[0];
  s.p1 = c->pic[1];
  s.w = c->cfg.width;
  s.h = c->cfg.height;
  s.bd = c->cfg.bit_depth;
  s.has_gcd = gcd != nullptr;
  if (gcd)
    for (int k = 0; k < 4; k++) s.gcd[k] = k < 3 ? gcd[k] : 0;
  if (c->mask) {
    s.msk = c->mask;
    s.msk_stride = c->mask_stride;
    s.msk_bps = c->mask_bps;
    s.msk_bd = c->mask_bd;
  }
  return IVQM_OK;
}

int run_sync(ivqm_ctx* c, uint32_t m, const int32_t* gcd, ivqm_frame_result* out) {
  PairSpec s;
  RCCHK(ctx_pair(c, s, gcd));
  c->eng->enable_best(c->best_enabled);
  RCCHK(c->eng->enqueue(&s, 1, m, 0));
  ivqm_frame_result r;
  RCCHK(c->eng->collect(0, &r));
  if (out) *out = r;
  return IVQM_OK;
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
const char* ivqm_last_error(void) { return ivqm::last_error_message(); }

int ivqm_config_init(ivqm_config* cfg) {
  if (!cfg) return set_error(IVQM_EINVAL, "cfg is NULL");
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
  cfg->ssim_mode = 4;  // BlockAveraged (xSSIM.h:35)
  cfg->ssim_border = 0;
  return IVQM_OK;
}

int ivqm_create(const ivqm_config* cfg, ivqm_ctx** out) {
  return guarded([&]() -> int {
    if (!cfg || !out) return set_error(IVQM_EINVAL, "NULL argument");
    *out = nullptr;
    RCCHK(validate_config(cfg, true));
    std::unique_ptr<ivqm_ctx> c(new ivqm_ctx());
    c->cfg = *cfg;
    c->eng.reset(new Engine(*cfg));
    RCCHK(c->eng->init());
    *out = c.release();
    return IVQM_OK;
  });
}

int ivqm_destroy(ivqm_ctx* c) {
  return guarded([&]() -> int {
    if (!c) return IVQM_OK;
    if (c->eng) (void)c->eng->sync();
    for (int s = 0; s < 2; s++)
      if (c->own[s]) (void)hipFree(c->own[s]);
    if (c->own_mask) (void)hipFree(c->own_mask);
    delete c;
    return IVQM_OK;
  });
}

int ivqm_set_stream(ivqm_ctx* c, void* s) {
  return guarded([&]() -> int {
    if (!c) return set_error(IVQM_EINVAL, "ctx is NULL");
    return c->eng->set_stream((hipStream_t)s);
  });
}

int ivqm_upload(ivqm_ctx* c, int32_t slot, const ivqm_picture* p) {
  return guarded([&]() -> int {
    if (!c || slot < 0 || slot > 1) return set_error(IVQM_EINVAL, "bad context or slot");
    RCCHK(validate_picture(c->cfg.width, c->cfg.height, c->cfg.bit_depth, p));
    RCCHK(copy_into_slot(c, slot, p, hipMemcpyHostToDevice));
    return c->eng->sync();
  });
}

int ivqm_bind(ivqm_ctx* c, int32_t slot, const ivqm_picture* p) {
  return guarded([&]() -> int {
    if (!c || slot < 0 || slot > 1) return set_error(IVQM_EINVAL, "bad context or slot");
    RCCHK(validate_picture(c->cfg.width, c->cfg.height, c->cfg.bit_depth, p));
    if (!aligned_for_kernels(p)) return copy_into_slot(c, slot, p, hipMemcpyDeviceToDevice);
    c->pic[slot] = to_dev(p);
    c->bound[slot] = true;
    c->eng->set_scp_valid(false);
    return IVQM_OK;
  });
}

int ivqm_bind_mask(ivqm_ctx* c, const ivqm_mask* m) {
  return guarded([&]() -> int {
    if (!c) return set_error(IVQM_EINVAL, "NULL context");
    if (!m) {
      c->mask = nullptr;
      return IVQM_OK;
    }
    RCCHK(validate_mask(c->cfg.width, c->cfg.height, m));
    c->mask = m->plane;
    c->mask_stride = m->stride;
    c->mask_bps = m->bytes_per_sample;
    c->mask_bd = m->bit_depth;
    return IVQM_OK;
  });
}

int ivqm_upload_mask(ivqm_ctx* c, const ivqm_mask* m) {
  return guarded([&]() -> int {
    if (!c) return set_error(IVQM_EINVAL, "NULL context");
    RCCHK(validate_mask(c->cfg.width, c->cfg.height, m));
    const int W = c->cfg.width, H = c->cfg.height, bps = m->bytes_per_sample;
    const size_t bytes = (size_t)W * H * bps;
    RCCHK(c->eng->sync());
    if (c->own_mask_bytes < bytes) {
      if (c->own_mask) (void)hipFree(c->own_mask);
      c->own_mask = nullptr;
      c->own_mask_bytes = 0;
      HIPCHK(hipMalloc(&c->own_mask, bytes));
      c->own_mask_bytes = bytes;
    }
    HIPCHK(hipMemcpy2DAsync(c->own_mask, (size_t)W * bps, m->plane, (size_t)m->stride * bps, (size_t)W * bps, H,
                            hipMemcpyHostToDevice, c->eng->stream()));
    c->mask = c->own_mask;
    c->mask_stride = W;
    c->mask_bps = bps;
    c->mask_bd = m->bit_depth;
    return c->eng->sync();
  });
}

int ivqm_gcd(ivqm_ctx* c, int32_t g[4]) {
  return guarded([&]() -> int {
    if (!c || !g) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_MSE, nullptr, &r));
    std::memcpy(g, r.gcd_r2t, sizeof(r.gcd_r2t));
    return IVQM_OK;
  });
}
int ivqm_ssd(ivqm_ctx* c, uint64_t s[4]) {
  return guarded([&]() -> int {
    if (!c || !s) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_MSE, nullptr, &r));
    for (int k = 0; k < 3; k++) s[k] = r.ssd[k];
    s[3] = 0;
    return IVQM_OK;
  });
}
int ivqm_mse(ivqm_ctx* c, double v[4]) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_MSE, nullptr, &r));
    std::memcpy(v, r.mse, sizeof(r.mse));
    return IVQM_OK;
  });
}
int ivqm_psnr(ivqm_ctx* c, double v[4]) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_PSNR, nullptr, &r));
    std::memcpy(v, r.psnr, sizeof(r.psnr));
    return IVQM_OK;
  });
}
int ivqm_wspsnr(ivqm_ctx* c, double v[4]) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_WSPSNR, nullptr, &r));
    std::memcpy(v, r.wspsnr, sizeof(r.wspsnr));
    return IVQM_OK;
  });
}
int ivqm_ivpsnr(ivqm_ctx* c, const int32_t* gcd, double* v, double* r2t, double* t2r) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_IVPSNR, gcd, &r));
    *v = r.ivpsnr;
    if (r2t) *r2t = r.ivpsnr_r2t;
    if (t2r) *t2r = r.ivpsnr_t2r;
    return IVQM_OK;
  });
}
int ivqm_gen_scp(ivqm_ctx* c, const int32_t* gcd) {
  return guarded([&]() -> int {
    if (!c) return set_error(IVQM_EINVAL, "NULL argument");
    return run_sync(c, kScpOnly, gcd, nullptr);
  });
}
int ivqm_ssim(ivqm_ctx* c, double v[4]) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_SSIM, nullptr, &r));
    std::memcpy(v, r.ssim, sizeof(r.ssim));
    return IVQM_OK;
  });
}
int ivqm_msssim(ivqm_ctx* c, double v[4]) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_MSSSIM, nullptr, &r));
    std::memcpy(v, r.msssim, sizeof(r.msssim));
    return IVQM_OK;
  });
}
int ivqm_ivssim(ivqm_ctx* c, double* v, double* r2t, double* t2r) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_IVSSIM | kReuseScp, nullptr, &r));
    *v = r.ivssim;
    if (r2t) *r2t = r.ivssim_r2t;
    if (t2r) *t2r = r.ivssim_t2r;
    return IVQM_OK;
  });
}
int ivqm_ivmsssim(ivqm_ctx* c, double* v, double* r2t, double* t2r) {
  return guarded([&]() -> int {
    if (!c || !v) return set_error(IVQM_EINVAL, "NULL argument");
    ivqm_frame_result r;
    RCCHK(run_sync(c, IVQM_M_IVMSSSIM | kReuseScp, nullptr, &r));
    *v = r.ivmsssim;
    if (r2t) *r2t = r.ivmsssim_r2t;
    if (t2r) *t2r = r.ivmsssim_t2r;
    return IVQM_OK;
  });
}
int ivqm_upload_scp(ivqm_ctx* c, int32_t idx, const ivqm_picture* p) {
  return guarded([&]() -> int {
    if (!c || idx < 0 || idx > 1) return set_error(IVQM_EINVAL, "bad context or index");
    const int W = c->cfg.width, H = c->cfg.height;
    RCCHK(validate_picture(W, H, c->cfg.bit_depth, p));
    if (p->chroma_420 || p->bytes_per_sample != 2) return set_error(IVQM_EINVAL, "SCP pictures are 4:4:4 uint16");
    const int pitch = Engine::scp_pitch(W);
    RCCHK(c->eng->ensure_scp(6 * (int64_t)pitch * H));
    const size_t plane = (size_t)pitch * H;
    for (int k = 0; k < 3; k++)
      HIPCHK(hipMemcpy2DAsync(c->eng->d_scp() + ((size_t)idx * 3 + k) * plane, sizeof(uint16_t) * pitch, p->plane[k],
                              sizeof(uint16_t) * p->stride[k], sizeof(uint16_t) * W, H, hipMemcpyHostToDevice,
                              c->eng->stream()));
    RCCHK(c->eng->sync());
    c->eng->set_scp_valid(true);
    return IVQM_OK;
  });
}

int ivqm_frame(ivqm_ctx* c, uint32_t m, ivqm_frame_result* out) {
  return guarded([&]() -> int {
    if (!c || !out) return set_error(IVQM_EINVAL, "NULL argument");
    return run_sync(c, m & 0xFFu, nullptr, out);
  });
}
int ivqm_frame_enqueue(ivqm_ctx* c, uint32_t m, int32_t ring) {
  return guarded([&]() -> int {
    if (!c) return set_error(IVQM_EINVAL, "NULL argument");
    PairSpec s;
    RCCHK(ctx_pair(c, s, nullptr));
    c->eng->enable_best(c->best_enabled);
    return c->eng->enqueue(&s, 1, m & 0xFFu, ring);
  });
}
int ivqm_frame_collect(ivqm_ctx* c, int32_t ring, ivqm_frame_result* out) {
  return guarded([&]() -> int {
    if (!c) return set_error(IVQM_EINVAL, "NULL argument");
    return c->eng->collect(ring, out);
  });
}

int ivqm_get_rows(ivqm_ctx* c, int32_t dir, uint64_t* rows) {
  return guarded([&]() -> int {
    if (!c || !rows || dir < 0 || dir > 1) return set_error(IVQM_EINVAL, "bad argument");
    const BatchPlan& P = c->eng->last_plan();
    if (P.n != 1 || !P.need_search) return set_error(IVQM_ESTATE, "no search ran on this context");
    const int H = c->cfg.height;
    std::vector<uint32_t> h((size_t)3 * H);
    RCCHK(c->eng->sync());
    HIPCHK(hipMemcpy(h.data(), c->eng->d_rows() + P.pp[0].rows_off + (size_t)dir * 3 * H, sizeof(uint32_t) * 3 * H,
                     hipMemcpyDeviceToHost));
    for (int y = 0; y < H; y++) {
      for (int k = 0; k < 3; k++) rows[(size_t)y * 4 + k] = (uint64_t)(int64_t)(int32_t)h[(size_t)k * H + y];
      rows[(size_t)y * 4 + 3] = 0;
    }
    return IVQM_OK;
  });
}
int ivqm_get_scp(ivqm_ctx* c, int32_t idx, uint16_t* const planes[3]) {
  return guarded([&]() -> int {
    if (!c || !planes || idx < 0 || idx > 1) return set_error(IVQM_EINVAL, "bad argument");
    if (!c->eng->scp_valid()) return set_error(IVQM_ESTATE, "no shift-compensated pictures computed");
    const int W = c->cfg.width, H = c->cfg.height, pitch = Engine::scp_pitch(W);
    const size_t plane = (size_t)pitch * H;
    RCCHK(c->eng->sync());
    for (int k = 0; k < 3; k++)
      HIPCHK(hipMemcpy2D(planes[k], sizeof(uint16_t) * W, c->eng->d_scp() + ((size_t)idx * 3 + k) * plane,
                         sizeof(uint16_t) * pitch, sizeof(uint16_t) * W, H, hipMemcpyDeviceToHost));
    return IVQM_OK;
  });
}
int ivqm_enable_best_map(ivqm_ctx* c, int32_t en) {
  return guarded([&]() -> int {
    if (!c) return set_error(IVQM_EINVAL, "NULL argument");
    c->best_enabled = en != 0;
    return IVQM_OK;
  });
}
int ivqm_get_best_map(ivqm_ctx* c, int32_t dir, uint8_t* map) {
  return guarded([&]() -> int {
    if (!c || !map || dir < 0 || dir > 1) return set_error(IVQM_EINVAL, "bad argument");
    const BatchPlan& P = c->eng->last_plan();
    if (!c->eng->d_best() || P.n != 1 || P.pairs.empty() || P.pairs[0].best_off < 0 || !P.need_search)
      return set_error(IVQM_ESTATE, "best map not enabled for the last search");
    const size_t n = (size_t)c->cfg.width * c->cfg.height;
    RCCHK(c->eng->sync());
    HIPCHK(hipMemcpy(map, c->eng->d_best() + (size_t)dir * n, n, hipMemcpyDeviceToHost));
    return IVQM_OK;
  });
}

int ivqm_profile(ivqm_ctx* c, int32_t kernel, int32_t iters, double* avg_ms) {
  return guarded([&]() -> int {
    if (!c || !avg_ms || iters <= 0) return set_error(IVQM_EINVAL, "bad argument");
    return c->eng->profile(kernel, iters, avg_ms);
  });
}

static int pic_from(const ivqm_picture* p, PicDev& d) {
  if (!p || (p->bytes_per_sample != 1 && p->bytes_per_sample != 2)) return set_error(IVQM_EINVAL, "bad picture");
  if (p->chroma_420 < 0 || p->chroma_420 > 2) return set_error(IVQM_EINVAL, "bad chroma format");
  for (int k = 0; k < 3; k++)
    if (!p->plane[k]) return set_error(IVQM_EINVAL, "plane pointer is NULL");
  d = to_dev(p);
  return IVQM_OK;
}
int ivqm_upsample420(const ivqm_picture* src, int32_t w, int32_t h, uint16_t* dst, int32_t dst_stride, void* stream) {
  return guarded([&]() -> int {
    PicDev d;
    RCCHK(pic_from(src, d));
    if (!dst || w <= 0 || h <= 0 || dst_stride < w || (d.cs && (w & 1)) || (d.csy && (h & 1)))
      return set_error(IVQM_EINVAL, "bad size");
    HIPCHK(launch_upsample420(d, w, h, dst, dst_stride, (hipStream_t)stream));
    return IVQM_OK;
  });
}
int ivqm_convert_colour(const ivqm_picture* src, int32_t w, int32_t h, int32_t bd, int32_t direction, int32_t space,
                        uint16_t* dst, int32_t dst_stride, void* stream) {
  return guarded([&]() -> int {
    PicDev d;
    RCCHK(pic_from(src, d));
    if (!dst || w <= 0 || h <= 0 || dst_stride < w || (d.cs && (w & 1)) || (d.csy && (h & 1)))
      return set_error(IVQM_EINVAL, "bad size");
    if (bd < 8 || bd > 16 || (d.bps == 1 && bd != 8)) return set_error(IVQM_EINVAL, "bad bit depth");
    if (direction != 0 && direction != 1) return set_error(IVQM_EINVAL, "direction must be 0 or 1");
    if (space < 0 || space > 3) return set_error(IVQM_EINVAL, "space must be 0..3");
    HIPCHK(launch_colour(d, w, h, bd, direction, space, dst, dst_stride, (hipStream_t)stream));
    return IVQM_OK;
  });
}
int ivqm_downsample(const ivqm_picture* src, int32_t w, int32_t h, uint16_t* dst, void* stream) {
  return guarded([&]() -> int {
    PicDev d;
    RCCHK(pic_from(src, d));
    if (!dst || w < 2 || h < 2) return set_error(IVQM_EINVAL, "bad size");
    HIPCHK(launch_downsample_one(d, w, h, dst, w >> 1, (hipStream_t)stream));
    return IVQM_OK;
  });
}

// ------------------------------- batches -----------------------------------

int ivqm_batch_create(const ivqm_config* cfg, ivqm_batch** out) {
  return guarded([&]() -> int {
    if (!cfg || !out) return set_error(IVQM_EINVAL, "NULL argument");
    *out = nullptr;
    RCCHK(validate_config(cfg, false));
    std::unique_ptr<ivqm_batch> b(new ivqm_batch());
    b->cfg = *cfg;
    b->eng.reset(new Engine(*cfg));
    RCCHK(b->eng->init());
    b->eng2.reset(new Engine(*cfg));
    RCCHK(b->eng2->init());
    *out = b.release();
    return IVQM_OK;
  });
}
int ivqm_batch_destroy(ivqm_batch* b) {
  return guarded([&]() -> int {
    if (!b) return IVQM_OK;
    if (b->eng) (void)b->eng->sync();
    if (b->eng2) (void)b->eng2->sync();
    delete b;
    return IVQM_OK;
  });
}
int ivqm_batch_set_stream(ivqm_batch* b, void* s) {
  return guarded([&]() -> int {
    if (!b) return set_error(IVQM_EINVAL, "NULL argument");
    // a caller's stream serializes both lanes on it
    RCCHK(b->eng->set_stream((hipStream_t)s));
    return b->eng2->set_stream((hipStream_t)s);
  });
}
int ivqm_batch_set_tuning(ivqm_batch* b, int32_t search_tile_rows, int32_t ssim_band_rows, int32_t use_graph) {
  return guarded([&]() -> int {
    if (!b) return set_error(IVQM_EINVAL, "NULL argument");
    b->eng->set_tuning(search_tile_rows, ssim_band_rows);
    b->eng2->set_tuning(search_tile_rows, ssim_band_rows);
    b->use_graph = use_graph != 0;
    return IVQM_OK;
  });
}
int ivqm_batch_get_tuning(ivqm_batch* b, int32_t* search_tile_rows, int32_t* ssim_band_rows) {
  return guarded([&]() -> int {
    if (!b) return set_error(IVQM_EINVAL, "NULL argument");
    if (search_tile_rows) *search_tile_rows = b->eng->tuning_py();
    if (ssim_band_rows) *ssim_band_rows = b->eng->tuning_nb();
    return IVQM_OK;
  });
}

static int batch_specs(const ivqm_pair* pairs, int32_t n, std::vector<PairSpec>& specs) {
  if (!pairs || n <= 0) return set_error(IVQM_EINVAL, "empty batch");
  specs.resize(n);
  for (int i = 0; i < n; i++) {
    const ivqm_pair& p = pairs[i];
    if (p.width < 8 || p.height < 8 || p.bit_depth < 8 || p.bit_depth > 16)
      return set_error(IVQM_EINVAL, "bad pair size or bit depth");
    RCCHK(validate_picture(p.width, p.height, p.bit_depth, &p.p0));
    RCCHK(validate_picture(p.width, p.height, p.bit_depth, &p.p1));
    if (p.p0.chroma_420 != p.p1.chroma_420 || p.p0.bytes_per_sample != p.p1.bytes_per_sample)
      return set_error(IVQM_EINVAL, "pictures of a pair must share chroma format and sample size");
    if (!aligned_for_kernels(&p.p0) || !aligned_for_kernels(&p.p1))
      return set_error(IVQM_EINVAL, "batch pictures need 4-sample-aligned full-resolution planes and strides");
    PairSpec& s = specs[i];
    s.p0 = to_dev(&p.p0);
    s.p1 = to_dev(&p.p1);
    s.w = p.width;
    s.h = p.height;
    s.bd = p.bit_depth;
    s.has_gcd = false;
  }
  return IVQM_OK;
}

int ivqm_batch_enqueue(ivqm_batch* b, const ivqm_pair* pairs, int32_t n, uint32_t metrics, int32_t ring) {
  return guarded([&]() -> int {
    if (!b) return set_error(IVQM_EINVAL, "NULL argument");
    std::vector<PairSpec> specs;
    RCCHK(batch_specs(pairs, n, specs));
    if (ring < 0 || ring >= kRing) return set_error(IVQM_EINVAL, "ring slot out of range");
    return b->lane(ring).enqueue(specs.data(), n, metrics & 0xFFu, ring, b->use_graph);
  });
}
int ivqm_batch_enqueue_masked(ivqm_batch* b, const ivqm_pair* pairs, const ivqm_mask* masks, int32_t n,
                              uint32_t metrics, int32_t ring) {
  return guarded([&]() -> int {
    if (!b || !masks) return set_error(IVQM_EINVAL, "NULL argument");
    std::vector<PairSpec> specs;
    RCCHK(batch_specs(pairs, n, specs));
    for (int i = 0; i < n; i++) {
      if (!masks[i].plane) continue;
      RCCHK(validate_mask(pairs[i].width, pairs[i].height, &masks[i]));
      mask_into_spec(masks[i], specs[i]);
    }
    if (ring < 0 || ring >= kRing) return set_error(IVQM_EINVAL, "ring slot out of range");
    return b->lane(ring).enqueue(specs.data(), n, metrics & 0xFFu, ring, b->use_graph);
  });
}
int ivqm_batch_collect(ivqm_batch* b, int32_t ring, ivqm_frame_result* results) {
  return guarded([&]() -> int {
    if (!b || !results) return set_error(IVQM_EINVAL, "NULL argument");
    if (ring < 0 || ring >= kRing) return set_error(IVQM_EINVAL, "ring slot out of range");
    return b->lane(ring).collect(ring, results);
  });
}
int ivqm_batch_run(ivqm_batch* b, const ivqm_pair* pairs, int32_t n, uint32_t metrics, ivqm_frame_result* results) {
  return guarded([&]() -> int {
    RCCHK(ivqm_batch_enqueue(b, pairs, n, metrics, 0));
    return ivqm_batch_collect(b, 0, results);
  });
}
int ivqm_batch_profile(ivqm_batch* b, int32_t kernel, int32_t iters, double* avg_ms) {
  return guarded([&]() -> int {
    if (!b || !avg_ms || iters <= 0) return set_error(IVQM_EINVAL, "bad argument");
    return b->eng->profile(kernel, iters, avg_ms);
  });
}

// Times every search tile height and SSIM band height on this batch and keeps
// the fastest of each (the two kernels do not share a shape).
int ivqm_batch_autotune(ivqm_batch* b, const ivqm_pair* pairs, int32_t n, uint32_t metrics, int32_t* search_tile_rows,
                        int32_t* ssim_band_rows) {
  if (!b) return set_error(IVQM_EINVAL, "NULL argument");
  std::vector<PairSpec> specs;
  RCCHK(batch_specs(pairs, n, specs));
  double best_j = 1e30;
  int py_best = b->eng->tuning_py(), nb_best = b->eng->tuning_nb();
  // the keyed 4:2:0 search's tile rows (the other search variants run 64 x 16 tiles)
  const int ths[2] = {16, 32}, nbs[3] = {4, 8, 16};
  // A warm-up trial, then each tile height twice in alternating order, keeping each one's best: the
  // first trials after an idle device run at lower clocks and would favour whichever height goes last.
  double tth[2] = {1e30, 1e30};
  bool search_tuned = true;
  for (int it = 0; it < 5; it++) {
    const int pi = it == 0 ? 0 : (it - 1) & 1;
    b->eng->set_tuning(ths[pi], nb_best);
    RCCHK(b->eng->enqueue(specs.data(), n, metrics & 0xFFu, 0, false));
    RCCHK(b->eng->collect(0, nullptr));
    if (!b->eng->last_plan().need_search || !b->eng->last_plan().tiles420) { search_tuned = false; break; }
    double t = 0;
    RCCHK(b->eng->profile(1, 5, &t));
    if (it > 0) tth[pi] = std::min(tth[pi], t);
  }
  if (search_tuned) py_best = tth[1] < tth[0] ? ths[1] : ths[0];
  for (int ni = 0; ni < 3; ni++) {
    b->eng->set_tuning(py_best, nbs[ni]);
    RCCHK(b->eng->enqueue(specs.data(), n, metrics & 0xFFu, 0, false));
    RCCHK(b->eng->collect(0, nullptr));
    if (b->eng->last_plan().jobs.empty()) break;
    double t = 0;
    RCCHK(b->eng->profile(2, 5, &t));
    if (t < best_j) { best_j = t; nb_best = nbs[ni]; }
  }
  b->eng->set_tuning(py_best, nb_best);
  b->eng2->set_tuning(py_best, nb_best);
  if (search_tile_rows) *search_tile_rows = py_best;
  if (ssim_band_rows) *ssim_band_rows = nb_best;
  return IVQM_OK;
}

}  // extern "C"
