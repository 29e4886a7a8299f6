// Batch engine of the gfx950 metric hot path (host side).
//
// An Engine owns one HIP stream and the device buffers of a batch of frame
// pairs. enqueue() plans a batch (offsets, block counts, job tables), writes
// the tables into a pinned staging area of the chosen ring slot and enqueues,
// in stream order: table upload -> stats -> GCD -> search (both directions,
// fused SCP) -> pyramid levels -> SSIM jobs -> FP64 totals -> one copy per
// result region into the slot's pinned arena. collect() waits for the slot
// and runs the host FP64 epilogue (ivqm_epilogue.h) per pair. Every phase is
// one launch for the whole batch; a batch of one pair is the plain API.
// Optionally the enqueued sequence of a slot is captured once as a hipGraph
// and replayed while the batch shape stays the same.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <map>
#include <string>
#include <vector>

#include "../../include/ivqm_hip.h"
#include "ivqm_device.h"

namespace ivqm {

int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what);

constexpr int kRing = 8;
constexpr int kPitchAlign = 64;  // samples
constexpr uint32_t kReuseScp = 1u << 31;  // IV-SSIM metrics use the stored SCP pictures
constexpr uint32_t kScpOnly = 1u << 30;   // run the search for the SCP pictures only

struct PairSpec {
  PicDev p0, p1;
  int w = 0, h = 0, bd = 8;
  bool has_gcd = false;
  int32_t gcd[4] = {0, 0, 0, 0};
  uint32_t metrics = 0;  // this pair's IVQM_M_* set; 0 = the batch's set
  // mask mode: weight plane at luma resolution (null = no mask)
  const void* msk = nullptr;
  int msk_stride = 0, msk_bps = 0, msk_bd = 0;
};

// Per-pair bookkeeping the epilogue needs.
struct PairPlan {
  uint32_t m;  // metrics of this pair
  int w, h, bd, cs, csy;  // chroma column / row shifts
  int64_t rows_off, rowsd_off, scp_off;
  int ns;
  int ssim = -1, iv_t2r = -1, iv_r2t = -1;
  int ms[5] = {-1, -1, -1, -1, -1}, ivms_t2r[5] = {-1, -1, -1, -1, -1}, ivms_r2t[5] = {-1, -1, -1, -1, -1};
  int erp_h = -1;  // height whose ERP weights apply
  bool masked = false;
  int bdm = 0;            // mask bit depth
  int64_t rows64_off = 0;
  bool fused = false;     // IV-SSIM from the search's sub-block moments (no SCP pictures)
};

// A picture of a plan: a caller's picture, or three planes in one of the
// engine's arenas (turned into device pointers, with a bounds check, when the
// batch is issued).
struct PicRef {
  int arena = 0;        // 0: caller picture, 1: SCP arena, 2: pyramid arena, 3: none (fused moments)
  int64_t off = 0;      // element offset of plane 0 in the arena
  int pitch = 0, w = 0, h = 0;
  PicDev user;
  // Pyramid level 1 of a 4:2:0 uint16 caller picture: the 2x2 mean of replicated chroma is the chroma
  // plane itself, so planes 1 and 2 are the caller's chroma planes (no copy).
  bool chroma_from_user = false;
};

// Device planes of a picture reference: the caller's picture, nothing (fused moments), or three planes
// of pitch x h elements from element offset `off` of the SCP or pyramid arena. Every table entry that
// names an arena picture goes through this check on the host before any launch reads the tables: an
// arena picture that would leave its allocation is rejected with IVQM_EINVAL (*why says which rule).
inline int resolve_picture(const PicRef& r, uint16_t* scp, int64_t cap_scp, uint16_t* pyr, int64_t cap_pyr,
                           PicDev& out, const char** why) {
  if (r.arena == 0) {
    out = r.user;
    return IVQM_OK;
  }
  if (r.arena == 3) {
    out = PicDev{};
    return IVQM_OK;
  }
  if (r.arena != 1 && r.arena != 2) {
    *why = "internal: unknown picture arena";
    return IVQM_EINVAL;
  }
  uint16_t* base = r.arena == 1 ? scp : pyr;
  const int64_t cap = r.arena == 1 ? cap_scp : cap_pyr;
  const int64_t plane = (int64_t)r.pitch * r.h;
  if (!base || r.off < 0 || r.w <= 0 || r.h <= 0 || r.w > r.pitch || r.off + 3 * plane > cap) {
    *why = "internal: arena picture outside its allocation";
    return IVQM_EINVAL;
  }
  for (int k = 0; k < 3; k++) {
    out.p[k] = base + r.off + (int64_t)k * plane;
    out.stride[k] = r.pitch;
  }
  out.cs = 0;
  out.csy = 0;
  out.bps = 2;
  if (r.chroma_from_user) {  // level 1 of a 4:2:0 uint16 picture: its chroma planes at (w >> 1, h >> 1)
    if (r.arena != 2 || chroma_format(r.user) != 1 || r.user.bps != 2) {
      *why = "internal: pyramid chroma source";
      return IVQM_EINVAL;
    }
    for (int k = 1; k < 3; k++) {
      out.p[k] = r.user.p[k];
      out.stride[k] = r.user.stride[k];
    }
  }
  return IVQM_OK;
}

struct BatchPlan {
  uint32_t metrics = 0;
  bool need_stats = false, need_search = false, need_scp = false;
  int n = 0;
  std::vector<PairPlan> pp;
  std::vector<PairDev> pairs;
  std::vector<SsimJobDev> jobs;      // a / b filled from job_a / job_b when issued
  std::vector<PicRef> job_a, job_b;
  std::vector<DownJob> down[5];      // src / dst filled from down_src / down_dst when issued
  std::vector<PicRef> down_src[5], down_dst[5];
  int stat_blocks = 0, tiles = 0, ssim_blocks = 0, down_blocks[5] = {0, 0, 0, 0, 0};
  bool ssim_gen = false;  // SSIM jobs run k_ssim_gen (any configuration but block-averaged 8 / 4)
  bool any_mask = false;
  int mtiles = 0;                              // k_search_masked tiles
  uint32_t mforms = 0;                         // masked search forms present (bit 0 reference, 1 narrow, 2 dot,
                                               // 3 4:2:0)
  uint32_t sforms = 0;                         // k_search kernels needed (search_form_bit of each pair's mode)
  int64_t aux_total = 0, rows64_total = 0;     // masked pairs' aux rows and uint64 search rows
  int64_t rows_total = 0, rowsd_total = 0, scp_total = 0, pyr_total = 0, partial_total = 0;
  int th420 = 32, nb = 8;  // k_search420 tile rows, SSIM band rows
  int tiles420 = 0;         // k_search420 tiles
  std::string signature;
};

struct RingSlot {
  // pinned host staging (tables) and results
  PairDev* t_pairs = nullptr;
  SsimJobDev* t_jobs = nullptr;
  DownJob* t_down = nullptr;
  int32_t* r_gcd = nullptr;
  int64_t* r_sd = nullptr;
  uint64_t* r_ssd = nullptr;
  uint64_t* r_row_ssd = nullptr;
  uint32_t* r_rows = nullptr;
  double* r_ssim = nullptr;
  int64_t* r_mcount = nullptr;   // per pair: non-zero mask samples, differing samples of Y, Cb, Cr
  uint64_t* r_rows64 = nullptr;  // masked search rows
  int64_t c_pairs = 0, c_rows = 0, c_rowsd = 0, c_jobs = 0, c_down = 0, c_rows64 = 0;  // capacities of the above
  hipEvent_t ev = nullptr;
  bool busy = false;
  BatchPlan plan;
  hipGraphExec_t graph = nullptr;
  std::string graph_sig;
};

// Search variant of a pair (PairDev::mode) from its sample maximum and the largest |GCD| it can see:
// the keyed forms where the host-checked ranges allow, the reference-form arithmetic otherwise.
int pick_search_mode(const ivqm_config& cfg, int maxv, int32_t gmax);
// Whether a 4:2:0 pair with sample maximum maxv and largest |GCD| gmax can run the keyed 4:2:0 search
// (k_search420): its int16 operands and int32 keys stay exact.
bool search420_ok(const ivqm_config& cfg, int maxv, int32_t gmax);

// Checks the SSIM mode / window / stride / border of a configuration
// (the rules of xAppQMIV.cpp:325-332 and xStructSimSTD.cpp:353-419).
int validate_ssim_config(const ivqm_config& cfg);
// Window weights of a Gaussian SSIM mode, row-major wnd x wnd (empty for averaged modes).
std::vector<double> ssim_window_weights(int mode, int wnd);

class Engine {
 public:
  explicit Engine(const ivqm_config& cfg);
  ~Engine();
  int init();
  int set_stream(hipStream_t s);
  hipStream_t stream() const { return stream_; }
  const ivqm_config& cfg() const { return cfg_; }

  // Plans and enqueues one batch into ring slot `ring`.
  int enqueue(const PairSpec* pairs, int n, uint32_t metrics, int ring, bool use_graph = false);
  // Waits for ring slot `ring` and writes one result per pair.
  int collect(int ring, ivqm_frame_result* out);
  // 1 when ring slot `ring` has finished on the device (or holds no batch), 0 if not, < 0 on error.
  int query(int ring);
  int sync();

  void set_tuning(int th, int nb) {
    th420_ = search420_tile_rows(th, th420_);
    if (nb == 4 || nb == 8 || nb == 16) nb_ = nb;
  }
  int tuning_py() const { return th420_; }
  int tuning_nb() const { return nb_; }
  void enable_best(bool on) { best_on_ = on; }
  // Times `iters` launches of one phase of the last enqueued batch (HIP events).
  int profile(int kernel, int iters, double* avg_ms);

  // device buffers (last batch) for debug readers
  uint32_t* d_rows() const { return d_rows_; }
  uint16_t* d_scp() const { return d_scp_; }
  uint8_t* d_best() const { return d_best_; }
  const BatchPlan& last_plan() const { return last_; }
  bool scp_valid() const { return scp_valid_; }
  void set_scp_valid(bool v) { scp_valid_ = v; }
  int ensure_scp(int64_t elems);
  static int scp_pitch(int w) { return (w + kPitchAlign - 1) / kPitchAlign * kPitchAlign; }

 private:
  int plan(const PairSpec* pairs, int n, uint32_t metrics, BatchPlan& P);
  int ensure_capacity(const BatchPlan& P);
  int ensure_slot_capacity(RingSlot& S, const BatchPlan& P);
  int issue(const BatchPlan& P, RingSlot& S);
  hipError_t launch_search_all(const BatchPlan& P, hipStream_t st);
  int fill_tables(const BatchPlan& P, RingSlot& S, int* ndown, int down_first[5]);
  int materialize(const PicRef& r, PicDev& out) const;
  void drop_graphs();
  int erp_offset(int h, int64_t* off);
  int masked_epilogue(const RingSlot& S, const PairPlan& pp, int i, ivqm_frame_result& r);
  int search_mode(int maxv, int32_t gcd_abs_max) const;
  int32_t masked_key_base(int maxv, int32_t gcd_abs_max, int tile_h) const;  // tile_h: the form's tile height
  bool masked_dot_ok(int maxv, int32_t gcd_abs_max) const;

  ivqm_config cfg_;
  hipStream_t stream_ = nullptr;
  bool own_stream_ = false;
  int th420_ = 32, nb_ = 8;
  bool best_on_ = false;
  bool scp_valid_ = false;

  // capacities (elements)
  int64_t cap_pairs_ = 0, cap_rows_ = 0, cap_rowsd_ = 0, cap_scp_ = 0, cap_pyr_ = 0, cap_partial_ = 0,
          cap_jobs_ = 0, cap_down_ = 0, cap_best_ = 0;

  PairDev* d_pairs_ = nullptr;
  SsimJobDev* d_jobs_ = nullptr;
  DownJob* d_down_ = nullptr;
  int64_t* d_row_sd_ = nullptr;
  uint64_t* d_row_ssd_ = nullptr;
  uint32_t* d_rows_ = nullptr;
  int32_t* d_gcd_ = nullptr;
  int64_t* d_sd_ = nullptr;
  uint64_t* d_ssd_ = nullptr;
  uint16_t* d_scp_ = nullptr;
  uint8_t* d_best_ = nullptr;
  uint16_t* d_pyr_ = nullptr;
  double* d_partial_ = nullptr;
  double* d_ssim_out_ = nullptr;
  double* d_erp_ = nullptr;
  double* d_wts_ = nullptr;  // SSIM window weights of the configured mode (Gaussian modes)
  int64_t* d_aux_ = nullptr;      // masked stats: differing / non-zero-mask samples per row
  int64_t* d_mcount_ = nullptr;   // 4 per pair
  uint64_t* d_rows64_ = nullptr;  // masked search rows
  int64_t cap_aux_ = 0, cap_rows64_ = 0;
  int64_t erp_used_ = 0, erp_cap_ = 0;
  std::map<int, int64_t> erp_off_;

 public:
  std::map<int, std::pair<std::vector<double>, double>> erp_host;  // height -> weights, correction
  // Cosine-latitude row weights and correction of height h (computed once).
  const std::pair<std::vector<double>, double>& erp_weights(int h);
 private:
  RingSlot ring_[kRing];
  BatchPlan last_;
};

}  // namespace ivqm
