// Batch engine of the gfx950 metric hot path: planning, buffers, launches,
// result ring and the host FP64 epilogue. See ivqm_engine.h.
#include "ivqm_engine.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "ivqm_epilogue.h"

namespace ivqm {

namespace {
thread_local std::string g_last_error;

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

template <typename T> int dmalloc(T** p, int64_t n) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (n <= 0) return IVQM_OK;
  hipError_t e = hipMalloc((void**)p, sizeof(T) * (size_t)n);
  if (e != hipSuccess) return hip_error(e, "hipMalloc");
  return IVQM_OK;
}
template <typename T> int hmalloc(T** p, int64_t n) {
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  if (n <= 0) n = 1;
  hipError_t e = hipHostMalloc((void**)p, sizeof(T) * (size_t)n, hipHostMallocDefault);
  if (e != hipSuccess) return hip_error(e, "hipHostMalloc");
  return IVQM_OK;
}
int64_t grow(int64_t need, int64_t have) { return need <= have ? have : std::max<int64_t>(need, have + have / 2); }

// Normalised n x n Gaussian (centre (n - 1) / 2, sum 1) in FP64.
std::vector<double> gaussian(int n, double sigma) {
  std::vector<double> g((size_t)n * n);
  const double c = 0.5 * (n - 1);
  double sum = 0.0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      const double v = std::exp(-((i - c) * (i - c) + (j - c) * (j - c)) / (2.0 * sigma * sigma));
      g[(size_t)i * n + j] = v;
      sum += v;
    }
  for (double& v : g) v /= sum;
  return g;
}

}  // namespace

// libs/pmbb-ivqm/src/xStructSimConsts.h: the float weights (:51-66) are the
// 11 x 11 Gaussian of sigma 1.5 (c_Sigma, :35) rounded to float32; the Q18
// tables (c_FltrIntShift = 18, :95-98) are round(2^18 x Gaussian) with one
// added at a few positions -- 11 x 11 sigma 1.5 (:100-113), 8 x 8 sigma 1.5
// (:134-144), 16 x 16 sigma 2.4 (:146-164).
std::vector<double> ssim_window_weights(int mode, int wnd) {
  if (mode == 0) {
    std::vector<double> g = gaussian(11, 1.5);
    for (double& v : g) v = (double)(float)v;
    return g;
  }
  struct Q18 {
    int n;
    double sigma;
    std::vector<int> plus_one;  // (row, column) pairs
  };
  static const Q18 k11 = {11, 1.5, {3, 4, 3, 6, 4, 3, 4, 7, 5, 4, 5, 6, 6, 3, 6, 7, 7, 4, 7, 6}};
  static const Q18 k8 = {8, 1.5, {2, 2, 2, 5, 5, 2, 5, 5}};
  static const Q18 k16 = {16, 2.4, {0, 3, 0, 12, 3, 0, 3, 15, 12, 0, 12, 15, 15, 3, 15, 12}};
  const Q18* q = mode == 1 ? &k11 : (mode == 3 && wnd == 8) ? &k8 : (mode == 3 && wnd == 16) ? &k16 : nullptr;
  if (!q) return {};
  std::vector<double> g = gaussian(q->n, q->sigma);
  for (double& v : g) v = std::floor(v * 262144.0 + 0.5);
  for (size_t k = 0; k + 1 < q->plus_one.size(); k += 2) g[(size_t)q->plus_one[k] * q->n + q->plus_one[k + 1]] += 1.0;
  return g;
}

int validate_ssim_config(const ivqm_config& cfg) {
  const int mode = cfg.ssim_mode, wnd = cfg.ssim_window, s = cfg.ssim_stride;
  if (mode < 0 || mode > 4) return set_error(IVQM_EINVAL, "ssim_mode must be 0..4 (xSSIM::eMode)");
  if (cfg.ssim_border < 0 || cfg.ssim_border > 5) return set_error(IVQM_EINVAL, "ssim_border must be 0..5 (eMrgExt)");
  const int ew = mode <= 2 ? 11 : wnd;  // regular windows are 11 x 11 whatever ssim_window says (xSSIM.h:95)
  if (mode == 3 && wnd != 8 && wnd != 16)
    return set_error(IVQM_EINVAL, "BlockGaussianInt supports windows 8 and 16 (xStructSimSTD.cpp:353-419)");
  if (mode == 4 && (wnd < 1 || wnd > 64)) return set_error(IVQM_EINVAL, "BlockAveraged window must be 1..64");
  if (s < 1 || s > ew) return set_error(IVQM_EINVAL, "ssim_stride must be in range 1..window (xAppQMIV.cpp:331)");
  return IVQM_OK;
}

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
int hip_error(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return IVQM_EHIP - (int)e;
}
const char* last_error_message() { return g_last_error.c_str(); }

#define HIPCHK(expr)                                   \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_error(_e, #expr); \
  } while (0)
#define RCCHK(expr)          \
  do {                       \
    int _rc = (expr);        \
    if (_rc) return _rc;     \
  } while (0)

Engine::Engine(const ivqm_config& cfg) : cfg_(cfg) { th420_ = search420_tile_rows(cfg.search_tile_rows, th420_); }

Engine::~Engine() {
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (auto& S : ring_) {
    if (S.graph) (void)hipGraphExecDestroy(S.graph);
    if (S.ev) (void)hipEventDestroy(S.ev);
    void* hs[] = {S.t_pairs, S.t_jobs, S.t_down, S.r_gcd, S.r_sd, S.r_ssd, S.r_row_ssd, S.r_rows, S.r_ssim, S.r_mcount,
                  S.r_rows64};
    for (void* h : hs)
      if (h) (void)hipHostFree(h);
  }
  void* ds[] = {d_pairs_, d_jobs_, d_down_, d_row_sd_, d_row_ssd_, d_rows_, d_gcd_, d_sd_, d_ssd_,
                d_scp_,   d_best_, d_pyr_,  d_partial_, d_ssim_out_, d_erp_, d_wts_, d_aux_, d_mcount_, d_rows64_};
  for (void* d : ds)
    if (d) (void)hipFree(d);
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
}

int Engine::init() {
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) return set_error(IVQM_EHIP - (int)(e ? e : hipErrorNoDevice), "no HIP device");
  HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  own_stream_ = true;
  for (auto& S : ring_) HIPCHK(hipEventCreateWithFlags(&S.ev, hipEventDisableTiming));
  std::vector<double> wts = ssim_window_weights(cfg_.ssim_mode, cfg_.ssim_window);
  if (!wts.empty()) {  // row-major table, then its transpose (column-outer kernels)
    const int n = (int)std::lround(std::sqrt((double)wts.size()));
    wts.resize(2 * (size_t)n * n);
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) wts[(size_t)n * n + (size_t)j * n + i] = wts[(size_t)i * n + j];
    RCCHK(dmalloc(&d_wts_, (int64_t)wts.size()));
    HIPCHK(hipMemcpy(d_wts_, wts.data(), sizeof(double) * wts.size(), hipMemcpyHostToDevice));
  }
  return IVQM_OK;
}

void Engine::drop_graphs() {
  for (auto& S : ring_)
    if (S.graph) { (void)hipGraphExecDestroy(S.graph); S.graph = nullptr; S.graph_sig.clear(); }
}

int Engine::set_stream(hipStream_t s) {
  RCCHK(sync());
  drop_graphs();
  if (own_stream_ && stream_) (void)hipStreamDestroy(stream_);
  own_stream_ = false;
  stream_ = s;
  if (!s) {
    HIPCHK(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    own_stream_ = true;
  }
  return IVQM_OK;
}

int Engine::sync() {
  HIPCHK(hipStreamSynchronize(stream_));
  return IVQM_OK;
}

int Engine::search_mode(int maxv, int32_t gmax) const { return pick_search_mode(cfg_, maxv, gmax); }

int pick_search_mode(const ivqm_config& cfg, int maxv, int32_t gmax) {
  if (cfg.search_mode_cap == 1) return 0;  // reference-form arithmetic only
  const int32_t* w = cfg.cmp_weights_search;
  const int64_t A = (int64_t)maxv + gmax;
  int64_t wabs = 0;
  for (int k = 0; k < 3; k++) wabs = std::max<int64_t>(wabs, std::llabs((long long)w[k]));
  // (d*w)*d by 24-bit multiplies: |d| <= A and |d * w| < 2^23; with the default weights {4, 1, 1}
  // the luma term is (2d)^2 (mode 5, kSearchWide24D in the kernels): |2d| < 2^23. The expanded form
  // T + Q(r) - 2 sum w t r mod 2^32 (mode 6, kSearchQuad) takes the multipliers -2 w t as 24-bit
  // operands: 2 |w| A < 2^23.
  const bool w411 = w[0] == 4 && w[1] == 1 && w[2] == 1;
  const int wide = (2 * A * wabs < (1ll << 23)) ? 6 : (A * wabs < (1ll << 23)) ? (w411 ? 5 : 3) : 0;
  for (int k = 0; k < 3; k++)
    if (w[k] < 0) return wide;
  const int64_t S = 1 << key_bits(cfg.search_range);
  const int64_t wsum = (int64_t)w[0] + w[1] + w[2];
  const int64_t wmax = std::max<int64_t>(w[0], std::max<int64_t>(w[1], w[2]));
  // the keyed variants' epilogue takes t - r as int16
  if (A >= 32768) return wide;
  // key = S * (Q(r) - 2 sum w t r) + idx stays inside int32 and the 24-bit multiplier operands
  if (3 * S * wsum * A * A >= (1ll << 31)) return wide;
  if (2 * S * wmax * A >= (1ll << 23)) return wide;
  // packed variant: S * r and 2 * w * t are int16 operands
  if (cfg.search_mode_cap != 2 && S * maxv <= 32767 && 2 * wmax * A <= 32767) return 2;
  return 1;
}

// Keyed masked search (ivqm_masked.hip): key = S * (Q(r) - 2 sum_c w_c t_c r_c) + offset, a masked-out
// candidate's entry holds the constant B instead of S * Q(r) + offset. With |t| <= A = Max + |gcd|,
// r <= Max and T = 2 S wsum A Max bounding the per-pixel term, B = 2^31 - 1 - T never wraps and exceeds
// every considered key (<= S wsum Max^2 + T + offset) when 2^31 - 1 - 2 T > that bound. Returns B, or 0
// for the reference-form masked search.
int32_t Engine::masked_key_base(int maxv, int32_t gmax, int tile_h) const {
  if (cfg_.search_mode_cap == 1) return 0;
  const int32_t* w = cfg_.cmp_weights_search;
  for (int k = 0; k < 3; k++)
    if (w[k] < 0) return 0;
  const int R = cfg_.search_range, D = 2 * R + 1;
  const int64_t S = 1 << key_bits(R);
  const int64_t A = (int64_t)maxv + gmax;
  const int64_t wsum = (int64_t)w[0] + w[1] + w[2];
  const int64_t wmax = std::max<int64_t>(w[0], std::max<int64_t>(w[1], w[2]));
  if (A >= 32768 || 2 * S * wmax * A >= (1ll << 23) || wmax * maxv >= (1ll << 24)) return 0;
  const int64_t off_max = (int64_t)(tile_h + 2 * R) * D + kMaskTileW + 2 * R;
  const int64_t T = 2 * S * wsum * A * maxv;
  const int64_t vmax = S * wsum * (int64_t)maxv * maxv + T + off_max;
  const int64_t B = (1ll << 31) - 1 - T;
  if (B - T <= vmax) return 0;
  return (int32_t)B;
}

// The packed form of the keyed masked search: S * r and 2 * w * t are int16 operands of v_dot2
// (as the unmasked search's dot variant); search_mode_cap 2 forbids it.
// k_search420 operands (C0 = Max + max|gcd| <= A): samples below 2^12, weights below 128 and 2 C0 below
// 2^14 (the kernel's 24-bit multiplies), chroma samples times Sy and the chroma multipliers
// 2 F w (C0 - t) <= 2 F w 2A (F = Sk / Sy) are u16, the luma multiplier 2 Sk wY (C0 - t) takes a u16 scale
// 2 Sk wY, and the keys Sk * E' + offset stay inside int32 (|E'| <= 3 wsum A^2 bounds Q(r) + 2 sum w |t| r).
bool search420_ok(const ivqm_config& cfg, int maxv, int32_t gmax) {
  const int32_t* w = cfg.cmp_weights_search;
  for (int k = 0; k < 3; k++)
    if (w[k] < 0) return false;
  const int R = cfg.search_range;
  const int64_t Sk = 1ll << search420_key_bits(R), Sy = 1ll << search420_sample_bits(R), F = Sk / Sy;
  const int64_t A = (int64_t)maxv + gmax;
  const int64_t wsum = (int64_t)w[0] + w[1] + w[2];
  const int64_t wmax = std::max<int64_t>(w[0], std::max<int64_t>(w[1], w[2]));
  if (maxv >= 4096 || A >= 8192 || wmax >= 128) return false;  // the kernel's operand masks
  if (Sy * maxv > 65535 || 2 * F * wmax * 2 * A > 65535 || 2 * Sk * wmax > 65535) return false;
  if (2 * Sk * wmax * 2 * A >= (1ll << 24)) return false;  // luma multiplier: an operand of v_mad_u32_u24
  return 3 * Sk * wsum * A * A + (1ll << 16) < (1ll << 31);
}

bool Engine::masked_dot_ok(int maxv, int32_t gmax) const {
  if (cfg_.search_mode_cap == 2) return false;
  const int32_t* w = cfg_.cmp_weights_search;
  const int64_t S = 1 << key_bits(cfg_.search_range);
  const int64_t A = (int64_t)maxv + gmax;
  const int64_t wmax = std::max<int64_t>(w[0], std::max<int64_t>(w[1], w[2]));
  return S * maxv <= 32767 && 2 * wmax * A <= 32767;
}

const std::pair<std::vector<double>, double>& Engine::erp_weights(int h) {
  auto it = erp_host.find(h);
  if (it != erp_host.end()) return it->second;
  auto& hw = erp_host[h];
  host::ws_weights(h, cfg_.lat_range_deg, hw.first, hw.second);
  return hw;
}

int Engine::erp_offset(int h, int64_t* off) {
  auto it = erp_off_.find(h);
  if (it != erp_off_.end()) { *off = it->second; return IVQM_OK; }
  const auto& hw = erp_weights(h);
  if (erp_used_ + h > erp_cap_) {
    // grow: keep existing weights (copy old contents)
    const int64_t nc = grow(erp_used_ + h, std::max<int64_t>(erp_cap_, 4096));
    double* nd = nullptr;
    HIPCHK(hipMalloc((void**)&nd, sizeof(double) * nc));
    if (d_erp_) {
      HIPCHK(hipStreamSynchronize(stream_));
      HIPCHK(hipMemcpy(nd, d_erp_, sizeof(double) * erp_used_, hipMemcpyDeviceToDevice));
      (void)hipFree(d_erp_);
    }
    d_erp_ = nd;
    drop_graphs();  // captured SSIM totals read the old weights
    erp_cap_ = nc;
  }
  HIPCHK(hipStreamSynchronize(stream_));
  HIPCHK(hipMemcpy(d_erp_ + erp_used_, hw.first.data(), sizeof(double) * h, hipMemcpyHostToDevice));
  erp_off_[h] = erp_used_;
  *off = erp_used_;
  erp_used_ += h;
  return IVQM_OK;
}

// Which k_stats launches a plan needs (stats_form_bit over its pairs).
static uint32_t stats_forms(const BatchPlan& P) {
  uint32_t f = 0;
  for (int i = 0; i < P.n; i++) f |= stats_form_bit(P.pairs[i]);
  return f;
}

// Which k_ssim launches a plan needs: bit 0 jobs that read pictures, bit 1 jobs that read fused moments.
static uint32_t ssim_forms(const BatchPlan& P) {
  uint32_t f = 0;
  // jobs own consecutive block ranges in table order; twins own none and launch nothing
  const size_t nj = P.jobs.size();
  for (size_t j = 0; j < nj; j++) {
    const SsimJobDev& J = P.jobs[j];
    const int nb = (j + 1 < nj ? P.jobs[j + 1].blk0 : P.ssim_blocks) - J.blk0;
    if (nb > 0) f |= J.mom_off >= 0 ? 2u : (J.wsq == 0 ? 1u : 4u);
  }
  return f;
}

int Engine::plan(const PairSpec* pairs, int n, uint32_t metrics, BatchPlan& P) {
  const bool reuse = (metrics & kReuseScp) != 0, scp_only = (metrics & kScpOnly) != 0;
  uint32_t m = metrics & 0xFFu;
  for (int i = 0; i < n; i++) m |= pairs[i].metrics & 0xFFu;  // the batch covers every pair's set
  const uint32_t ivs = IVQM_M_IVSSIM | IVQM_M_IVMSSSIM;
  P = BatchPlan();
  P.metrics = m;
  P.n = n;
  P.th420 = th420_;
  P.nb = nb_;
  P.need_search = (m & IVQM_M_IVPSNR) || (!reuse && (m & ivs)) || scp_only;
  P.need_scp = (!reuse && (m & ivs)) || scp_only;
  P.need_stats = (m & (IVQM_M_PSNR | IVQM_M_MSE | IVQM_M_WSPSNR)) || P.need_search;
  if ((m & ivs) && reuse && !(n == 1 && scp_valid_))
    return set_error(IVQM_ESTATE, "IV-SSIM needs shift-compensated pictures (ivqm_gen_scp / ivqm_upload_scp)");
  const int R = cfg_.search_range;
  const int TH = kSearchTileH;
  const int wnd = cfg_.ssim_window, str = cfg_.ssim_stride;
  P.pp.resize(n);
  P.pairs.resize(n);
  char sigbuf[96];
  std::snprintf(sigbuf, sizeof sigbuf, "m%x py%d nb%d n%d|", m | (reuse ? 0x100u : 0u) | (scp_only ? 0x200u : 0u),
                P.th420, P.nb, n);
  P.signature = sigbuf;
  const bool ms_any = (m & (IVQM_M_MSSSIM | IVQM_M_IVMSSSIM)) != 0;
  const bool regular = cfg_.ssim_mode <= 2;
  P.ssim_gen = !(cfg_.ssim_mode == 4 && wnd == 8 && str == 4);
  // Fused IV-SSIM (block-averaged 8 / 4): the search writes 4x4 sub-block moments
  // instead of the SCP pictures when nothing else reads them -- batches of two or
  // more pairs (a single pair keeps its SCP for ivqm_get_scp and the IV-SSIM
  // entry points that reuse it), 32-bit sums. With IV-MS-SSIM (k_search420 pairs)
  // the search also writes MS-SSIM level 1 of the SCP pictures: level 0 of
  // IV-MS-SSIM reads the IV-SSIM windows (a twin of the moment job), the pyramid
  // continues from level 1, and no full-size SCP picture is written.
  const bool fuse_ok = n > 1 && !reuse && !scp_only && !P.ssim_gen && !cfg_.ssim_wide_sums;
  for (int i = 0; i < n; i++) {
    const PairSpec& s = pairs[i];
    PairPlan& pp = P.pp[i];
    PairDev& d = P.pairs[i];
    std::memset(&d, 0, sizeof(d));
    const uint32_t mp = s.metrics ? (s.metrics & 0xFFu) : m;
    pp.m = mp;
    pp.w = s.w; pp.h = s.h; pp.bd = s.bd; pp.cs = s.p0.cs; pp.csy = s.p0.csy;
    d.p0 = s.p0;
    d.p1 = s.p1;
    d.w = s.w;
    d.h = s.h;
    d.maxv = (1 << s.bd) - 1;
    int32_t gmax = 0;
    for (int k = 0; k < 3; k++) {
      d.md[k] = host::round_f32(cfg_.unnoticeable_coef[k] * (float)d.maxv);  // xGlobClrDiff.cpp:29
      gmax = std::max<int32_t>(gmax, s.has_gcd ? std::abs(s.gcd[k]) : std::abs(d.md[k]));
    }
    d.mode = search_mode(d.maxv, gmax);
    d.wrap32 = (cfg_.gcd_sd_wrap32 && s.bd <= 14) ? 1 : 0;
    d.gcd_given = s.has_gcd ? 1 : 0;
    for (int k = 0; k < 4; k++) d.gcd_in[k] = s.has_gcd ? s.gcd[k] : 0;
    d.res_idx = i;
    if (s.msk) {  // mask mode (xAppQMIV.cpp:360-365: no structural-similarity metric with a mask)
      if (mp & (IVQM_M_SSIM | IVQM_M_MSSSIM | ivs))
        return set_error(IVQM_EINVAL, "structural similarity metrics cannot be combined with a mask");
      d.msk = s.msk;
      d.msk_stride = s.msk_stride;
      d.msk_bps = s.msk_bps;
      d.msk_max = (1 << s.msk_bd) - 1;
      d.aux_off = P.aux_total;
      P.aux_total += 4 * (int64_t)s.h;
      d.rows64_off = P.rows64_total;
      pp.rows64_off = d.rows64_off;
      P.rows64_total += 6 * (int64_t)s.h;
      pp.masked = true;
      pp.bdm = s.msk_bd;
      d.mkey = masked_key_base(d.maxv, gmax, kMaskTileH);
      d.mdot = (d.mkey != 0 && masked_dot_ok(d.maxv, gmax)) ? 1 : 0;
      // 4:2:0 pairs: the 4:2:0 form (the packed form's operand ranges, keys over its taller tiles)
      if (d.mdot && chroma_format(s.p0) == 1 && chroma_format(s.p1) == 1 && cfg_.search_mode_cap == 0 &&
          masked_key_base(d.maxv, gmax, kMask420TileH) != 0)
        d.mdot = 2;
      P.any_mask = true;
    }
    d.stat_blk0 = P.stat_blocks;
    P.stat_blocks += s.msk ? 3 * s.h : s.h + 2 * (s.h >> s.p0.csy);  // masked: every component at 4:4:4
    d.tile0 = P.tiles;  // pairs without a search own no tiles (the table lookup skips them)
    d.tile0b = P.tiles420;
    if (!s.msk && chroma_format(s.p0) == 1 && (d.mode == 1 || d.mode == 2) && cfg_.search_mode_cap == 0 &&
        search420_ok(cfg_, d.maxv, gmax))
      d.mode = kSearchDot420;
    d.mtile0 = P.mtiles;
    if (s.msk) {
      if ((mp & IVQM_M_IVPSNR) || scp_only) {  // the masked search also writes the SCP pictures
        const int th = d.mdot == 2 ? kMask420TileH : kMaskTileH;
        P.mtiles += 2 * ((s.w + kMaskTileW - 1) / kMaskTileW) * ((s.h + th - 1) / th);
        P.mforms |= 1u << (d.mkey == 0 ? 0 : (d.mdot == 2 ? 3 : (d.mdot ? 2 : 1)));
      }
    } else if ((mp & IVQM_M_IVPSNR) || (P.need_scp && ((mp & ivs) || scp_only))) {
      if (d.mode == kSearchDot420)
        P.tiles420 += ((s.w + kSearchTileW - 1) / kSearchTileW) * ((s.h + P.th420 - 1) / P.th420);
      else {
        P.tiles += ((s.w + kSearchTileW - 1) / kSearchTileW) * ((s.h + TH - 1) / TH);
        P.sforms |= search_form_bit(d.mode);
      }
    }
    d.rowsd_off = P.rowsd_total;
    pp.rowsd_off = d.rowsd_off;
    P.rowsd_total += 3 * (int64_t)s.h;
    d.rows_off = P.rows_total;
    pp.rows_off = d.rows_off;
    P.rows_total += 6 * (int64_t)s.h;
    d.scp_pitch = scp_pitch(s.w);
    // the kernels address a plane with 32-bit byte offsets (SCP, buffer resources)
    for (int c = 0; c < 3; c++) {
      const int64_t rows = c ? (s.h + s.p0.csy) >> s.p0.csy : s.h;
      if ((int64_t)s.p0.stride[c] * rows * s.p0.bps >= (1ll << 32) ||
          (int64_t)s.p1.stride[c] * rows * s.p1.bps >= (1ll << 32))
        return set_error(IVQM_EINVAL, "picture plane of 4 GiB or more");
    }
    if ((int64_t)d.scp_pitch * s.h * 2 >= (1ll << 32)) return set_error(IVQM_EINVAL, "picture plane of 4 GiB or more");
    // the keyed search variants (1 narrow, 2 dot) write the moments; k_search420 also the SCP level 1
    const bool fused = fuse_ok && (mp & IVQM_M_IVSSIM) &&
                       (!(mp & IVQM_M_IVMSSSIM) || (d.mode == kSearchDot420 && ms_num_scales(s.w, s.h) > 1)) &&
                       !s.msk &&
                       s.bd <= kMomMaxBits && (d.mode == 1 || d.mode == 2 || d.mode == kSearchDot420) &&
                       num_blocks(s.w, 8, 4) > 0 &&
                       num_blocks(s.h, 8, 4) > 0;
    d.mom_off = -1;
    d.scp1_off = -1;
    d.scp1_pitch = 0;
    pp.ns = ms_num_scales(s.w, s.h);
    if (fused && (mp & IVQM_M_IVMSSSIM) && pp.ns > 1) {
      d.scp1_pitch = scp_pitch(s.w >> 1);
      P.scp_total = (P.scp_total + 127) & ~(int64_t)127;
      d.scp1_off = P.scp_total;
      P.scp_total += 2 * 3 * (int64_t)d.scp1_pitch * (s.h >> 1);
    }
    if (fused) {
      // sub-blocks 0 .. num_blocks of each axis: those the 8 / 4 windows cover, all inside the picture
      d.scp_off = -1;
      d.mom_rows = num_blocks(s.h, 8, 4) + 1;
      d.mom_cols = num_blocks(s.w, 8, 4) + 1;
      d.mom_pitch = (d.mom_cols + 15) & ~15;  // a 64-column search tile stores 16 sub-blocks = 256 bytes
      P.scp_total = (P.scp_total + 127) & ~(int64_t)127;  // 256-byte aligned
      d.mom_off = P.scp_total;
      P.scp_total += 2 * 3 * (int64_t)d.mom_rows * d.mom_pitch * 8;  // 4 x uint32 = 8 uint16 per sub-block
    } else if ((P.need_scp && ((mp & ivs) || scp_only)) || (reuse && (mp & ivs))) {
      d.scp_off = P.scp_total;
      P.scp_total += 6 * (int64_t)d.scp_pitch * s.h;
    } else {
      d.scp_off = -1;
    }
    pp.scp_off = d.scp_off;
    pp.fused = fused;
    d.best_off = (best_on_ && n == 1) ? 0 : -1;
    if (cfg_.erp) pp.erp_h = s.h;
    std::snprintf(sigbuf, sizeof sigbuf, "%dx%d:%d:%d:%d:%x:%d|", s.w, s.h, s.bd, chroma_format(s.p0), s.p0.bps, mp,
                  s.msk ? s.msk_bps : 0);
    P.signature += sigbuf;
  }
  // pyramid offsets: per pair, pictures 0..3 (P0, P1, SCP0, SCP1), levels 1..ns-1
  std::vector<std::array<std::array<int64_t, 5>, 4>> pyr(n);
  for (int i = 0; i < n; i++) {
    PairPlan& pp = P.pp[i];
    if (!ms_any || pp.ns <= 1) continue;
    const uint32_t mp = pp.m;
    if (!(mp & (IVQM_M_MSSSIM | IVQM_M_IVMSSSIM))) continue;
    for (int which = 0; which < 4; which++) {
      if (which >= 2 && !(mp & IVQM_M_IVMSSSIM)) continue;
      for (int l = 1; l < pp.ns; l++) {
        if (which >= 2 && l == 1 && P.pairs[i].scp1_off >= 0) continue;  // written by the search
        pyr[i][which][l] = P.pyr_total;
        P.pyr_total += 3 * (int64_t)scp_pitch(pp.w >> l) * (pp.h >> l);
      }
    }
  }
  // Pictures of level l: caller pictures at level 0, SCP arena for SCP0/SCP1 at
  // level 0, pyramid arena above.
  auto ref_for = [&](int i, int which, int l) -> PicRef {
    const PairDev& d = P.pairs[i];
    PicRef r;
    r.w = d.w >> l;
    r.h = d.h >> l;
    if (l == 0 && which < 2) {
      r.arena = 0;
      r.user = which == 0 ? d.p0 : d.p1;
    } else if (l == 0) {
      r.arena = d.scp_off >= 0 ? 1 : 3;  // fused pairs have no full-size SCP (their jobs read moments)
      r.pitch = d.scp_pitch;
      r.off = d.scp_off + (int64_t)(which - 2) * 3 * d.scp_pitch * d.h;
    } else if (l == 1 && which >= 2 && d.scp1_off >= 0) {  // SCP level 1 from the search
      r.arena = 1;
      r.pitch = d.scp1_pitch;
      r.off = d.scp1_off + (int64_t)(which - 2) * 3 * d.scp1_pitch * (d.h >> 1);
    } else {
      r.arena = 2;
      r.pitch = scp_pitch(d.w >> l);
      r.off = pyr[i][which][l];
      const PicDev& u = which == 0 ? d.p0 : d.p1;
      // the SSIM kernel reads a 4:4:4 picture's rows as 4-sample vectors: the chroma planes need the
      // alignment the full-resolution planes have (8-byte base, stride a multiple of 4 samples)
      const bool aligned = ((reinterpret_cast<uintptr_t>(u.p[1]) | reinterpret_cast<uintptr_t>(u.p[2])) & 7) == 0 &&
                           u.stride[1] % 4 == 0 && u.stride[2] % 4 == 0;
      if (l == 1 && which < 2 && chroma_format(u) == 1 && u.bps == 2 && aligned) {
        r.chroma_from_user = true;
        r.user = u;
      }
    }
    return r;
  };
  // downsample jobs per level
  for (int i = 0; i < n; i++) {
    const PairPlan& pp = P.pp[i];
    if (!ms_any || pp.ns <= 1) continue;
    const uint32_t mp = pp.m;
    if (!(mp & (IVQM_M_MSSSIM | IVQM_M_IVMSSSIM))) continue;
    for (int which = 0; which < 4; which++) {
      if (which >= 2 && !(mp & IVQM_M_IVMSSSIM)) continue;
      for (int l = 1; l < pp.ns; l++) {
        if (which >= 2 && l == 1 && P.pairs[i].scp1_off >= 0) continue;  // written by the search
        DownJob dj;
        std::memset(&dj, 0, sizeof(dj));
        dj.w = pp.w >> (l - 1);
        dj.h = pp.h >> (l - 1);
        dj.pitch = scp_pitch(pp.w >> l);
        const int dw = dj.w >> 1, dh = dj.h >> 1;
        dj.blk0 = P.down_blocks[l];
        const PicRef dst = ref_for(i, which, l);
        dj.ncomp = dst.chroma_from_user ? 1 : 3;
        P.down_blocks[l] += dj.ncomp * ((dw + kDownBlkW - 1) / kDownBlkW) * ((dh + kDownBlkH - 1) / kDownBlkH);
        P.down[l].push_back(dj);
        P.down_src[l].push_back(ref_for(i, which, l - 1));
        P.down_dst[l].push_back(dst);
      }
    }
  }
  // SSIM jobs
  auto add_job = [&](int i, int* slot, const PicRef& a, const PicRef& b, int w, int h, int calc_l, int erp) -> int {
    SsimJobDev J;
    std::memset(&J, 0, sizeof(J));
    J.w = w;
    J.h = h;
    // window grid: xSSIM::xInitLoopRanges, libs/pmbb-ivqm/src/xSSIM.cpp:161-196
    J.mode = cfg_.ssim_mode;
    J.stride = str;
    if (regular) {
      const int b = cfg_.ssim_border ? 0 : 5;  // centres over the whole plane with an extension
      J.ext = cfg_.ssim_border;
      J.wnd = 11;
      J.gx0 = J.gy0 = b;
      J.nux = (w - 2 * b + str - 1) / str;
      J.nuy = (h - 2 * b + str - 1) / str;
      J.erp_base = b;  // ERP weight at the window's centre row (xIVSSIM.cpp:152)
    } else {
      J.wnd = wnd;
      J.nux = num_blocks(w, wnd, str);
      J.nuy = num_blocks(h, wnd, str);
      J.erp_base = wnd >> 1;  // ... at its top row + half the window
    }
    if (J.nux <= 0 || J.nuy <= 0)
      return set_error(IVQM_EINVAL, "picture (or MS-SSIM scale) smaller than the SSIM window");
    if (P.ssim_gen) {
      const SsimGenShape sh = ssim_gen_shape(cfg_.ssim_mode, cfg_.ssim_window, cfg_.ssim_border, str);
      J.nchunk = (J.nux + sh.cols - 1) / sh.cols;
      J.nbg = (J.nuy + 4 * sh.rows - 1) / (4 * sh.rows);
    } else {
      J.nchunk = (J.nux + 62) / 63;
      J.nbg = ((J.nuy + P.nb - 1) / P.nb + 3) / 4;
    }
    J.calc_l = calc_l;
    J.erp = erp;
    // moment sum widths: 32-bit up to 12-bit samples, 64-bit windows up to 14-bit, 64-bit throughout above
    J.wsq = (P.pp[i].bd > 14 || cfg_.ssim_wide_sums) ? 2 : (P.pp[i].bd > 12 ? 1 : 0);
    J.erp_off = -1;
    if (erp) RCCHK(erp_offset(h, &J.erp_off));
    J.blk0 = P.ssim_blocks;
    P.ssim_blocks += 3 * J.nchunk * J.nbg;
    J.twin_part_off = -1;
    J.grp = P.ssim_gen ? 1 : 4;
    J.part_off = P.partial_total;
    P.partial_total += 3 * (int64_t)((J.nuy + J.grp - 1) / J.grp) * J.nchunk;
    J.out_idx = (int)P.jobs.size();
    const double maxv = (double)((1 << P.pp[i].bd) - 1);
    const double c1 = (0.01 * maxv) * (0.01 * maxv), c2 = (0.03 * maxv) * (0.03 * maxv);  // xSSIM.cpp:50-52
    const double n2 = (double)(wnd * wnd) * (double)(wnd * wnd);
    J.c1n = (float)(c1 * n2);
    J.c2n = (float)(c2 * n2);
    J.c1 = c1;
    J.c2 = c2;
    // sums -> moments: xStructSimSTD.cpp:38-42 (float weights), :104-108 (Q18), :176-180 (1 / 121),
    // :446-450 (1 / window area)
    J.kscale = J.mode == 0 ? 1.0 : (J.mode == 1 || J.mode == 3) ? (double)1.0 / (double)262144
               : J.mode == 2 ? (double)1.0 / (double)121 : (double)1.0 / (double)(wnd * wnd);
    J.wt = d_wts_;
    J.mom_off = -1;
    *slot = J.out_idx;
    P.jobs.push_back(J);
    P.job_a.push_back(a);
    P.job_b.push_back(b);
    return IVQM_OK;
  };
  for (int i = 0; i < n; i++) {
    PairPlan& pp = P.pp[i];
    const uint32_t mp = pp.m;
    if (mp & IVQM_M_SSIM) RCCHK(add_job(i, &pp.ssim, ref_for(i, 0, 0), ref_for(i, 1, 0), pp.w, pp.h, 1, 0));
    if (mp & IVQM_M_IVSSIM) {
      RCCHK(add_job(i, &pp.iv_t2r, ref_for(i, 0, 0), ref_for(i, 3, 0), pp.w, pp.h, 1, cfg_.erp));
      RCCHK(add_job(i, &pp.iv_r2t, ref_for(i, 1, 0), ref_for(i, 2, 0), pp.w, pp.h, 1, cfg_.erp));
      if (pp.fused) {  // search direction 0 pairs P0 with SCP1 (T2R), direction 1 P1 with SCP0 (R2T)
        const PairDev& d = P.pairs[i];
        for (int dir = 0; dir < 2; dir++) {
          const int j = dir ? pp.iv_r2t : pp.iv_t2r;
          SsimJobDev& J = P.jobs[j];
          if (J.nux + 1 != d.mom_cols || J.nuy + 1 != d.mom_rows)
            return set_error(IVQM_EINVAL, "internal: fused moment grid mismatch");
          J.mom_off = d.mom_off + (int64_t)dir * 3 * d.mom_rows * d.mom_pitch * 8;
          J.mom_pitch = d.mom_pitch;
          J.mom_rows = d.mom_rows;
          P.job_b[j].arena = 3;
        }
      }
    }
    if (mp & (IVQM_M_MSSSIM | IVQM_M_IVMSSSIM)) {
      for (int l = 0; l < pp.ns; l++) {
        const int w = pp.w >> l, h = pp.h >> l, cl = (l == pp.ns - 1) ? 1 : 0;
        if (mp & IVQM_M_MSSSIM) {
          RCCHK(add_job(i, &pp.ms[l], ref_for(i, 0, l), ref_for(i, 1, l), w, h, cl, 0));
          // Level 0 of MS-SSIM reads the same (P0, P1) windows as SSIM: with the block-averaged 8 / 4
          // kernel the SSIM job also writes this job's contrast-structure partials, and this job
          // owns no blocks (a zero-block entry: the block lookup finds the next job).
          if (l == 0 && cl == 0 && pp.ssim >= 0 && !P.ssim_gen) {
            SsimJobDev& T = P.jobs[pp.ms[0]];
            SsimJobDev& S = P.jobs[pp.ssim];
            if (T.nux == S.nux && T.nuy == S.nuy && T.nchunk == S.nchunk && T.wsq == S.wsq && !T.erp &&
                T.blk0 + 3 * T.nchunk * T.nbg == P.ssim_blocks) {
              P.ssim_blocks = T.blk0;
              S.twin_part_off = T.part_off;
            }
          }
        }
        if (mp & IVQM_M_IVMSSSIM) {
          // Level 0 of IV-MS-SSIM reads the IV-SSIM jobs' windows, (P0, SCP1) and (P1, SCP0): twins of
          // those jobs as above (a twin has no ERP weights: the IV-SSIM job weights only its own partials).
          bool twin_fail = false;
          auto twin_iv = [&](int twin_job, int iv_job) {
            if (l != 0 || cl != 0 || iv_job < 0 || P.ssim_gen) return;
            SsimJobDev& T = P.jobs[twin_job];
            SsimJobDev& S = P.jobs[iv_job];
            if (S.twin_part_off < 0 && T.nux == S.nux && T.nuy == S.nuy && T.nchunk == S.nchunk &&
                T.wsq == S.wsq && !T.erp && T.blk0 + 3 * T.nchunk * T.nbg == P.ssim_blocks) {
              P.ssim_blocks = T.blk0;
              S.twin_part_off = T.part_off;
            } else if (S.mom_off >= 0) {
              twin_fail = true;  // a fused pair has no SCP picture to read
            }
          };
          RCCHK(add_job(i, &pp.ivms_t2r[l], ref_for(i, 0, l), ref_for(i, 3, l), w, h, cl, 0));
          twin_iv(pp.ivms_t2r[l], pp.iv_t2r);
          RCCHK(add_job(i, &pp.ivms_r2t[l], ref_for(i, 1, l), ref_for(i, 2, l), w, h, cl, 0));
          twin_iv(pp.ivms_r2t[l], pp.iv_r2t);
          if (twin_fail || (l == 0 && pp.fused && cl != 0))
            return set_error(IVQM_EINVAL, "internal: fused IV-MS-SSIM level 0 without its IV-SSIM twin");
        }
      }
    }
  }
  // A captured graph replays its launches with their captured grids and its copies with their captured
  // sizes; only the tables are refreshed. Everything that shapes a launch or a copy therefore enters the
  // signature: besides the pair shapes, the block counts of every phase, the masked search forms, the
  // SSIM kernel forms and the pyramid jobs' component counts (level 1 of a 4:2:0 uint16 picture writes
  // luma only when the picture's chroma planes are aligned, which depends on the caller's pointers).
  std::string launch_sig;
  char lb[160];
  std::snprintf(lb, sizeof lb, "s%d t%d/%d/%x mt%d mf%x j%zu sb%d sf%x g%d am%d r%lld rd%lld r64%lld|",
                P.stat_blocks, P.tiles, P.tiles420, P.sforms, P.mtiles, P.mforms, P.jobs.size(), P.ssim_blocks,
                ssim_forms(P), P.ssim_gen ? 1 : 0,
                P.any_mask ? 1 : 0, (long long)P.rows_total, (long long)P.rowsd_total, (long long)P.rows64_total);
  launch_sig += lb;
  for (int l = 1; l < 5; l++) {
    std::snprintf(lb, sizeof lb, "L%d:%zu:%d:", l, P.down[l].size(), P.down_blocks[l]);
    launch_sig += lb;
    for (const DownJob& dj : P.down[l]) launch_sig += (char)('0' + dj.ncomp);
  }
  P.signature += launch_sig;
  return IVQM_OK;
}


int Engine::ensure_scp(int64_t elems) {
  if (elems <= cap_scp_) return IVQM_OK;
  RCCHK(sync());
  cap_scp_ = grow(elems, cap_scp_);
  RCCHK(dmalloc(&d_scp_, cap_scp_));
  scp_valid_ = false;
  drop_graphs();
  return IVQM_OK;
}

int Engine::ensure_capacity(const BatchPlan& P) {
  const int64_t njobs = (int64_t)P.jobs.size();
  int64_t ndown = 0;
  for (int l = 1; l < 5; l++) ndown += (int64_t)P.down[l].size();
  const bool need = P.n > cap_pairs_ || P.rows_total > cap_rows_ || P.rowsd_total > cap_rowsd_ ||
                    P.scp_total > cap_scp_ || P.pyr_total > cap_pyr_ || P.partial_total > cap_partial_ ||
                    njobs > cap_jobs_ || ndown > cap_down_ || P.aux_total > cap_aux_ ||
                    P.rows64_total > cap_rows64_;
  const int64_t best_need = best_on_ && P.n == 1 ? 2 * (int64_t)P.pp[0].w * P.pp[0].h : 0;
  if (!need && best_need <= cap_best_) return IVQM_OK;
  RCCHK(sync());
  if (P.n > cap_pairs_) {
    cap_pairs_ = grow(P.n, cap_pairs_);
    RCCHK(dmalloc(&d_pairs_, cap_pairs_));
    RCCHK(dmalloc(&d_gcd_, 4 * cap_pairs_));
    RCCHK(dmalloc(&d_sd_, 3 * cap_pairs_));
    RCCHK(dmalloc(&d_ssd_, 3 * cap_pairs_));
    RCCHK(dmalloc(&d_mcount_, 4 * cap_pairs_));
  }
  if (P.aux_total > cap_aux_) { cap_aux_ = grow(P.aux_total, cap_aux_); RCCHK(dmalloc(&d_aux_, cap_aux_)); }
  if (P.rows64_total > cap_rows64_) {
    cap_rows64_ = grow(P.rows64_total, cap_rows64_);
    RCCHK(dmalloc(&d_rows64_, cap_rows64_));
  }
  if (P.rows_total > cap_rows_) { cap_rows_ = grow(P.rows_total, cap_rows_); RCCHK(dmalloc(&d_rows_, cap_rows_)); }
  if (P.rowsd_total > cap_rowsd_) {
    cap_rowsd_ = grow(P.rowsd_total, cap_rowsd_);
    RCCHK(dmalloc(&d_row_sd_, cap_rowsd_));
    RCCHK(dmalloc(&d_row_ssd_, cap_rowsd_));
  }
  if (P.scp_total > cap_scp_) RCCHK(ensure_scp(P.scp_total));
  if (P.pyr_total > cap_pyr_) { cap_pyr_ = grow(P.pyr_total, cap_pyr_); RCCHK(dmalloc(&d_pyr_, cap_pyr_)); }
  if (P.partial_total > cap_partial_) {
    cap_partial_ = grow(P.partial_total, cap_partial_);
    RCCHK(dmalloc(&d_partial_, cap_partial_));
  }
  if (njobs > cap_jobs_) {
    cap_jobs_ = grow(njobs, cap_jobs_);
    RCCHK(dmalloc(&d_jobs_, cap_jobs_));
    RCCHK(dmalloc(&d_ssim_out_, 6 * cap_jobs_));
  }
  if (ndown > cap_down_) { cap_down_ = grow(ndown, cap_down_); RCCHK(dmalloc(&d_down_, cap_down_)); }
  if (best_need > cap_best_) { cap_best_ = best_need; RCCHK(dmalloc(&d_best_, cap_best_)); }
  // device buffers moved: captured graphs are stale
  drop_graphs();
  return IVQM_OK;
}

// Pinned tables and result buffers belong to one slot: growing them never
// touches another slot, whose results may still wait to be collected.
int Engine::ensure_slot_capacity(RingSlot& S, const BatchPlan& P) {
  const int64_t nj = (int64_t)P.jobs.size();
  int64_t nd = 0;
  for (int l = 1; l < 5; l++) nd += (int64_t)P.down[l].size();
  bool moved = false;
  if (P.n > S.c_pairs) {
    S.c_pairs = grow(P.n, S.c_pairs);
    RCCHK(hmalloc(&S.t_pairs, S.c_pairs));
    RCCHK(hmalloc(&S.r_gcd, 4 * S.c_pairs));
    RCCHK(hmalloc(&S.r_sd, 3 * S.c_pairs));
    RCCHK(hmalloc(&S.r_ssd, 3 * S.c_pairs));
    RCCHK(hmalloc(&S.r_mcount, 4 * S.c_pairs));
    moved = true;
  }
  if (P.rows64_total > S.c_rows64 || !S.r_rows64) {
    S.c_rows64 = grow(std::max<int64_t>(P.rows64_total, 1), S.c_rows64);
    RCCHK(hmalloc(&S.r_rows64, S.c_rows64));
    moved = true;
  }
  if (nj > S.c_jobs || !S.t_jobs) {
    S.c_jobs = grow(std::max<int64_t>(nj, 1), S.c_jobs);
    RCCHK(hmalloc(&S.t_jobs, S.c_jobs));
    RCCHK(hmalloc(&S.r_ssim, 6 * S.c_jobs));
    moved = true;
  }
  if (nd > S.c_down || !S.t_down) {
    S.c_down = grow(std::max<int64_t>(nd, 1), S.c_down);
    RCCHK(hmalloc(&S.t_down, S.c_down));
    moved = true;
  }
  if (P.rowsd_total > S.c_rowsd || !S.r_row_ssd) {
    S.c_rowsd = grow(std::max<int64_t>(P.rowsd_total, 1), S.c_rowsd);
    RCCHK(hmalloc(&S.r_row_ssd, S.c_rowsd));
    moved = true;
  }
  if (P.rows_total > S.c_rows || !S.r_rows) {
    S.c_rows = grow(std::max<int64_t>(P.rows_total, 1), S.c_rows);
    RCCHK(hmalloc(&S.r_rows, S.c_rows));
    moved = true;
  }
  if (moved && S.graph) {  // its copy nodes refer to the old buffers
    (void)hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
    S.graph_sig.clear();
  }
  return IVQM_OK;
}

int Engine::materialize(const PicRef& r, PicDev& out) const {
  const char* why = nullptr;
  const int rc = resolve_picture(r, d_scp_, cap_scp_, d_pyr_, cap_pyr_, out, &why);
  return rc ? set_error(rc, why) : IVQM_OK;
}

// Writes the slot's pinned tables (pairs, SSIM jobs, pyramid jobs) with device pointers.
int Engine::fill_tables(const BatchPlan& P, RingSlot& S, int* ndown, int down_first[5]) {
  std::memcpy(S.t_pairs, P.pairs.data(), sizeof(PairDev) * P.n);
  const int nj = (int)P.jobs.size();
  for (int j = 0; j < nj; j++) {
    S.t_jobs[j] = P.jobs[j];
    RCCHK(materialize(P.job_a[j], S.t_jobs[j].a));
    RCCHK(materialize(P.job_b[j], S.t_jobs[j].b));
    SsimJobDev& J = S.t_jobs[j];
    J.mom = nullptr;
    if (J.mom_off >= 0) {
      if (!d_scp_ || J.mom_off % 128 || J.mom_off + 3 * (int64_t)J.mom_rows * J.mom_pitch * 8 > cap_scp_)
        return set_error(IVQM_EINVAL, "internal: fused moments outside the SCP arena");
      J.mom = reinterpret_cast<const uint32_t*>(d_scp_ + J.mom_off);
    }
  }
  int nd = 0;
  for (int l = 0; l < 5; l++) down_first[l] = 0;
  for (int l = 1; l < 5; l++) {
    down_first[l] = nd;
    for (size_t k = 0; k < P.down[l].size(); k++) {
      DownJob d = P.down[l][k];
      RCCHK(materialize(P.down_src[l][k], d.src));
      PicDev dst;
      RCCHK(materialize(P.down_dst[l][k], dst));
      if (P.down_dst[l][k].arena != 2 || dst.stride[0] != d.pitch)
        return set_error(IVQM_EINVAL, "internal: pyramid destination mismatch");
      d.dst = (uint16_t*)dst.p[0];
      S.t_down[nd++] = d;
    }
  }
  *ndown = nd;
  return IVQM_OK;
}

// Both search kernels of a batch: k_search over its tiles, k_search420 over the keyed 4:2:0 pairs' tiles.
hipError_t Engine::launch_search_all(const BatchPlan& P, hipStream_t st) {
  hipError_t e = launch_search(d_pairs_, P.n, P.tiles, cfg_.search_range, cfg_.cmp_weights_search, d_gcd_, d_rows_,
                               d_scp_, d_best_, P.sforms, st);
  if (e != hipSuccess) return e;
  return launch_search420(d_pairs_, P.n, P.tiles420, cfg_.search_range, P.th420, cfg_.cmp_weights_search, d_gcd_,
                          d_rows_, d_scp_, d_best_, st);
}

int Engine::issue(const BatchPlan& P, RingSlot& S) {
  const hipStream_t st = stream_;
  const int n = P.n;
  const int nj = (int)P.jobs.size();
  int nd = 0, down_first[5];
  RCCHK(fill_tables(P, S, &nd, down_first));
  HIPCHK(hipMemcpyAsync(d_pairs_, S.t_pairs, sizeof(PairDev) * n, hipMemcpyHostToDevice, st));
  if (nj) HIPCHK(hipMemcpyAsync(d_jobs_, S.t_jobs, sizeof(SsimJobDev) * nj, hipMemcpyHostToDevice, st));
  if (nd) HIPCHK(hipMemcpyAsync(d_down_, S.t_down, sizeof(DownJob) * nd, hipMemcpyHostToDevice, st));
  if (P.need_stats) {
    HIPCHK(launch_stats(d_pairs_, n, P.stat_blocks, d_row_sd_, d_row_ssd_, d_rows_, d_aux_, d_rows64_, stats_forms(P),
                        st));
    HIPCHK(launch_gcd_finalize(d_pairs_, n, d_row_sd_, d_row_ssd_, d_aux_, d_gcd_, d_sd_, d_ssd_, d_mcount_, st));
  }
  if (P.mtiles)
    HIPCHK(launch_search_masked(d_pairs_, n, P.mtiles, cfg_.search_range, cfg_.cmp_weights_search, d_gcd_, d_rows64_,
                                d_scp_, P.mforms, st));
  if (P.need_search)
    HIPCHK(launch_search_all(P, st));
  for (int l = 1; l < 5; l++)
    if (!P.down[l].empty())
      HIPCHK(launch_downsample(d_down_ + down_first[l], (int)P.down[l].size(), P.down_blocks[l], st));
  if (nj) {
    if (P.ssim_gen)
      HIPCHK(launch_ssim_gen(d_jobs_, nj, P.ssim_blocks, cfg_.ssim_mode, cfg_.ssim_window, cfg_.ssim_border, cfg_.ssim_stride,
                             d_partial_, st));
    else
      HIPCHK(launch_ssim(d_jobs_, nj, P.ssim_blocks, P.nb, d_erp_, d_partial_, ssim_forms(P), st));
    HIPCHK(launch_ssim_finalize(d_jobs_, nj, d_partial_, d_erp_, d_ssim_out_, st));
  }
  if (P.need_stats) {
    HIPCHK(hipMemcpyAsync(S.r_gcd, d_gcd_, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(S.r_sd, d_sd_, sizeof(int64_t) * 3 * n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(S.r_ssd, d_ssd_, sizeof(uint64_t) * 3 * n, hipMemcpyDeviceToHost, st));
    if ((P.metrics & IVQM_M_WSPSNR) && cfg_.erp)
      HIPCHK(hipMemcpyAsync(S.r_row_ssd, d_row_ssd_, sizeof(uint64_t) * P.rowsd_total, hipMemcpyDeviceToHost, st));
  }
  if (P.metrics & IVQM_M_IVPSNR)
    HIPCHK(hipMemcpyAsync(S.r_rows, d_rows_, sizeof(uint32_t) * P.rows_total, hipMemcpyDeviceToHost, st));
  if (P.any_mask && P.need_stats)
    HIPCHK(hipMemcpyAsync(S.r_mcount, d_mcount_, sizeof(int64_t) * 4 * n, hipMemcpyDeviceToHost, st));
  if (P.mtiles)
    HIPCHK(hipMemcpyAsync(S.r_rows64, d_rows64_, sizeof(uint64_t) * P.rows64_total, hipMemcpyDeviceToHost, st));
  if (nj) HIPCHK(hipMemcpyAsync(S.r_ssim, d_ssim_out_, sizeof(double) * 6 * nj, hipMemcpyDeviceToHost, st));
  return IVQM_OK;
}

int Engine::enqueue(const PairSpec* pairs, int n, uint32_t metrics, int ring, bool use_graph) {
  if (n <= 0) return set_error(IVQM_EINVAL, "empty batch");
  if (ring < 0 || ring >= kRing) return set_error(IVQM_EINVAL, "ring slot out of range");
  RingSlot& S = ring_[ring];
  if (S.busy) {
    HIPCHK(hipEventSynchronize(S.ev));
    S.busy = false;
  }
  BatchPlan P;
  RCCHK(plan(pairs, n, metrics, P));
  RCCHK(ensure_capacity(P));
  RCCHK(ensure_slot_capacity(S, P));
  // Device pointers reach the kernels only through the tables, which a replay refreshes through the
  // captured staging copy; the signature (plan()) covers everything that shapes a launch or a copy.
  if (use_graph) {
    if (S.graph && S.graph_sig != P.signature) {
      (void)hipGraphExecDestroy(S.graph);
      S.graph = nullptr;
    }
    if (!S.graph) {
      hipGraph_t g = nullptr;
      HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
      int rc = issue(P, S);
      hipError_t e = hipStreamEndCapture(stream_, &g);
      if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
      if (e != hipSuccess) return hip_error(e, "hipStreamEndCapture");
      e = hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) { S.graph = nullptr; return hip_error(e, "hipGraphInstantiate"); }
      S.graph_sig = P.signature;
    } else {
      // same shapes: refresh the staging tables the graph's upload nodes read
      int nd = 0, down_first[5];
      RCCHK(fill_tables(P, S, &nd, down_first));
    }
    HIPCHK(hipGraphLaunch(S.graph, stream_));
  } else {
    RCCHK(issue(P, S));
  }
  HIPCHK(hipEventRecord(S.ev, stream_));
  S.busy = true;
  S.plan = P;
  last_ = P;
  if (P.need_scp) scp_valid_ = (n == 1);
  return IVQM_OK;
}

namespace {
double ssim_value(const BatchPlan& P, const double* out, int idx, int comp) {
  const SsimJobDev& J = P.jobs[idx];
  const double sum = out[(size_t)idx * 6 + comp * 2 + 0];
  const double wsum = out[(size_t)idx * 6 + comp * 2 + 1];
  const double active = (double)((int64_t)J.nuy * J.nux);
  if (J.erp) return sum / (active * (wsum / J.nuy));  // xIVSSIM.cpp:158-163
  return sum / active;                                 // xSSIM.cpp:233-236
}
// libs/pmbb-ivqm/src/xSSIM.cpp:132-140: product of ReLU(score)^exponent
void ms_combine(const BatchPlan& P, const double* out, int ns, const int* idx, double v4[4]) {
  for (int comp = 0; comp < 3; comp++) {
    double v = 1.0;
    for (int l = 0; l < ns; l++) {
      const double s = idx[l] >= 0 ? ssim_value(P, out, idx[l], comp) : 0.0;
      v *= pow(s > 0 ? s : 0.0, kMsExp[ns - 1][l]);
    }
    v4[comp] = v;
  }
  double v3 = 1.0;  // lane 3 starts at 1 and is multiplied by pow(ReLU(0), w_i)
  for (int l = 0; l < ns; l++) v3 *= pow(0.0, kMsExp[ns - 1][l]);
  v4[3] = v3;
}
}  // namespace

int Engine::query(int ring) {
  if (ring < 0 || ring >= kRing) return set_error(IVQM_EINVAL, "ring slot out of range");
  RingSlot& S = ring_[ring];
  if (!S.busy) return 1;
  const hipError_t e = hipEventQuery(S.ev);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  return hip_error(e, "hipEventQuery");
}

// Mask mode as the app computes it (xAppQMIV.cpp:948-1029): the result's ssd[] holds the masked
// SSD (sum of (P0 - P1)^2 x mask); MSE = SSDM / (NNM x MaxM); PSNR uses the mask's bit depth for
// the picture peak, as calcPicPSNRM passes it (xPSNR.cpp:60), and so does WS-PSNR without ERP
// (xWSPSNR.cpp:61-64); WS-PSNR with ERP and IV-PSNR use the picture's (xWSPSNR.cpp:128-133,
// xIVPSNR.cpp:190). Exact components get the app's fake PSNR. IV-PSNR labels follow
// calcPicIVPSNRM (xIVPSNR.cpp:161-162): R2T is P1 - GCD searched in P0.
int Engine::masked_epilogue(const RingSlot& S, const PairPlan& pp, int i, ivqm_frame_result& r) {
  const uint32_t m = pp.m;
  const int H = pp.h, bd = pp.bd, bdm = pp.bdm;
  const int64_t area = (int64_t)pp.w * H;
  const int32_t* cwa = cfg_.cmp_weights_average;
  const int64_t nnm = S.r_mcount[4 * i];
  if (nnm <= 0) return set_error(IVQM_EINVAL, "the mask has no non-zero sample");
  const int64_t maxm = (1 << bdm) - 1;
  const double fake = host::psnr_from_ssd(1.0, area, bd);
  if (m & (IVQM_M_PSNR | IVQM_M_MSE)) {
    for (int k = 0; k < 4; k++) { r.mse[k] = DBL_MAX; r.psnr[k] = DBL_MAX; }
    for (int k = 0; k < 3; k++) {
      const double ssdm = (double)r.ssd[k];
      r.mse[k] = ssdm / (double)(nnm * maxm);
      r.psnr[k] = host::psnr_masked(ssdm, nnm, bdm, bdm);
      if (cfg_.fake_vals_for_exact && r.exact[k]) r.psnr[k] = fake;
    }
    r.psnr_pic = host::pic_value(r.psnr, cwa);
  }
  if (m & IVQM_M_WSPSNR) {
    for (int k = 0; k < 4; k++) r.wspsnr[k] = DBL_MAX;
    if (!cfg_.erp) {
      for (int k = 0; k < 3; k++) r.wspsnr[k] = host::psnr_masked((double)r.ssd[k], nnm, bdm, bdm);
    } else {
      const auto& hw = erp_weights(H);
      const uint64_t* rs = S.r_row_ssd + pp.rowsd_off;  // masked rows are 4:4:4 rows of every component
      for (int k = 0; k < 3; k++) {
        host::Kbns kb;
        for (int y = 0; y < H; y++) kb.acc((double)rs[(size_t)k * H + y] * hw.first[y]);
        r.wspsnr[k] = host::psnr_masked(kb.result() * hw.second, nnm, bd, bdm);
      }
    }
    for (int k = 0; k < 3; k++)
      if (cfg_.fake_vals_for_exact && r.exact[k]) r.wspsnr[k] = fake;
    r.wspsnr_pic = host::pic_value(r.wspsnr, cwa);
  }
  if (m & IVQM_M_IVPSNR) {
    const uint64_t* rows = S.r_rows64 + pp.rows64_off;
    const double* ew = cfg_.erp ? erp_weights(H).first.data() : nullptr;
    r.ivpsnr_t2r = host::ivpsnr_quality_masked(rows, H, nnm, bd, bdm, ew, cwa);                   // P0 + GCD in P1
    r.ivpsnr_r2t = host::ivpsnr_quality_masked(rows + (size_t)3 * H, H, nnm, bd, bdm, ew, cwa);  // P1 - GCD in P0
    r.ivpsnr = std::min(r.ivpsnr_r2t, r.ivpsnr_t2r);
  }
  return IVQM_OK;
}

int Engine::collect(int ring, ivqm_frame_result* out) {
  if (ring < 0 || ring >= kRing) return set_error(IVQM_EINVAL, "ring slot out of range");
  RingSlot& S = ring_[ring];
  if (!S.busy) return set_error(IVQM_ESTATE, "ring slot has no pending batch");
  HIPCHK(hipEventSynchronize(S.ev));
  S.busy = false;
  const BatchPlan& P = S.plan;
  const int32_t* cwa = cfg_.cmp_weights_average;
  for (int i = 0; i < P.n; i++) {
    const PairPlan& pp = P.pp[i];
    const uint32_t m = pp.m;
    const int W = pp.w, H = pp.h, bd = pp.bd;
    const int64_t area = (int64_t)W * H;
    ivqm_frame_result r;
    std::memset(&r, 0, sizeof(r));
    r.metrics = m;
    if (P.need_stats) {
      for (int k = 0; k < 4; k++) r.gcd_r2t[k] = S.r_gcd[4 * i + k];
      for (int k = 0; k < 3; k++) {
        r.ssd[k] = S.r_ssd[3 * i + k];
        r.exact[k] = pp.masked ? S.r_mcount[4 * i + 1 + k] == 0 : r.ssd[k] == 0;
      }
    }
    if (pp.masked) {
      RCCHK(masked_epilogue(S, pp, i, r));
      if (out) out[i] = r;
      continue;
    }
    const double* ew = nullptr;
    double corr = 1.0;
    if (cfg_.erp && (m & (IVQM_M_WSPSNR | IVQM_M_IVPSNR))) {
      const auto& hw = erp_weights(H);
      ew = hw.first.data();
      corr = hw.second;
    }
    host::psnr_family(m, W, H, bd, pp.cs, pp.csy, cfg_.fake_vals_for_exact != 0, cwa, S.r_row_ssd + pp.rowsd_off,
                      S.r_rows + pp.rows_off, ew, corr, r);
    const double* so = S.r_ssim;
    if ((m & IVQM_M_SSIM) && pp.ssim >= 0) {
      for (int k = 0; k < 3; k++) r.ssim[k] = ssim_value(P, so, pp.ssim, k);
      r.ssim_pic = host::pic_value(r.ssim, cwa);
    }
    if ((m & IVQM_M_MSSSIM) && pp.ns >= 1) {
      ms_combine(P, so, pp.ns, pp.ms, r.msssim);
      r.msssim_pic = host::pic_value(r.msssim, cwa);
    }
    if ((m & IVQM_M_IVSSIM) && pp.iv_t2r >= 0) {
      double a[4] = {0, 0, 0, 0}, b[4] = {0, 0, 0, 0};
      for (int k = 0; k < 3; k++) {
        a[k] = ssim_value(P, so, pp.iv_t2r, k);
        b[k] = ssim_value(P, so, pp.iv_r2t, k);
      }
      r.ivssim_t2r = host::weighted_mean4(a, cwa);
      r.ivssim_r2t = host::weighted_mean4(b, cwa);
      r.ivssim = std::min(r.ivssim_t2r, r.ivssim_r2t);  // xIVSSIM.cpp:56
    }
    if ((m & IVQM_M_IVMSSSIM) && pp.ns >= 1) {
      double a[4], b[4];
      ms_combine(P, so, pp.ns, pp.ivms_t2r, a);
      ms_combine(P, so, pp.ns, pp.ivms_r2t, b);
      a[3] = 0.0;
      b[3] = 0.0;
      r.ivmsssim_t2r = host::weighted_mean4(a, cwa);
      r.ivmsssim_r2t = host::weighted_mean4(b, cwa);
      r.ivmsssim = std::min(r.ivmsssim_t2r, r.ivmsssim_r2t);  // xIVSSIM.cpp:87
    }
    if (out) out[i] = r;
  }
  return IVQM_OK;
}

int Engine::profile(int kernel, int iters, double* avg_ms) {
  const BatchPlan& P = last_;
  if (P.n <= 0) return set_error(IVQM_ESTATE, "no batch enqueued yet");
  RCCHK(sync());
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  const int nj = (int)P.jobs.size();
  // one untimed launch first: the first launch of a kernel also loads its code
  for (int i = -1; i < iters; i++) {
    if (i == 0) HIPCHK(hipEventRecord(e0, stream_));
    switch (kernel) {
      case 0:
        HIPCHK(launch_stats(d_pairs_, P.n, P.stat_blocks, d_row_sd_, d_row_ssd_, d_rows_, d_aux_, d_rows64_,
                            stats_forms(P), stream_));
        break;
      case 1:
        HIPCHK(launch_search_all(P, stream_));
        break;
      case 2:
        if (P.ssim_gen)
          HIPCHK(launch_ssim_gen(d_jobs_, nj, P.ssim_blocks, cfg_.ssim_mode, cfg_.ssim_window, cfg_.ssim_border, cfg_.ssim_stride,
                                 d_partial_, stream_));
        else
          HIPCHK(launch_ssim(d_jobs_, nj, P.ssim_blocks, P.nb, d_erp_, d_partial_, ssim_forms(P), stream_));
        break;
      case 3:
        HIPCHK(launch_search_masked(d_pairs_, P.n, P.mtiles, cfg_.search_range, cfg_.cmp_weights_search, d_gcd_,
                                    d_rows64_, d_scp_, P.mforms, stream_));
        break;
      default: return set_error(IVQM_EINVAL, "unknown kernel id");
    }
  }
  HIPCHK(hipEventRecord(e1, stream_));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = (double)ms / iters;
  return IVQM_OK;
}

}  // namespace ivqm
