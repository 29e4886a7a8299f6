// No-allocation form of the batched PSNR family: ivqm_batch_run_noalloc (include/ivqm_batch.h).
//
// The kernels' scratch memory -- the pair table, the per-row partial sums and
// the frame totals -- is one file-scope device array, `scratch`. Every GPU
// holds its own copy at its own address: the address is taken once per device,
// with that device current, by hipGetSymbolAddress, kept in g_scratch[device],
// and only that address reaches kernels and copies (in host code the array's
// name is not a device address). Results are staged in file-scope host arrays.
//
// A process-wide lock is held over a whole call, which ends with a
// synchronization of the caller's stream, so calls never overlap and the
// scratch needs no clearing beyond what the kernels already do: k_stats writes
// every SD / SSD row partial and clears the search row sums before k_search
// adds into them, and k_gcd_finalize writes every frame total. The same
// launchers as the batch engine run the same kernels, so the results equal
// ivqm_batch_run's for the same pairs and metrics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "../../include/ivqm_batch.h"
#include "../../include/ivqm_hip.h"
#include "ivqm_abi_util.h"
#include "ivqm_device.h"
#include "ivqm_engine.h"
#include "ivqm_epilogue.h"

namespace ivqm {
namespace {

constexpr int kMaxPairs = IVQM_NOALLOC_MAX_PAIRS;
constexpr int kMaxRows = IVQM_NOALLOC_MAX_ROWS;  // sum of the pairs' heights
constexpr int kMaxDevices = 64;
constexpr uint32_t kFamily = IVQM_M_PSNR | IVQM_M_MSE | IVQM_M_WSPSNR | IVQM_M_IVPSNR;

constexpr size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }
// scratch layout (bytes)
constexpr size_t kOffPairs = 0;
constexpr size_t kOffGcd = align256(kOffPairs + sizeof(PairDev) * kMaxPairs);
constexpr size_t kOffSd = align256(kOffGcd + sizeof(int32_t) * 4 * kMaxPairs);
constexpr size_t kOffSsd = align256(kOffSd + sizeof(int64_t) * 3 * kMaxPairs);
constexpr size_t kOffRowSd = align256(kOffSsd + sizeof(uint64_t) * 3 * kMaxPairs);
constexpr size_t kOffRowSsd = align256(kOffRowSd + sizeof(int64_t) * 3 * (size_t)kMaxRows);
constexpr size_t kOffRows = align256(kOffRowSsd + sizeof(uint64_t) * 3 * (size_t)kMaxRows);
constexpr size_t kScratchBytes = align256(kOffRows + sizeof(uint32_t) * 6 * (size_t)kMaxRows);
static_assert(kScratchBytes <= IVQM_NOALLOC_SCRATCH_BYTES, "scratch layout exceeds the documented size");

static __device__ __align__(256) unsigned char scratch[IVQM_NOALLOC_SCRATCH_BYTES];

std::mutex g_lock;
void* g_scratch[kMaxDevices];  // scratch's address on each device (null until first used there)

// host staging, used under g_lock only
PairDev h_pairs[kMaxPairs];
int32_t h_gcd[4 * kMaxPairs];
uint64_t h_ssd[3 * kMaxPairs];
uint64_t h_row_ssd[3 * (size_t)kMaxRows];
uint32_t h_rows[6 * (size_t)kMaxRows];
double h_erp[kMaxRows];

int scratch_address(int dev, unsigned char** out) {
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!p) return set_error(IVQM_EHIP, "hipGetSymbolAddress(scratch) returned no address");
    g_scratch[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[dev]);
  return IVQM_OK;
}

#define NA_HIPCHK(expr)                                \
  do {                                                 \
    hipError_t _e = (expr);                            \
    if (_e != hipSuccess) return hip_error(_e, #expr); \
  } while (0)

int run(const ivqm_config& cfg, const ivqm_pair* pairs, int n, uint32_t m, ivqm_frame_result* out, hipStream_t st) {
  // checks before anything runs (the batch ABI's picture rules plus this entry point's limits)
  int64_t total_rows = 0;
  for (int i = 0; i < n; i++) {
    const ivqm_pair& p = pairs[i];
    if (p.width < 8 || p.height < 8 || p.bit_depth < 8 || p.bit_depth > 16)
      return set_error(IVQM_EINVAL, "bad pair size or bit depth");
    IVQM_ABI_RCCHK(validate_picture(p.width, p.height, p.bit_depth, &p.p0));
    IVQM_ABI_RCCHK(validate_picture(p.width, p.height, p.bit_depth, &p.p1));
    if (p.p0.chroma_420 != p.p1.chroma_420 || p.p0.bytes_per_sample != p.p1.bytes_per_sample)
      return set_error(IVQM_EINVAL, "pictures of a pair must share chroma format and sample size");
    if (!aligned_for_kernels(&p.p0) || !aligned_for_kernels(&p.p1))
      return set_error(IVQM_EINVAL, "batch pictures need 4-sample-aligned full-resolution planes and strides");
    const int csy = p.p0.chroma_420 == 1 ? 1 : 0;
    for (int c = 0; c < 3; c++) {  // the kernels address a plane with 32-bit byte offsets
      const int64_t rows = c ? (p.height + csy) >> csy : p.height;
      if ((int64_t)p.p0.stride[c] * rows * p.p0.bytes_per_sample >= (1ll << 32) ||
          (int64_t)p.p1.stride[c] * rows * p.p1.bytes_per_sample >= (1ll << 32))
        return set_error(IVQM_EINVAL, "picture plane of 4 GiB or more");
    }
    total_rows += p.height;
  }
  if (total_rows > kMaxRows)
    return set_error(IVQM_EINVAL, "the pairs' heights add up to more than IVQM_NOALLOC_MAX_ROWS");

  std::lock_guard<std::mutex> lock(g_lock);
  int dev = 0;
  NA_HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return set_error(IVQM_EINVAL, "device index above the supported range");
  unsigned char* base = nullptr;
  IVQM_ABI_RCCHK(scratch_address(dev, &base));
  PairDev* d_pairs = reinterpret_cast<PairDev*>(base + kOffPairs);
  int32_t* d_gcd = reinterpret_cast<int32_t*>(base + kOffGcd);
  int64_t* d_sd = reinterpret_cast<int64_t*>(base + kOffSd);
  uint64_t* d_ssd = reinterpret_cast<uint64_t*>(base + kOffSsd);
  int64_t* d_row_sd = reinterpret_cast<int64_t*>(base + kOffRowSd);
  uint64_t* d_row_ssd = reinterpret_cast<uint64_t*>(base + kOffRowSsd);
  uint32_t* d_rows = reinterpret_cast<uint32_t*>(base + kOffRows);

  // pair table: the batch engine's plan of an unmasked pair without SSIM metrics
  const int TH = kSearchTileH, TH420 = search420_tile_rows(cfg.search_tile_rows, 32);
  int stat_blocks = 0, tiles = 0, tiles420 = 0;
  uint32_t sforms = 0;  // k_search kernels needed (launch_search)
  int64_t rowsd_off = 0, rows_off = 0;
  for (int i = 0; i < n; i++) {
    const ivqm_pair& p = pairs[i];
    PairDev& d = h_pairs[i];
    std::memset(&d, 0, sizeof(d));
    d.p0 = to_dev(&p.p0);
    d.p1 = to_dev(&p.p1);
    d.w = p.width;
    d.h = p.height;
    d.maxv = (1 << p.bit_depth) - 1;
    int32_t gmax = 0;
    for (int k = 0; k < 3; k++) {
      d.md[k] = host::round_f32(cfg.unnoticeable_coef[k] * (float)d.maxv);  // xGlobClrDiff.cpp:29
      gmax = std::max<int32_t>(gmax, std::abs(d.md[k]));
    }
    d.mode = pick_search_mode(cfg, d.maxv, gmax);
    d.wrap32 = (cfg.gcd_sd_wrap32 && p.bit_depth <= 14) ? 1 : 0;
    d.res_idx = i;
    d.stat_blk0 = stat_blocks;
    stat_blocks += p.height + 2 * (p.height >> d.p0.csy);
    d.tile0 = tiles;
    d.tile0b = tiles420;
    if (chroma_format(d.p0) == 1 && (d.mode == 1 || d.mode == 2) && cfg.search_mode_cap == 0 &&
        search420_ok(cfg, d.maxv, gmax))
      d.mode = kSearchDot420;
    if ((m & IVQM_M_IVPSNR) && d.mode == kSearchDot420)
      tiles420 += ((p.width + kSearchTileW - 1) / kSearchTileW) * ((p.height + TH420 - 1) / TH420);
    else if (m & IVQM_M_IVPSNR) {
      tiles += ((p.width + kSearchTileW - 1) / kSearchTileW) * ((p.height + TH - 1) / TH);
      sforms |= search_form_bit(d.mode);
    }
    d.rowsd_off = rowsd_off;
    rowsd_off += 3 * (int64_t)p.height;
    d.rows_off = rows_off;
    rows_off += 6 * (int64_t)p.height;
    d.scp_pitch = 0;
    d.scp_off = -1;
    d.best_off = -1;
    d.mom_off = -1;
    d.scp1_off = -1;
  }
  // every offset stays inside its scratch region
  if (rowsd_off > 3 * (int64_t)kMaxRows || rows_off > 6 * (int64_t)kMaxRows)
    return set_error(IVQM_EINVAL, "internal: row partials outside the scratch array");

  NA_HIPCHK(hipMemcpyAsync(d_pairs, h_pairs, sizeof(PairDev) * n, hipMemcpyHostToDevice, st));
  uint32_t stforms = 0;  // k_stats kernels needed
  for (int i = 0; i < n; i++) stforms |= stats_form_bit(h_pairs[i]);
  NA_HIPCHK(launch_stats(d_pairs, n, stat_blocks, d_row_sd, d_row_ssd, d_rows, nullptr, nullptr, stforms, st));
  NA_HIPCHK(launch_gcd_finalize(d_pairs, n, d_row_sd, d_row_ssd, nullptr, d_gcd, d_sd, d_ssd, nullptr, st));
  if (tiles)
    NA_HIPCHK(launch_search(d_pairs, n, tiles, cfg.search_range, cfg.cmp_weights_search, d_gcd, d_rows, nullptr,
                            nullptr, sforms, st));
  if (tiles420)
    NA_HIPCHK(launch_search420(d_pairs, n, tiles420, cfg.search_range, TH420, cfg.cmp_weights_search, d_gcd, d_rows,
                               nullptr, nullptr, st));
  NA_HIPCHK(hipMemcpyAsync(h_gcd, d_gcd, sizeof(int32_t) * 4 * n, hipMemcpyDeviceToHost, st));
  NA_HIPCHK(hipMemcpyAsync(h_ssd, d_ssd, sizeof(uint64_t) * 3 * n, hipMemcpyDeviceToHost, st));
  if ((m & IVQM_M_WSPSNR) && cfg.erp)
    NA_HIPCHK(hipMemcpyAsync(h_row_ssd, d_row_ssd, sizeof(uint64_t) * rowsd_off, hipMemcpyDeviceToHost, st));
  if (m & IVQM_M_IVPSNR)
    NA_HIPCHK(hipMemcpyAsync(h_rows, d_rows, sizeof(uint32_t) * rows_off, hipMemcpyDeviceToHost, st));
  NA_HIPCHK(hipStreamSynchronize(st));

  for (int i = 0; i < n; i++) {
    const PairDev& d = h_pairs[i];
    ivqm_frame_result r;
    std::memset(&r, 0, sizeof(r));
    r.metrics = m;
    for (int k = 0; k < 4; k++) r.gcd_r2t[k] = h_gcd[4 * i + k];
    for (int k = 0; k < 3; k++) {
      r.ssd[k] = h_ssd[3 * i + k];
      r.exact[k] = r.ssd[k] == 0;
    }
    const double* ew = nullptr;
    double corr = 1.0;
    if (cfg.erp && (m & (IVQM_M_WSPSNR | IVQM_M_IVPSNR))) {
      corr = host::ws_weights_into(d.h, cfg.lat_range_deg, h_erp);
      ew = h_erp;
    }
    host::psnr_family(m, d.w, d.h, pairs[i].bit_depth, d.p0.cs, d.p0.csy, cfg.fake_vals_for_exact != 0,
                      cfg.cmp_weights_average, h_row_ssd + d.rowsd_off, h_rows + d.rows_off, ew, corr, r);
    out[i] = r;
  }
  return IVQM_OK;
}

}  // namespace
}  // namespace ivqm

extern "C" int ivqm_batch_run_noalloc(const ivqm_config* cfg, const ivqm_pair* pairs, int32_t n, uint32_t metrics,
                                      ivqm_frame_result* results, void* hip_stream) {
  using namespace ivqm;
  return guarded([&]() -> int {
    IVQM_ABI_RCCHK(validate_config(cfg, false));
    if (!pairs || !results) return set_error(IVQM_EINVAL, "NULL argument");
    if (n <= 0 || n > kMaxPairs) return set_error(IVQM_EINVAL, "pair count outside 1..IVQM_NOALLOC_MAX_PAIRS");
    if (!metrics || (metrics & ~kFamily))
      return set_error(IVQM_EINVAL, "ivqm_batch_run_noalloc computes IVQM_M_PSNR, MSE, WSPSNR and IVPSNR only");
    return run(*cfg, pairs, n, metrics, results, static_cast<hipStream_t>(hip_stream));
  });
}
