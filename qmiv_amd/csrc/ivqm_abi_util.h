// Argument checks and exception guard shared by the C ABI entry points (ivqm_runtime.hip,
// ivqm_noalloc.hip). Host code only.
#pragma once
#include <new>
#include <stdexcept>
#include <string>

#include "../../include/ivqm_hip.h"
#include "ivqm_engine.h"

namespace ivqm {

#ifndef IVQM_ABI_RCCHK
#define IVQM_ABI_RCCHK(expr) \
  do {                       \
    int _rc = (expr);        \
    if (_rc) return _rc;     \
  } while (0)
#endif

inline int validate_config(const ivqm_config* cfg, bool need_size) {
  if (!cfg) return set_error(IVQM_EINVAL, "cfg is NULL");
  if (need_size && (cfg->width < 8 || cfg->height < 8)) return set_error(IVQM_EINVAL, "picture must be at least 8x8");
  if (need_size && (cfg->bit_depth < 8 || cfg->bit_depth > 16)) return set_error(IVQM_EINVAL, "bit_depth must be 8..16");
  if (cfg->search_range < 1 || cfg->search_range > 4) return set_error(IVQM_EINVAL, "search_range must be 1..4");
  if (cfg->search_mode_cap < 0 || cfg->search_mode_cap > 2) return set_error(IVQM_EINVAL, "search_mode_cap must be 0..2");
  IVQM_ABI_RCCHK(validate_ssim_config(*cfg));
  if (cfg->erp && (cfg->lat_range_deg <= 0 || cfg->lat_range_deg > 180)) return set_error(IVQM_EINVAL, "lat_range_deg");
  if (cfg->cmp_weights_average[0] + cfg->cmp_weights_average[1] + cfg->cmp_weights_average[2] +
          cfg->cmp_weights_average[3] == 0)
    return set_error(IVQM_EINVAL, "component weights sum to zero");
  return IVQM_OK;
}

inline int validate_picture(int W, int H, int bd, const ivqm_picture* p) {
  if (!p) return set_error(IVQM_EINVAL, "picture is NULL");
  if (p->bytes_per_sample != 1 && p->bytes_per_sample != 2) return set_error(IVQM_EINVAL, "bytes_per_sample must be 1 or 2");
  if (p->bytes_per_sample == 1 && bd != 8) return set_error(IVQM_EINVAL, "8-bit samples require bit_depth 8");
  if (p->chroma_420 < 0 || p->chroma_420 > 2) return set_error(IVQM_EINVAL, "chroma_420 must be 0 (4:4:4), 1 (4:2:0) or 2 (4:2:2)");
  if (p->chroma_420 == 1 && ((W & 1) || (H & 1))) return set_error(IVQM_EINVAL, "4:2:0 requires even width and height");
  if (p->chroma_420 == 2 && (W & 1)) return set_error(IVQM_EINVAL, "4:2:2 requires an even width");
  const int cs = p->chroma_420 ? 1 : 0;
  for (int k = 0; k < 3; k++) {
    if (!p->plane[k]) return set_error(IVQM_EINVAL, "plane pointer is NULL");
    const int wp = k ? (W >> cs) : W;
    if (p->stride[k] < wp) return set_error(IVQM_EINVAL, "stride smaller than plane width");
  }
  return IVQM_OK;
}

// The SSIM kernel reads full-resolution planes four samples at a time.
inline bool aligned_for_kernels(const ivqm_picture* p) {
  const int bps = p->bytes_per_sample;
  for (int k = 0; k < 3; k++) {
    const bool full = (k == 0) || !p->chroma_420;
    if (full && ((((uintptr_t)p->plane[k]) % (size_t)(4 * bps)) || (p->stride[k] % 4))) return false;
  }
  return true;
}

inline PicDev to_dev(const ivqm_picture* p) {
  PicDev d;
  for (int k = 0; k < 3; k++) {
    d.p[k] = p->plane[k];
    d.stride[k] = p->stride[k];
  }
  d.cs = p->chroma_420 ? 1 : 0;
  d.csy = p->chroma_420 == 1 ? 1 : 0;
  d.bps = p->bytes_per_sample;
  return d;
}

// C++ exceptions (allocation failures, container errors) must not cross the C ABI.
template <typename F>
int guarded(F&& f) noexcept {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return set_error(IVQM_ENOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return set_error(IVQM_EINVAL, std::string("internal error: ") + e.what());
  } catch (...) {
    return set_error(IVQM_EINVAL, "internal error");
  }
}

}  // namespace ivqm
