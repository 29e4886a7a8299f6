/*
 * ivqm_hip_profile.h - diagnostics of the gfx950 metric kernels.
 *
 * No counterpart in the reference (QMIV times whole stages with RDTSC,
 * apps/QMIV/src/xAppQMIV.cpp:763-826); used by bench.py to time one kernel
 * family with HIP events on the context's stream.
 */
#ifndef IVQM_HIP_PROFILE_H
#define IVQM_HIP_PROFILE_H
#include <stdint.h>
#include "ivqm_hip.h"
#include "ivqm_batch.h"
#ifdef __cplusplus
extern "C" {
#endif

#define IVQM_K_STATS 0  /* k_stats: SD / SSD rows (streaming) */
#define IVQM_K_SEARCH 1 /* k_search: both directions + SCP */
#define IVQM_K_SSIM 2   /* k_ssim (or k_ssim_static / k_ssim_gen in the other SSIM modes): all SSIM jobs */
#define IVQM_K_SEARCH_MASKED 3 /* k_search_masked: both directions of the masked pairs */

/* Relaunches `kernel` `iters` times for the last enqueued pair / batch;
 * *avg_ms = mean duration (HIP events on the context's stream). */
int ivqm_profile(ivqm_ctx* ctx, int32_t kernel, int32_t iters, double* avg_ms);
int ivqm_batch_profile(ivqm_batch* b, int32_t kernel, int32_t iters, double* avg_ms);

#ifdef __cplusplus
}
#endif
#endif
