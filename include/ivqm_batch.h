/*
 * ivqm_batch.h - batched form of the C ABI: many frame pairs of mixed sizes,
 * bit depths and chroma formats per call, one kernel launch per phase for the
 * whole batch, optionally replayed from a captured hipGraph.
 *
 * No counterpart in the reference: QMIV processes one frame pair at a time
 * (apps/QMIV/src/xAppQMIV.cpp:761). Each pair's results are those the
 * per-frame calls of include/ivqm_hip.h return for it; a pair's result does
 * not depend on the other pairs of its batch.
 */
#ifndef IVQM_BATCH_H
#define IVQM_BATCH_H
#include <stdint.h>
#include "ivqm_hip.h"
#ifdef __cplusplus
extern "C" {
#endif

/* One frame pair of a batch: device-resident planes (P0 = test, P1 = reference). */
typedef struct ivqm_pair {
  ivqm_picture p0, p1;
  int32_t width, height, bit_depth;
} ivqm_pair;

typedef struct ivqm_batch ivqm_batch;

/* cfg supplies the shared parameters (search range, weights, coefficients,
 * ERP, SSIM); its width / height / bit_depth are ignored. */
int ivqm_batch_create(const ivqm_config* cfg, ivqm_batch** out);
int ivqm_batch_destroy(ivqm_batch* b);
int ivqm_batch_set_stream(ivqm_batch* b, void* hip_stream);
/* Kernel shapes: the search tile height of 4:2:0 pairs (16 or 32 rows of 64
 * pixels; 4 and 8 are taken as 16 and 32, 0 keeps the current value; 4:4:4
 * pairs always use 64 x 16 tiles), SSIM window rows per wave (4, 8 or 16), and
 * whether each ring slot replays a captured hipGraph while the batch shape
 * stays the same (re-captured when it changes). */
int ivqm_batch_set_tuning(ivqm_batch* b, int32_t search_tile_rows, int32_t ssim_band_rows, int32_t use_graph);
int ivqm_batch_get_tuning(ivqm_batch* b, int32_t* search_tile_rows, int32_t* ssim_band_rows);
/* Times the candidate shapes on this batch and keeps the fastest. */
int ivqm_batch_autotune(ivqm_batch* b, const ivqm_pair* pairs, int32_t n, uint32_t metrics,
                        int32_t* search_tile_rows, int32_t* ssim_band_rows);
/* Synchronous: results[i] for pairs[i]. */
int ivqm_batch_run(ivqm_batch* b, const ivqm_pair* pairs, int32_t n, uint32_t metrics, ivqm_frame_result* results);
/* Asynchronous over a ring of 8 result slots. Even and odd slots run on two
 * device streams with their own buffers, so batches enqueued on consecutive
 * slots overlap on the device (a stream given by ivqm_batch_set_stream
 * serializes them). */
int ivqm_batch_enqueue(ivqm_batch* b, const ivqm_pair* pairs, int32_t n, uint32_t metrics, int32_t ring);
/* As ivqm_batch_enqueue, with one device-resident mask per pair (ivqm_hip.h,
 * mask mode); masks[i].plane == NULL leaves pair i unmasked. The app's mask
 * mode (apps/QMIV/src/xAppQMIV.cpp:930-1029) for whole batches. */
int ivqm_batch_enqueue_masked(ivqm_batch* b, const ivqm_pair* pairs, const ivqm_mask* masks, int32_t n,
                              uint32_t metrics, int32_t ring);
int ivqm_batch_collect(ivqm_batch* b, int32_t ring, ivqm_frame_result* results);

/* No-allocation form of ivqm_batch_run for the PSNR family (GCD, MSE, PSNR,
 * WS-PSNR, IV-PSNR). No counterpart in the reference: it is for callers that
 * cannot allocate device memory -- code that must not call hipMalloc /
 * hipFree (fixed-memory serving loops, harnesses that forbid allocation).
 * There is no batch object: cfg gives the shared parameters (its width,
 * height and bit_depth are ignored), pairs are device-resident as for
 * ivqm_batch_run, and results[i] equals what ivqm_batch_run returns for
 * pairs[i] with the same metrics. The work runs on hip_stream (NULL: the null
 * stream), ordered after what the caller enqueued there before; the call
 * returns once that stream has finished it.
 * Scratch memory (pair table, row partial sums, frame totals) is one fixed
 * device array of IVQM_NOALLOC_SCRATCH_BYTES per GPU, held until the process
 * ends and reused by every call on that GPU; results are staged in fixed host
 * arrays. A lock held over each whole call serializes calls from all threads,
 * so they never overlap on the shared scratch. On success nothing is
 * allocated. Limits (IVQM_EINVAL beyond them, before anything runs):
 * 1..IVQM_NOALLOC_MAX_PAIRS pairs whose heights add up to at most
 * IVQM_NOALLOC_MAX_ROWS, metrics within IVQM_M_PSNR | IVQM_M_MSE |
 * IVQM_M_WSPSNR | IVQM_M_IVPSNR, no mask. */
#define IVQM_NOALLOC_MAX_PAIRS 16
#define IVQM_NOALLOC_MAX_ROWS 32768
#define IVQM_NOALLOC_SCRATCH_BYTES (5u << 19) /* 2.5 MiB */
int ivqm_batch_run_noalloc(const ivqm_config* cfg, const ivqm_pair* pairs, int32_t n, uint32_t metrics,
                           ivqm_frame_result* results, void* hip_stream);

#ifdef __cplusplus
}
#endif
#endif
