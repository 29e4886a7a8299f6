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
/* Kernel shapes: search tile rows per wave (4 or 8 accepted; the search runs
 * 4-row waves, 64 x 16 tiles), SSIM window rows per
 * wave (4, 8 or 16), and whether each ring slot replays a captured hipGraph
 * while the batch shape stays the same (re-captured when it changes). */
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

#ifdef __cplusplus
}
#endif
#endif
