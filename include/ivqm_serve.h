/*
 * Serving entry points of the gfx950 metric hot path: a request loop that
 * accepts frame-pair jobs of mixed resolution and bit depth as they arrive,
 * keeps reference sequences resident in device memory, admits jobs against a
 * device-memory budget and batches the frame pairs of every admitted job
 * continuously (pairs of different jobs, sizes and metric sets share one
 * launch per kernel phase; batches overlap on the device).
 *
 * The reference has no serving mode: a job is the per-frame loop of
 * apps/QMIV/src/xAppQMIV.cpp:927-1108 (GCD -> SCP -> IV-PSNR -> IV-SSIM per
 * frame pair) run for `num_frames` pairs, with the same results the
 * per-frame calls of ivqm_hip.h give for each pair (bit-exact IV-PSNR, PSNR,
 * WS-PSNR; the SSIM family from the same kernels). These entry points are the
 * "submit / poll" boundary of a C-ABI replacement; they replace no reference
 * function.
 *
 * Threading: submit, poll, drain, stats and the sequence calls may be made
 * from any host thread; one worker thread per server drives the device.
 */
#ifndef IVQM_SERVE_H
#define IVQM_SERVE_H

#include <stdint.h>

#include "ivqm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define IVQM_MAX_JOB_FRAMES 8

typedef struct ivqm_server ivqm_server;

typedef struct ivqm_server_config {
  ivqm_config params;          /* metric parameters (search range, weights, ERP, ...); width, height and
                                  bit_depth are taken from each sequence */
  uint64_t memory_budget;      /* device bytes for staged test frames of admitted jobs; 0 = 64 GiB */
  int32_t max_batch_pairs;     /* frame pairs per batch launch (1..256); 0 = 16 */
  int32_t max_inflight;        /* batches on the device at once (1..8); 0 = 4 */
  int32_t reserved[8];
} ivqm_server_config;

/* A reference sequence kept resident on the device. Frames are copied once
 * (from host or device memory) into the server's own allocation. */
typedef struct ivqm_sequence_desc {
  int32_t width, height, bit_depth;
  int32_t chroma_420;          /* chroma format: 1 4:2:0, 2 4:2:2, 0 4:4:4 (as ivqm_picture) */
  int32_t bytes_per_sample;    /* 1 (8-bit only) or 2 */
  int32_t num_frames;
  const void* const* planes;   /* num_frames * 3 plane pointers, frame-major */
  const int32_t* strides;      /* num_frames * 3 strides in samples; NULL = tight */
} ivqm_sequence_desc;

/* One request: test frames i = 0..num_frames-1 against reference frames
 * first_frame + i of sequence seq_id (P0 = test, P1 = reference, as the
 * app's InputFile0 / InputFile1). Test frames have the sequence's format. */
typedef struct ivqm_job {
  uint64_t tag;                /* caller's id, returned with the result */
  int32_t seq_id;
  int32_t first_frame;
  int32_t num_frames;          /* 1..IVQM_MAX_JOB_FRAMES */
  uint32_t metrics;            /* IVQM_M_* set */
  int32_t planes_on_device;    /* 0: host planes, staged into the budget at admission; 1: device planes, read in place */
  int32_t reserved;
  const void* const* planes;   /* num_frames * 3 plane pointers; must stay valid until the result is polled */
  const int32_t* strides;      /* num_frames * 3 strides in samples; NULL = tight */
  /* Mask mode, the app's InputFileM (apps/QMIV/src/xAppQMIV.cpp:899-906, :930-1029): NULL, or num_frames
   * weight planes at luma resolution, one per test frame, in host or device memory as planes_on_device
   * says (staged or read in place like the test planes). The job's metrics are then the masked MSE,
   * PSNR, WS-PSNR and IV-PSNR of ivqm_bind_mask; structural-similarity metrics are rejected, as the app
   * rejects them (xAppQMIV.cpp:360). */
  const void* const* mask_planes;
  const int32_t* mask_strides; /* num_frames strides in samples; NULL = tight */
  int32_t mask_bit_depth;      /* BitDepthM, 8..14 */
  int32_t mask_bytes_per_sample; /* 1 (bit depth 8 only) or 2 */
} ivqm_job;

typedef struct ivqm_job_result {
  uint64_t tag;
  int32_t status;              /* IVQM_OK or a negative status */
  int32_t num_frames;
  double latency_us;           /* submit -> results ready on the host */
  double queue_us;             /* submit -> admitted (memory granted) */
  ivqm_frame_result frames[IVQM_MAX_JOB_FRAMES];
  /* Where the rest of latency_us went (queue_us + stage_us + wait_us + device_us <= latency_us): */
  double stage_us;             /* admitted -> test frames (and masks) on the device (0 for device planes) */
  double wait_us;              /* on the device -> the job's first frame put into a batch */
  double device_us;            /* first frame batched -> the job's last batch collected (launches, kernels, copies) */
} ivqm_job_result;

typedef struct ivqm_server_stats {
  uint64_t jobs_submitted, jobs_completed, jobs_failed;
  uint64_t pairs_completed, batches;
  uint64_t budget_bytes, budget_in_use, budget_peak;
  uint64_t resident_bytes;     /* reference sequences */
  double max_batch_fill;       /* largest batch seen, in pairs */
} ivqm_server_stats;

int ivqm_server_config_init(ivqm_server_config* cfg);
int ivqm_server_create(const ivqm_server_config* cfg, ivqm_server** out);
/* Waits for every submitted job, then frees the server. */
int ivqm_server_destroy(ivqm_server* s);

int ivqm_server_add_sequence(ivqm_server* s, const ivqm_sequence_desc* seq, int32_t* seq_id);
/* IVQM_EBUSY while a submitted job still refers to the sequence. */
int ivqm_server_remove_sequence(ivqm_server* s, int32_t seq_id);

/* Queues a job (the job description is copied; the planes are read later).
 * Invalid jobs are rejected here with IVQM_EINVAL. */
int ivqm_server_submit(ivqm_server* s, const ivqm_job* job);
/* Moves up to `max` finished results into out[]; waits up to wait_us
 * microseconds for the first one (0 = do not wait). *n = results written. */
int ivqm_server_poll(ivqm_server* s, ivqm_job_result* out, int32_t max, int32_t* n, int32_t wait_us);
/* Waits until every submitted job has finished (results stay pollable). */
int ivqm_server_drain(ivqm_server* s);
int ivqm_server_get_stats(ivqm_server* s, ivqm_server_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* IVQM_SERVE_H */
