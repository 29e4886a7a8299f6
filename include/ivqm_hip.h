/*
 * ivqm_hip.h - C ABI of the MI355X (gfx950) implementation of QMIV's metric
 * hot path (IV-PSNR, IV-SSIM, MS-SSIM, WS-PSNR, PSNR over planar pictures).
 *
 * Plain C types only: pointers, sizes, status codes. Every function returns an
 * int status: 0 on success, a negative IVQM_E* value on failure; nothing
 * aborts or throws across this boundary. A context is not thread-safe: use one
 * host thread per context. Device work runs on the context's HIP stream.
 *
 * Each entry point names the reference interface it replaces
 * (jstankowski/QMIV, file:line). The reference's processor objects are
 * configured by setters and then called once per frame with xPicP / xPicI
 * pictures; here the configuration is one struct and pictures are bound to the
 * context's two slots (0 = P0 = InputFile0 = "Tst", 1 = P1 = InputFile1 = "Ref").
 */
#ifndef IVQM_HIP_H
#define IVQM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IVQM_OK 0
#define IVQM_EINVAL (-22)   /* invalid argument / unsupported configuration */
#define IVQM_ENOMEM (-12)   /* device or host allocation failed, or memory budget exceeded */
#define IVQM_ESTATE (-71)   /* call out of order (e.g. no picture bound) */
#define IVQM_EBUSY (-16)    /* serving queue full */
#define IVQM_EHIP (-1000)   /* HIP runtime error; IVQM_EHIP - hipError_t value */

/* Metric bits for ivqm_frame() and the batch / serving interfaces. */
#define IVQM_M_PSNR (1u << 0)
#define IVQM_M_WSPSNR (1u << 1)
#define IVQM_M_IVPSNR (1u << 2)
#define IVQM_M_SSIM (1u << 3)
#define IVQM_M_MSSSIM (1u << 4)
#define IVQM_M_IVSSIM (1u << 5)
#define IVQM_M_IVMSSSIM (1u << 6)
#define IVQM_M_MSE (1u << 7)
#define IVQM_M_ALL 0xFFu

typedef struct ivqm_config {
  int32_t width, height;          /* luma size (even for 4:2:0) */
  int32_t bit_depth;              /* 8..16 */
  int32_t search_range;           /* R, 1..4; xCorrespPixelShiftPrms.h:20 (default 2) */
  int32_t cmp_weights_search[4];  /* xCorrespPixelShiftPrms.h:21 (default 4,1,1,0) */
  int32_t cmp_weights_average[4]; /* default 4,1,1,0 */
  float unnoticeable_coef[4];     /* xGlobClrDiff.h:20 (default 0.01,0.01,0.01,0) */
  int32_t erp;                    /* equirectangular weighting (xWeightedSpherically.h:21) */
  int32_t lon_range_deg, lat_range_deg; /* default 360, 180 */
  int32_t ssim_window, ssim_stride;     /* SSIM window and stride, default 8 and 4 (xSSIM.h:35-37); regular
                                           modes always use 11 x 11 windows (xSSIM.h:95) and ignore
                                           ssim_window; stride 1..window */
  int32_t fake_vals_for_exact;    /* PSNR of an exact component = PSNR(SSD=1) (xPSNR.cpp:147, xAppQMIV.cpp:974-981) */
  int32_t gcd_sd_wrap32;          /* reproduce the x86-64-v4 <=14-bit int32 SD accumulation (xDistortionAVX512.cpp:14-73) */
  int32_t search_tile_rows;       /* search tile height of 4:2:0 pairs: 16 or 32 rows (4 / 8 taken as
                                     16 / 32; 0 = 32); 4:4:4 pairs use 64 x 16 tiles */
  int32_t ssim_wide_sums;         /* 64-bit squared sums in SSIM at any bit depth (always on above 14-bit) */
  int32_t search_mode_cap;        /* search arithmetic: 0 = fastest exact variant the ranges allow,
                                     1 = reference-form int32 arithmetic only, 2 = no packed-int16 variant
                                     (all variants give identical results) */
  int32_t ssim_mode;              /* xSSIM::eMode (xSSIM.h:24-31): 0 RegularGaussianFlt, 1 RegularGaussianInt,
                                     2 RegularAveraged, 3 BlockGaussianInt (window 8 or 16),
                                     4 BlockAveraged (default; window 1..64) */
  int32_t ssim_border;            /* eMrgExt (xCommonDefCORE.h:322-331), regular modes only: 0 None (default),
                                     1 Nearest, 2 Reflect, 3 Mirror, 4 Constant (65535), 5 Zero */
  int32_t reserved[3];
} ivqm_config;

/* One picture: three planes of 8-bit (bytes_per_sample = 1) or 16-bit samples.
 * chroma_420: the chroma format, as the reference reads it (libs/pmbb-core/src/xSeq.cpp:231-301):
 * 1 = 4:2:0, chroma planes (width/2) x (height/2), even sizes; 2 = 4:2:2, chroma planes
 * (width/2) x height, even width; 0 = 4:4:4, full size. Subsampled chroma is read as its
 * nearest replication at 4:4:4, which is what the reference's upsampling stores.
 * Strides are in samples. */
typedef struct ivqm_picture {
  const void* plane[3];
  int32_t stride[3];
  int32_t chroma_420;
  int32_t bytes_per_sample;
} ivqm_picture;

/* Everything one frame pair produces. Values follow the reference's types:
 * per-component arrays hold Y, Cb, Cr and a 4th lane as the reference's
 * int32V4 / flt64V4 do. "pic" values are the 4:1:1 (cmp_weights_average)
 * means the app prints (xMetricQMIV.cpp:91-97). */
typedef struct ivqm_frame_result {
  uint32_t metrics;         /* bits computed */
  int32_t gcd_r2t[4];       /* xGlobClrDiffProc::CalcGlobalColorDiff(P0, P1) */
  uint64_t ssd[4];          /* sum((P0 - P1)^2) per component at 4:4:4 */
  int32_t exact[4];         /* component identical in both pictures */
  double mse[4], psnr[4], wspsnr[4];
  double psnr_pic, wspsnr_pic;
  double ivpsnr, ivpsnr_r2t, ivpsnr_t2r;
  double ssim[4], msssim[4];
  double ssim_pic, msssim_pic;
  double ivssim, ivssim_r2t, ivssim_t2r;
  double ivmsssim, ivmsssim_r2t, ivmsssim_t2r;
} ivqm_frame_result;

typedef struct ivqm_ctx ivqm_ctx;

/* ---- configuration and context ------------------------------------------ */
int ivqm_config_init(ivqm_config* cfg);            /* fills the reference defaults */
int ivqm_create(const ivqm_config* cfg, ivqm_ctx** out);
int ivqm_destroy(ivqm_ctx* ctx);
/* Runs the context's work on `hip_stream` (a hipStream_t; NULL = own stream). */
int ivqm_set_stream(ivqm_ctx* ctx, void* hip_stream);
const char* ivqm_strerror(int status);
/* Message of the last failing call on this thread. */
const char* ivqm_last_error(void);

/* ---- pictures ------------------------------------------------------------ */
/* Copies a host picture into the context's device buffer of `slot` (0/1).
 * Replaces xSeqPic::readFrame's unpack into xPicP (libs/pmbb-core/src/xSeq.cpp:45-63,
 * :231-301); 4:2:0 stays 4:2:0 on the device and is upsampled on the fly. */
int ivqm_upload(ivqm_ctx* ctx, int32_t slot, const ivqm_picture* host_pic);
/* Binds device-resident planes to `slot` without copying (pointers must stay
 * valid until rebound). */
int ivqm_bind(ivqm_ctx* ctx, int32_t slot, const ivqm_picture* dev_pic);

/* ---- per-metric calls (blocking; results as the reference returns them) --- */
/* xGlobClrDiffProc::CalcGlobalColorDiff(P0, P1): libs/pmbb-ivqm/src/xGlobClrDiff.h:41. */
int ivqm_gcd(ivqm_ctx* ctx, int32_t gcd_r2t[4]);
/* xPSNR::calcPicSSD / calcPicMSE / calcPicPSNR(P0, P1): libs/pmbb-ivqm/src/xPSNR.h:30-32. */
int ivqm_ssd(ivqm_ctx* ctx, uint64_t ssd[4]);
int ivqm_mse(ivqm_ctx* ctx, double mse[4]);
int ivqm_psnr(ivqm_ctx* ctx, double psnr[4]);
/* xWSPSNR::calcPicWSPSNR(P0, P1): libs/pmbb-ivqm/src/xWSPSNR.h:22. */
int ivqm_wspsnr(ivqm_ctx* ctx, double wspsnr[4]);
/* xIVPSNR::calcPicIVPSNR(I0, I1, GCD_R2T): libs/pmbb-ivqm/src/xIVPSNR.h:36.
 * gcd_r2t == NULL computes the GCD first (xIVPSNR.h:33). r2t/t2r may be NULL. */
int ivqm_ivpsnr(ivqm_ctx* ctx, const int32_t* gcd_r2t, double* ivpsnr, double* r2t, double* t2r);
/* xShftCompPicProc::GenShftCompPics(SCP1, SCP0, I1, I0, GCD_R2T): libs/pmbb-ivqm/src/xShftCompPic.h:44-47.
 * The pictures stay on the device for ivqm_ivssim / ivqm_ivmsssim; download with ivqm_get_scp. */
int ivqm_gen_scp(ivqm_ctx* ctx, const int32_t* gcd_r2t);
/* xSSIM::calcPicSSIM / calcPicMSSSIM(P0, P1): libs/pmbb-ivqm/src/xSSIM.h:83-84. */
int ivqm_ssim(ivqm_ctx* ctx, double ssim[4]);
int ivqm_msssim(ivqm_ctx* ctx, double msssim[4]);
/* xIVSSIM::calcPicIVSSIM / calcPicIVMSSSIM(P0, P1, SCP0, SCP1): libs/pmbb-ivqm/src/xIVSSIM.h:39-40.
 * Uses the SCP pictures of the last ivqm_gen_scp. */
int ivqm_ivssim(ivqm_ctx* ctx, double* ivssim, double* r2t, double* t2r);
/* Loads shift-compensated picture idx (0 = SCP0, 1 = SCP1) from host 4:4:4 planes,
 * for callers that hold their own SCP pictures (xIVSSIM.h:39 takes them as arguments). */
int ivqm_upload_scp(ivqm_ctx* ctx, int32_t idx, const ivqm_picture* host_pic);
int ivqm_ivmsssim(ivqm_ctx* ctx, double* ivmsssim, double* r2t, double* t2r);

/* ---- mask mode --------------------------------------------------------------
 * A weight mask, as the QMIV app's InputFileM (apps/QMIV/src/xAppQMIV.cpp:69-71,
 * :284-288): one plane at luma resolution, samples 0 .. 2^bit_depth - 1
 * (bit_depth 8..14 as BitDepthM; 1 byte per sample needs bit_depth 8); the
 * stride is in samples. */
typedef struct ivqm_mask {
  const void* plane;
  int32_t stride;
  int32_t bytes_per_sample;
  int32_t bit_depth;
  int32_t reserved;
} ivqm_mask;
/* Binds a device-resident mask to the context (NULL unbinds). With a mask
 * bound, ivqm_gcd / ivqm_ssd / ivqm_mse / ivqm_psnr / ivqm_wspsnr / ivqm_ivpsnr
 * and ivqm_frame run the masked variants the app uses
 * (xGlobClrDiffProc::CalcGlobalColorDiffM, xGlobClrDiff.h:42; xPSNR::calcPicMSEM /
 * calcPicPSNRM, xPSNR.h:37-39; xWSPSNR::calcPicWSPSNRM, xWSPSNR.h:23;
 * xIVPSNRM::calcPicIVPSNRM, xIVPSNR.h:55-56), ssd[] holds the masked SSD,
 * ivqm_gen_scp runs xShftCompPicProc::GenShftCompPicsM (xShftCompPic.h:48-50:
 * pixels whose mask sample is 0 are not searched and hold 0 in the SCP
 * pictures; the reference leaves them unwritten), and the structural-similarity
 * metrics return IVQM_EINVAL (the app rejects that combination,
 * xAppQMIV.cpp:360). A mask without any non-zero sample makes the metric calls
 * return IVQM_EINVAL. */
int ivqm_bind_mask(ivqm_ctx* ctx, const ivqm_mask* dev_mask);
/* Copies a host mask into the context's own buffer, then binds it. */
int ivqm_upload_mask(ivqm_ctx* ctx, const ivqm_mask* host_mask);

/* ---- fused frame pass ----------------------------------------------------- */
/* All requested metrics of the bound pair in one device pass (stats, GCD,
 * search with fused SCP, SSIM jobs, FP64 totals) and one result download;
 * the per-frame stage sequence of xAppQMIV::processAllFrames
 * (apps/QMIV/src/xAppQMIV.cpp:761-827). */
int ivqm_frame(ivqm_ctx* ctx, uint32_t metrics, ivqm_frame_result* out);
/* Asynchronous form: enqueue into result ring slot `ring` (0..7), collect later. */
int ivqm_frame_enqueue(ivqm_ctx* ctx, uint32_t metrics, int32_t ring);
int ivqm_frame_collect(ivqm_ctx* ctx, int32_t ring, ivqm_frame_result* out);

/* ---- debug outputs (per-row / per-pixel integers for parity checks) ------- */
/* Last search's per-row distances of direction dir (0 = R2T, 1 = T2R) as the
 * reference's uint64V4 rows: rows[y*4 + c] (xIVPSNR.cpp:126-131). */
int ivqm_get_rows(ivqm_ctx* ctx, int32_t dir, uint64_t* rows);
/* Shift-compensated picture idx (0 = SCP0 from P0, 1 = SCP1 from P1): three
 * width x height uint16 planes, tightly packed. */
int ivqm_get_scp(ivqm_ctx* ctx, int32_t idx, uint16_t* const planes[3]);
/* Enables the per-pixel winner map (raster candidate index) on the next searches. */
int ivqm_enable_best_map(ivqm_ctx* ctx, int32_t enable);
int ivqm_get_best_map(ivqm_ctx* ctx, int32_t dir, uint8_t* map);

/* ---- pixel operations on device buffers (async on `hip_stream`) ----------- */
/* 4:2:0 -> 4:4:4 nearest upsampling (xPixelOpsSTD.cpp:30-49 via xSeq.cpp:231-301):
 * dst = three width x height uint16 planes with row stride dst_stride. */
int ivqm_upsample420(const ivqm_picture* src, int32_t width, int32_t height, uint16_t* dst, int32_t dst_stride,
                     void* hip_stream);
/* Colour-space conversion into three width x height uint16 4:4:4 planes (plane k at
 * dst + k * dst_stride * height), as the app's preprocessFrames does it
 * (apps/QMIV/src/xAppQMIV.cpp:853-876) with xColorSpace::ConvertYCbCr2RGB /
 * ConvertRGB2YCbCr (libs/pmbb-core/src/xColorSpace.h, fixed-point kernels).
 * direction 0: YCbCr -> RGB (src 4:2:0 or 4:4:4), 1: RGB -> YCbCr (src planes
 * R, G, B); space: 0 BT.601 / SMPTE170M, 1 BT.709, 2 SMPTE240M, 3 BT.2020. */
int ivqm_convert_colour(const ivqm_picture* src, int32_t width, int32_t height, int32_t bit_depth,
                        int32_t direction, int32_t space, uint16_t* dst, int32_t dst_stride, void* hip_stream);
/* 2x2 rounded-mean downsampling (xPixelOpsSTD.cpp:50-67, xSSIM.cpp:418-424):
 * dst = three (width/2) x (height/2) uint16 planes, tightly packed. */
int ivqm_downsample(const ivqm_picture* src, int32_t width, int32_t height, uint16_t* dst, void* hip_stream);

#ifdef __cplusplus
}
#endif

#endif /* IVQM_HIP_H */
