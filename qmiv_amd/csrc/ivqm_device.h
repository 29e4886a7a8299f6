// Shared host/device definitions of the gfx950 metric kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ivqm {

// One picture as the kernels see it: three planes of 8-bit or 16-bit samples.
// Luma is at full resolution; chroma planes are subsampled by `cs` in both
// axes (cs = 1 for 4:2:0, 0 for 4:4:4). Strides are in samples.
struct PicDev {
  const void* p[3];
  int32_t stride[3];
  int32_t cs;
  int32_t bps;  // bytes per sample: 1 or 2
};

// One SSIM evaluation (a picture pair at one pyramid level).
struct SsimJob {
  PicDev a, b;
  int32_t w, h;         // luma-resolution size of this level
  int32_t nux, nuy;     // windows along x and y
  int32_t nchunk;       // 63-window column chunks
  int32_t calc_l;       // luminance term on/off
  int32_t part_off;     // offset (in doubles) of this job's partial block [3][nuy][nchunk]
  int32_t erp;          // weight rows by the ERP weights
};

constexpr int kMaxSsimJobs = 16;
struct SsimJobTable {
  SsimJob job[kMaxSsimJobs];
  int32_t njobs;
};

// Launchers (ivqm_kernels.hip). All are asynchronous on `st`.
hipError_t launch_stats(const PicDev& p0, const PicDev& p1, int w, int h, int64_t* row_sd, uint64_t* row_ssd,
                        uint32_t* zero_buf, int n_zero, hipStream_t st);
hipError_t launch_gcd_finalize(const int64_t* row_sd, const uint64_t* row_ssd, int w, int h, int cs, int bit_depth,
                               const int32_t maxdiff[3], int wrap32, int32_t* gcd_out, int64_t* sd_out,
                               uint64_t* ssd_out, hipStream_t st);
// rows: [2][3][h] uint32; scp: [2][3][h][scp_pitch] uint16 (picture 0 = SCP0, 1 = SCP1), may be null.
// best: optional [2][h][w] uint8 raster index of the winning candidate.
hipError_t launch_search(const PicDev& p0, const PicDev& p1, int w, int h, int bit_depth, const int32_t* gcd,
                         const int32_t cw[4], int R, bool narrow, uint32_t* rows, uint16_t* scp, int scp_pitch,
                         uint8_t* best, int tile_py, hipStream_t st);
hipError_t launch_ssim(const SsimJobTable& jobs, int max_nchunk, int max_bands, float c1n, float c2n, bool wide_sq,
                       double* partial, hipStream_t st);
// out: per job, per component: [job][3][2] = {sum, weight-sum}.
hipError_t launch_ssim_finalize(const SsimJobTable& jobs, const double* partial, const double* erp_w, int wnd,
                                int stride, double* out, hipStream_t st);
// dst: three planes (dw x dh, row pitch dst_pitch, plane size dst_pitch * dh) of the next pyramid level.
hipError_t launch_downsample(const PicDev& src, int w, int h, uint16_t* dst, int dst_pitch, hipStream_t st);
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t st);
hipError_t launch_upsample420(const PicDev& src, int w, int h, uint16_t* dst, int dst_stride, hipStream_t st);

}  // namespace ivqm
