// Shared host/device definitions of the gfx950 metric kernels.
//
// Every kernel works on a table of frame pairs (or SSIM / pyramid jobs) in
// device memory: one launch per phase covers all pairs of a batch, whatever
// their sizes and sample formats. A single pair is a batch of one.
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
  int32_t cs;   // chroma column shift: 1 for 4:2:0 and 4:2:2, 0 for 4:4:4
  int32_t bps;  // bytes per sample: 1 or 2
  int32_t csy;  // chroma row shift: 1 for 4:2:0, 0 for 4:2:2 and 4:4:4
};
// Chroma format of a picture: 0 4:4:4, 1 4:2:0, 2 4:2:2 (the ABI's ivqm_picture::chroma_420 values).
__host__ __device__ inline int chroma_format(const PicDev& p) { return p.cs ? (p.csy ? 1 : 2) : 0; }

// One frame pair of a batch and where its outputs live.
struct PairDev {
  PicDev p0, p1;
  int32_t w, h, maxv;
  int32_t mode;           // search variant: 0 reference form, 1 narrow key, 2 packed dot key, 3 reference form
                          // with 24-bit multiplies, 4 keyed 4:2:0 search (k_search420), 5 mode 3 with the
                          // default weights {4, 1, 1}, 6 reference score in expanded form (host-checked
                          // ranges)
  int32_t md[3];          // GCD clip limits round_f32(coef * Max)
  int32_t wrap32;         // x86-64-v4 int32 SD accumulation
  int32_t stat_blk0;      // first k_stats block of this pair
  int32_t tile0;          // first k_search tile of this pair
  int32_t gcd_given;      // 1: gcd[] already holds the caller's GCD
  int32_t scp_pitch;
  int64_t rowsd_off;      // [3][h] in row_sd (int64) and row_ssd (uint64)
  int64_t rows_off;       // [2][3][h] uint32
  int64_t scp_off;        // [2][3][h][scp_pitch] uint16, or -1
  int64_t best_off;       // [2][h][w] uint8, or -1
  int32_t res_idx;        // index of gcd[4], sd[3], ssd[3]
  int32_t gcd_in[4];      // the caller's GCD when gcd_given
  int32_t tile0b;         // first k_search420 tile of this pair (mode kSearchDot420)
  // Mask mode (weights 0 .. msk_max at luma resolution), or msk == null:
  const void* msk;
  int32_t msk_stride;     // in samples
  int32_t msk_bps;        // 1 or 2
  int32_t msk_max;        // 2^BitDepthM - 1
  int32_t mtile0;         // first k_search_masked tile of this pair
  int64_t aux_off;        // [4][h] int64: samples that differ per 4:4:4 row of each component; non-zero
                          // mask samples per row
  int64_t rows64_off;     // [2][3][h] uint64 masked search row sums (direction 0: P0 + GCD searched in P1)
  // Fused IV-SSIM: instead of the SCP pictures, the search writes the exact 4x4 sub-block moments of
  // (test picture, SCP) -- {S_a | S_b << 16, S_aa, S_bb, S_ab} as 4 x uint32 -- at mom_off (in uint16
  // units of the SCP arena, 256-byte aligned): [direction][component][mom_rows][mom_pitch], direction 0
  // pairing P0 with SCP1. Sub-blocks (sby, sbx) with sby < mom_rows, sbx < mom_cols are written; -1: off.
  // Keyed search modes (1, 2) only.
  int64_t mom_off;
  int32_t mom_rows, mom_cols, mom_pitch;
  // Masked search in keyed form: the key of a masked-out candidate's entry (host-checked to exceed every
  // considered candidate's key without wrapping), or 0 for the reference-form masked search.
  int32_t mkey;
  int32_t mdot;  // keyed masked search with the packed int16 {Y, Cb} key (1) or the 4:2:0 form with packed
                 // int16 {Cb, Cr} chroma words on 64 x 32 tiles (2) (host-checked ranges); 0: narrow key
  int32_t scp1_pitch;
  // Fused IV-MS-SSIM (k_search420 with the moments): the search also writes MS-SSIM level 1 of both SCP
  // pictures, [SCP0, SCP1][component][h >> 1][scp1_pitch] uint16 at scp1_off in the SCP arena
  // (direction d writes SCP 1 - d); -1: off.
  int64_t scp1_off;
};
// Largest bit depth of the fused IV-SSIM moments: 16 * 4095 fits the 16-bit halves of S_a | S_b << 16
// and 64 * 4095^2 the 32-bit window sums.
constexpr int kMomMaxBits = 12;

// One SSIM evaluation (a picture pair at one pyramid level).
struct SsimJobDev {
  PicDev a, b;
  int32_t w, h;
  int32_t nux, nuy, nchunk, nbg;  // windows, 63-window chunks, band groups (4 bands of NB rows)
  int32_t calc_l, erp, wsq;       // luminance term, ERP row weights, sum widths (0: 32-bit, 1: 64-bit
                                  // windows, 2: 64-bit throughout)
  int32_t blk0;                   // first block of this job (all 3 components)
  int64_t part_off;               // [3][ceil(nuy / grp)][nchunk] doubles in the partial buffer
  int64_t erp_off;                // offset of this height's weights, or -1
  int32_t out_idx;                // [3][2] doubles at out[6 * out_idx]
  float c1n, c2n;                 // C1 * n^2, C2 * n^2 of this job's bit depth (xSSIM.cpp:50-52)
  // Window grid and weights of k_ssim_gen (every mode other than block-averaged 8 / 4):
  int32_t mode, ext;              // window mode (xSSIM::eMode 0..4), border extension (eMrgExt 0..5)
  int32_t wnd, stride;            // window side (11 for regular modes) and stride
  int32_t gx0, gy0;               // first window position: regular = centre (5, or 0 with an extension),
                                  // block = top-left (0)
  int32_t erp_base;               // ERP weight row of window row j = erp_base + j * stride
  int32_t grp;                    // window rows per partial row: 1 (k_ssim_gen; k_ssim_finalize applies the
                                  // ERP weights) or 4 (k_ssim: the partial of rows 4i .. 4i + 3 is their
                                  // ERP-weighted sum, the weights applied in the kernel)
  double c1, c2;                  // xSSIM.cpp:50-52
  double kscale;                  // sums -> moments: 1 (float weights), 2^-18 (Q18), 1/area (averaged)
  const double* wt;               // wnd x wnd weights (row pitch wnd) followed by their transpose, or null
                                  // for averaged modes
  // Block-averaged 8 / 4 only: the search's 4x4 sub-block moments of (a, b) ([3][mom_rows][mom_pitch] x
  // 4 uint32, see PairDev::mom_off) replace reading the pictures; null: read a and b.
  const uint32_t* mom;
  int64_t mom_off;                // offset in the SCP arena (uint16 units) the host turns into `mom`
  int32_t mom_pitch, mom_rows;
  // Block-averaged 8 / 4 jobs: partials of a twin job over the same windows without the luminance term
  // (MS-SSIM level 0 beside SSIM, IV-MS-SSIM level 0 beside IV-SSIM), written by this job; -1: none.
  int64_t twin_part_off;
};

// One downsampling job of a pyramid level.
struct DownJob {
  PicDev src;
  int32_t w, h;      // source size
  int32_t pitch;     // destination pitch
  int32_t blk0;
  uint16_t* dst;     // three planes, plane size pitch * (h >> 1)
  int32_t ncomp;     // components written: 3, or 1 when the level takes a 4:2:0 source's chroma planes as they are
  int32_t pad;
};

// Launchers (ivqm_kernels.hip). All are asynchronous on `st`; tables live in
// device memory. `nblk` is the total block count of the flattened grid.
// Masked pairs take 3 x h stats rows (every component at 4:4:4), unmasked pairs h + 2 x (h >> cs).
// k_stats form of a pair: 0 unmasked 16-bit containers with samples below 2^12 (32-bit partial sums,
// a kernel of few registers), 1 everything else (8-bit samples, wider sums, masks). One launch per form
// present; `forms` is the bitmask of stats_form_bit over the batch's pairs with statistics.
__host__ __device__ inline int stats_form(const PairDev& p) {
  return (p.msk == nullptr && p.p0.bps == 2 && p.maxv < 4096) ? 0 : 1;
}
__host__ __device__ inline uint32_t stats_form_bit(const PairDev& p) { return 1u << stats_form(p); }
hipError_t launch_stats(const PairDev* pairs, int npairs, int nblk, int64_t* row_sd, uint64_t* row_ssd,
                        uint32_t* rows, int64_t* aux, uint64_t* rows64, uint32_t forms, hipStream_t st);
// mcount: per pair {non-zero mask samples, differing samples of Y, Cb, Cr} (masked pairs only).
hipError_t launch_gcd_finalize(const PairDev* pairs, int npairs, const int64_t* row_sd, const uint64_t* row_ssd,
                               const int64_t* aux, int32_t* gcd, int64_t* sd, uint64_t* ssd, int64_t* mcount,
                               hipStream_t st);
// Masked search of both directions: tiles of 64 x 16 pixels per direction (ivqm_masked.hip). forms: bit f
// set when a pair of form f (0 reference form, 1 keyed narrow, 2 keyed dot) is in the batch. Pairs with
// scp_off >= 0 also write their SCP pictures into scp.
constexpr int kMaskTileW = 64, kMaskTileH = 16;
constexpr int kMask420TileH = 32;  // tile height of the 4:2:0 keyed masked search (PairDev::mdot == 2)
hipError_t launch_search_masked(const PairDev* pairs, int npairs, int ntiles, int R, const int32_t cw[4],
                                const int32_t* gcd, uint64_t* rows64, uint16_t* scp, uint32_t forms, hipStream_t st);
// forms: one k_search kernel per form of the search arithmetic present among the batch's k_search pairs,
// each over all tiles: bit 0 keyed (modes 1, 2), bit 1 reference forms (modes 0, 3, 5), bit 2 mode 6.
constexpr int kSearchQuadMode = 6;
__host__ __device__ inline uint32_t search_form_bit(int mode) {
  return mode == 1 || mode == 2 ? 1u : (mode == kSearchQuadMode ? 4u : 2u);
}
hipError_t launch_search(const PairDev* pairs, int npairs, int ntiles, int R, const int32_t cw[4],
                         const int32_t* gcd, uint32_t* rows, uint16_t* scp, uint8_t* best, uint32_t forms,
                         hipStream_t st);
// forms: bit 0 when jobs read pictures with 32-bit sums (wsq 0), bit 2 when jobs read pictures with wider
// sums, bit 1 when jobs read fused moments (one launch each).
hipError_t launch_ssim(const SsimJobDev* jobs, int njobs, int nblk, int nb, const double* erp_w, double* partial,
                       uint32_t forms, hipStream_t st);
// Any window mode / size / stride / border extension: one wave per 64 window columns and
// kSsimGenRows window rows, 4 waves per block. Stride-4 windows of 8, 11 or 16 samples without a border
// extension (regular: window 11; block: BlockGaussianInt and BlockAveraged other than 8 / 4) run the
// rolling-row kernel instead: 64 - (N - 1) / 4 window columns and kSsimRollRows window rows per wave.
constexpr int kSsimGenRows = 8, kSsimRollRows = 16;
struct SsimGenShape {
  int cols, rows;  // window columns and window rows of one wave
  bool roll;
};
__host__ __device__ inline SsimGenShape ssim_gen_shape(int mode, int wnd, int ext, int stride) {
  const bool regular = mode <= 2;
  const int n = regular ? 11 : wnd;
  if (stride == 4 && (regular ? ext == 0 : (n == 8 || n == 16))) return SsimGenShape{64 - (n - 1) / 4, kSsimRollRows, true};
  return SsimGenShape{64, kSsimGenRows, false};
}
hipError_t launch_ssim_gen(const SsimJobDev* jobs, int njobs, int nblk, int mode, int wnd, int ext, int stride,
                           double* partial, hipStream_t st);
hipError_t launch_ssim_finalize(const SsimJobDev* jobs, int njobs, const double* partial, const double* erp_w,
                                double* out, hipStream_t st);
hipError_t launch_downsample(const DownJob* jobs, int njobs, int nblk, hipStream_t st);
// Standalone pixel operations on one picture.
hipError_t launch_downsample_one(const PicDev& src, int w, int h, uint16_t* dst, int dst_pitch, hipStream_t st);
hipError_t launch_upsample420(const PicDev& src, int w, int h, uint16_t* dst, int dst_stride, hipStream_t st);
// direction 0: YCbCr -> RGB, 1: RGB -> YCbCr; space 0 BT.601, 1 BT.709, 2 SMPTE240M, 3 BT.2020.
hipError_t launch_colour(const PicDev& src, int w, int h, int bit_depth, int direction, int space, uint16_t* dst,
                         int dst_stride, hipStream_t st);

constexpr int kSearchTileW = 64;
// k_search: waves of 4 pixel rows on 64 x 16 tiles for every search range and variant but the keyed
// 4:2:0 one; k_search420 (ivqm_search420.hip): 64 x TH tiles, TH = 16 or 32 (tuning).
constexpr int kSearchRowsPerWave = 4;
constexpr int kSearchTileH = 4 * kSearchRowsPerWave;
constexpr int kSearchDot420 = 4;  // PairDev::mode of the keyed 4:2:0 search
// Tile height of k_search420 from a tuning value: 16 or 32 (4 and 8, the earlier per-wave row counts, are
// taken as 16 and 32); anything else keeps `current`.
inline int search420_tile_rows(int v, int current) {
  if (v == 4 || v == 16) return 16;
  if (v == 8 || v == 32) return 32;
  return current;
}
// Key scale Sk = 2^search420_key_bits(R) and sample scale Sy = 2^search420_sample_bits(R) of k_search420.
int search420_key_bits(int R);
int search420_sample_bits(int R);
hipError_t launch_search420(const PairDev* pairs, int npairs, int ntiles, int R, int th, const int32_t cw[4],
                            const int32_t* gcd, uint32_t* rows, uint16_t* scp, uint8_t* best, hipStream_t st);
constexpr int kDownBlkW = 512, kDownBlkH = 4;  // k_downsample: outputs per block row (8 per lane), rows
// MI355X: 8 XCDs, each with its own L2; consecutive workgroups go to consecutive XCDs.
constexpr int kXcds = 8;

}  // namespace ivqm
