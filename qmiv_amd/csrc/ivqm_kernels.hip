// gfx950 (CDNA4) kernels of the QMIV metric hot path.
//
// Kernels:
//   k_stats          per plane row: sum(P1 - P0) and sum((P0 - P1)^2)
//                    (GCD input, PSNR / WS-PSNR row SSDs); zeroes the search
//                    row accumulators as a side job.
//   k_gcd_finalize   frame totals + global colour difference (FP64 mean,
//                    round half away, clip), written to device memory.
//   k_search         corresponding-pixel-shift search, both directions in one
//                    grid, LDS-staged reference tile with a +-R halo, per-row
//                    int32-wrapping distance sums and the shift-compensated
//                    pictures as a fused epilogue.
//   k_ssim           block-averaged 8x8 / stride-4 SSIM: exact integer 4x4
//                    sub-block moments, 2x2 combination, FP32 ratios, FP64
//                    per-window-row partial sums; many (picture pair, level)
//                    jobs per launch.
//   k_ssim_finalize  FP64 frame totals per job and component (ERP-weighted
//                    for IV-SSIM).
//   k_downsample     2x2 rounded-mean pyramid level (MS-SSIM).
//   k_upsample420    4:2:0 -> 4:4:4 nearest replication.
//
// 4:2:0 input is never materialised at 4:4:4 by the metric kernels: a chroma
// sample at luma-grid (x, y) is read at (x >> 1, y >> 1), which is exactly the
// value 2x2 replication (libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49) would have
// stored there. The reference's Nearest margin (xMarginOps.h:40-73) is
// reproduced by clamping coordinates.
#include "ivqm_device.h"

#include <limits.h>
#include <type_traits>

namespace ivqm {

namespace {

template <typename T>
__device__ __forceinline__ int ldpix(const PicDev& P, int c, int x, int y) {
  const int s = c ? P.cs : 0;
  const T* p = (const T*)P.p[c];
  return (int)p[(size_t)(y >> s) * (size_t)P.stride[c] + (x >> s)];
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
  return v;
}

// ---------------------------------------------------------------------------
// k_stats: one 256-thread block per plane row.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(256) void k_stats(PicDev P0, PicDev P1, int W, int H, int64_t* __restrict__ row_sd,
                                               uint64_t* __restrict__ row_ssd, uint32_t* __restrict__ zero_buf,
                                               int n_zero) {
  const int tid = threadIdx.x;
  const int b = blockIdx.x;
  {
    const int z = b * 256 + tid;
    if (z < n_zero) zero_buf[z] = 0u;
  }
  const int Hc = H >> P0.cs, Wc = W >> P0.cs;
  int c, y, wp;
  if (b < H) { c = 0; y = b; wp = W; }
  else if (b < H + Hc) { c = 1; y = b - H; wp = Wc; }
  else { c = 2; y = b - H - Hc; wp = Wc; }
  const T* a = (const T*)P0.p[c] + (size_t)y * P0.stride[c];
  const T* r = (const T*)P1.p[c] + (size_t)y * P1.stride[c];
  int64_t sd = 0;
  uint64_t ssd = 0;
  for (int x = tid; x < wp; x += 256) {
    const int d = (int)r[x] - (int)a[x];
    sd += d;
    ssd += (uint64_t)((uint32_t)d * (uint32_t)d);
  }
  __shared__ int64_t s_sd[4];
  __shared__ uint64_t s_ssd[4];
#pragma unroll
  for (int m = 32; m >= 1; m >>= 1) {
    sd += __shfl_xor(sd, m, 64);
    ssd += __shfl_xor(ssd, m, 64);
  }
  if ((tid & 63) == 0) { s_sd[tid >> 6] = sd; s_ssd[tid >> 6] = ssd; }
  __syncthreads();
  if (tid == 0) {
    row_sd[(size_t)c * H + y] = s_sd[0] + s_sd[1] + s_sd[2] + s_sd[3];
    row_ssd[(size_t)c * H + y] = s_ssd[0] + s_ssd[1] + s_ssd[2] + s_ssd[3];
  }
}

// ---------------------------------------------------------------------------
// k_gcd_finalize: one block. libs/pmbb-ivqm/src/xGlobClrDiff.cpp:23-56 with
// the app's argument order (sum of P1 - P0), chroma sums of 4:2:0 input x4.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gcd_finalize(const int64_t* __restrict__ row_sd,
                                                      const uint64_t* __restrict__ row_ssd, int W, int H, int cs,
                                                      int md0, int md1, int md2, int wrap32, int32_t* __restrict__ gcd,
                                                      int64_t* __restrict__ sd_out, uint64_t* __restrict__ ssd_out) {
  __shared__ int64_t s_sd[4];
  __shared__ uint64_t s_ssd[4];
  const int tid = threadIdx.x;
  for (int c = 0; c < 3; c++) {
    const int rows = c ? (H >> cs) : H;
    int64_t sd = 0;
    uint64_t ssd = 0;
    for (int y = tid; y < rows; y += 256) {
      sd += row_sd[(size_t)c * H + y];
      ssd += row_ssd[(size_t)c * H + y];
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
      sd += __shfl_xor(sd, m, 64);
      ssd += __shfl_xor(ssd, m, 64);
    }
    if ((tid & 63) == 0) { s_sd[tid >> 6] = sd; s_ssd[tid >> 6] = ssd; }
    __syncthreads();
    if (tid == 0) {
      int64_t t = s_sd[0] + s_sd[1] + s_sd[2] + s_sd[3];
      uint64_t q = s_ssd[0] + s_ssd[1] + s_ssd[2] + s_ssd[3];
      if (c && cs) { t *= 4; q *= 4; }
      sd_out[c] = t;
      ssd_out[c] = q;
      if (wrap32) t = (int64_t)(int32_t)(uint32_t)(uint64_t)t;
      const double avg = (double)t / (double)((int64_t)W * (int64_t)H);
      const double rnd = avg + copysign(0.5, avg);
      int g = (int)rnd;
      const int md = c == 0 ? md0 : (c == 1 ? md1 : md2);
      g = g < -md ? -md : (g > md ? md : g);
      gcd[c] = g;
      if (c == 2) gcd[3] = 0;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// k_search: corresponding-pixel-shift search.
//
// Reference semantics (x86 default build): libs/pmbb-ivqm/src/
// xCorrespPixelShiftSSE.cpp:18-70 and :75-127 -- for every test pixel
// t = Tst + gcs, the first strict minimum in dy-outer raster order of
// E = sum_c w_c (t_c - r_c)^2 over the (2R+1)^2 reference candidates, int32
// arithmetic that wraps; per-row int32-wrapping sums of the winner's
// per-component squared differences; SCP = clip(r_best - gcs, 0, Max).
//
// Block = 4 waves; the block tile is 64 x (4*PY) pixels of one direction.
// Each lane owns one column and PY consecutive rows. The reference tile
// (64 + 2R) x (4*PY + 2R) is staged once in LDS as int4 {Q, Y, Cb, Cr}.
// Walking the tile rows top to bottom visits every pixel's candidates in
// dy-outer, dx-inner raster order, so "first strict minimum" holds for the
// exact (wide) variant, while each loaded LDS row serves up to 2R+1 pixels.
//
// NARROW variant (host checks the ranges): E = T + E', T = sum w t^2 constant
// per pixel and E' = Q(r) - 2 sum_c w_c t_c r_c with Q(r) = sum_c w_c r_c^2
// precomputed per reference pixel. Nothing wraps in this range, so
// argmin E == argmin E'; the candidate's raster index is packed into the low
// KB bits of key = 2^KB * E' + idx, which makes a plain integer min equal to
// "first strict minimum in raster order". Per candidate: three v_mad_i32_i24
// and one min.
// ---------------------------------------------------------------------------
template <int R> struct KeyBits { static constexpr int v = (2 * R + 1) * (2 * R + 1) <= 16 ? 4 : (2 * R + 1) * (2 * R + 1) <= 32 ? 5 : (2 * R + 1) * (2 * R + 1) <= 64 ? 6 : 7; };

template <typename T, int R, int PY, bool NARROW>
__global__ __launch_bounds__(256) void k_search(PicDev P0, PicDev P1, int W, int H, int maxv,
                                                const int32_t* __restrict__ gcd, int wY, int wCb, int wCr,
                                                uint32_t* __restrict__ rows, uint16_t* __restrict__ scp,
                                                int scp_pitch, uint8_t* __restrict__ best) {
  constexpr int TW = 64, TH = 4 * PY, D = 2 * R + 1;
  constexpr int LW = TW + 2 * R, LH = TH + 2 * R;
  constexpr int KB = KeyBits<R>::v, S = 1 << KB;
  __shared__ int4 lds[LH][LW];

  const int dir = blockIdx.z;  // 0: R2T (Tst = P0, Ref = P1, +gcd); 1: T2R (Tst = P1, Ref = P0, -gcd)
  const PicDev Tst = dir ? P1 : P0;
  const PicDev Ref = dir ? P0 : P1;
  const int g0 = dir ? -gcd[0] : gcd[0];
  const int g1 = dir ? -gcd[1] : gcd[1];
  const int g2 = dir ? -gcd[2] : gcd[2];
  const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;

  for (int i = threadIdx.x; i < LW * LH; i += 256) {
    const int ly = i / LW, lx = i - ly * LW;
    const int x = clampi(x0 + lx - R, 0, W - 1), y = clampi(y0 + ly - R, 0, H - 1);
    const int rY = ldpix<T>(Ref, 0, x, y), rCb = ldpix<T>(Ref, 1, x, y), rCr = ldpix<T>(Ref, 2, x, y);
    int q = 0;
    if (NARROW) q = (wY * __mul24(rY, rY) + wCb * __mul24(rCb, rCb) + wCr * __mul24(rCr, rCr)) << KB;
    lds[ly][lx] = make_int4(q, rY, rCb, rCr);
  }
  __syncthreads();

  const int tx = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int x = x0 + tx, xc = min(x, W - 1);
  const int yb = y0 + wv * PY;

  int t0[PY], t1[PY], t2[PY], bk[PY], bi[PY];
#pragma unroll
  for (int i = 0; i < PY; i++) {
    const int y = min(yb + i, H - 1);
    t0[i] = ldpix<T>(Tst, 0, xc, y) + g0;
    t1[i] = ldpix<T>(Tst, 1, xc, y) + g1;
    t2[i] = ldpix<T>(Tst, 2, xc, y) + g2;
    bk[i] = INT_MAX;
    bi[i] = 0;
  }

  if (NARROW) {
    int m0[PY], m1[PY], m2[PY];
#pragma unroll
    for (int i = 0; i < PY; i++) {
      m0[i] = -(2 * S) * wY * t0[i];
      m1[i] = -(2 * S) * wCb * t1[i];
      m2[i] = -(2 * S) * wCr * t2[i];
    }
#pragma unroll 1
    for (int ry = 0; ry < PY + 2 * R; ry++) {
      int4 v[D];
#pragma unroll
      for (int dx = 0; dx < D; dx++) {
        v[dx] = lds[wv * PY + ry][tx + dx];
        v[dx].x += dx;
      }
#pragma unroll
      for (int i = 0; i < PY; i++) {
        const int dy = ry - i;
        if ((unsigned)dy >= (unsigned)D) continue;
        int rm = INT_MAX;
#pragma unroll
        for (int dx = 0; dx < D; dx++) {
          int acc = __mul24(m0[i], v[dx].y) + v[dx].x;
          acc = __mul24(m1[i], v[dx].z) + acc;
          acc = __mul24(m2[i], v[dx].w) + acc;
          rm = min(rm, acc);
        }
        bk[i] = min(bk[i], rm + dy * D);
      }
    }
#pragma unroll
    for (int i = 0; i < PY; i++) bi[i] = bk[i] & (S - 1);
  } else {
#pragma unroll 1
    for (int ry = 0; ry < PY + 2 * R; ry++) {
      int4 v[D];
#pragma unroll
      for (int dx = 0; dx < D; dx++) v[dx] = lds[wv * PY + ry][tx + dx];
#pragma unroll
      for (int i = 0; i < PY; i++) {
        const int dy = ry - i;
        if ((unsigned)dy >= (unsigned)D) continue;
#pragma unroll
        for (int dx = 0; dx < D; dx++) {
          const uint32_t d0 = (uint32_t)__mul24(t0[i] - v[dx].y, t0[i] - v[dx].y);
          const uint32_t d1 = (uint32_t)__mul24(t1[i] - v[dx].z, t1[i] - v[dx].z);
          const uint32_t d2 = (uint32_t)__mul24(t2[i] - v[dx].w, t2[i] - v[dx].w);
          const int e = (int)(d0 * (uint32_t)wY + d1 * (uint32_t)wCb + d2 * (uint32_t)wCr);
          if (e < bk[i]) { bk[i] = e; bi[i] = dy * D + dx; }
        }
      }
    }
  }

  // Epilogue: winner's squared differences, row sums, SCP, debug map.
  // dir 0 writes SCP1 (picture 1), dir 1 writes SCP0 (picture 0); planes [c][H][scp_pitch].
  const size_t plane_sz = (size_t)scp_pitch * H;
  uint16_t* scp_pic = scp ? scp + (size_t)(1 - dir) * 3 * plane_sz : nullptr;
#pragma unroll
  for (int i = 0; i < PY; i++) {
    const int y = yb + i;
    const bool ok = (x < W) && (y < H);
    const int ddy = bi[i] / D, ddx = bi[i] - ddy * D;
    const int4 r = lds[wv * PY + i + ddy][tx + ddx];
    const int e0 = t0[i] - r.y, e1 = t1[i] - r.z, e2 = t2[i] - r.w;
    uint32_t q0 = ok ? (uint32_t)__mul24(e0, e0) : 0u;
    uint32_t q1 = ok ? (uint32_t)__mul24(e1, e1) : 0u;
    uint32_t q2 = ok ? (uint32_t)__mul24(e2, e2) : 0u;
    q0 = wave_sum_u32(q0);
    q1 = wave_sum_u32(q1);
    q2 = wave_sum_u32(q2);
    if (tx == 0 && y < H) {
      atomicAdd(&rows[((size_t)dir * 3 + 0) * H + y], q0);
      atomicAdd(&rows[((size_t)dir * 3 + 1) * H + y], q1);
      atomicAdd(&rows[((size_t)dir * 3 + 2) * H + y], q2);
    }
    if (ok) {
      if (scp_pic) {
        const size_t o = (size_t)y * scp_pitch + x;
        scp_pic[o] = (uint16_t)clampi(r.y - g0, 0, maxv);
        scp_pic[plane_sz + o] = (uint16_t)clampi(r.z - g1, 0, maxv);
        scp_pic[2 * plane_sz + o] = (uint16_t)clampi(r.w - g2, 0, maxv);
      }
      if (best) best[(size_t)dir * W * H + (size_t)y * W + x] = (uint8_t)bi[i];
    }
  }
}

// ---------------------------------------------------------------------------
// k_ssim: block-averaged SSIM, window 8, stride 4.
// libs/pmbb-ivqm/src/xStructSimSTD.cpp:420-461 (window), xSSIM.cpp:238-254
// (row loop), :381-404 (window counts). Each wave owns 64 sub-block columns
// (63 windows) and a band of NB window rows; it walks the band top to bottom,
// carrying the previous 4x4 sub-block row so every input row is read once per
// band. Window moments are exact integers; the ratio is FP32 on the exact
// int64 numerators n*S_ab - S_a*S_b; window values are summed in FP64.
// ---------------------------------------------------------------------------
template <typename T, int CS>
__device__ __forceinline__ void load4(const T* __restrict__ p, int stride, int x, int y, uint32_t v[4]) {
  if (CS == 0) {
    const T* q = p + (size_t)y * stride + x;
    if (sizeof(T) == 2) {
      const ushort4 u = *(const ushort4*)q;
      v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
    } else {
      const uchar4 u = *(const uchar4*)q;
      v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
    }
  } else {
    const T* q = p + (size_t)(y >> 1) * stride + (x >> 1);
    const uint32_t a = q[0], b = q[1];
    v[0] = a; v[1] = a; v[2] = b; v[3] = b;
  }
}

template <typename TA, int CSA, typename TB, int CSB, bool WSQ>
__device__ void ssim_wave(const SsimJob& J, int c, int chunk, int band, int lane, float c1n, float c2n,
                          double* __restrict__ partial) {
  constexpr int NB = 8;
  using SQ = typename std::conditional<WSQ, uint64_t, uint32_t>::type;
  const int wy0 = band * NB;
  if (wy0 >= J.nuy || chunk >= J.nchunk) return;
  const int bx = chunk * 63 + lane;
  const bool col_ok = bx <= J.nux;
  const TA* pa = (const TA*)J.a.p[c];
  const TB* pb = (const TB*)J.b.p[c];
  const int sa = J.a.stride[c], sb = J.b.stride[c];
  uint32_t pA = 0, pB = 0;
  SQ pAA = 0, pBB = 0, pAB = 0;
  for (int k = 0; k <= NB; k++) {
    const int by = wy0 + k;
    if (by > J.nuy) break;
    uint32_t A = 0, B = 0;
    SQ AA = 0, BB = 0, AB = 0;
    if (col_ok) {
#pragma unroll
      for (int r = 0; r < 4; r++) {
        uint32_t a[4], b[4];
        load4<TA, CSA>(pa, sa, bx * 4, by * 4 + r, a);
        load4<TB, CSB>(pb, sb, bx * 4, by * 4 + r, b);
#pragma unroll
        for (int j = 0; j < 4; j++) {
          A += a[j];
          B += b[j];
          AA += (SQ)__umul24(a[j], a[j]);
          BB += (SQ)__umul24(b[j], b[j]);
          AB += (SQ)__umul24(a[j], b[j]);
        }
      }
    }
    if (k > 0) {
      const int wy = by - 1;
      const uint32_t vA = pA + A, vB = pB + B;
      const uint64_t vAA = (uint64_t)pAA + AA, vBB = (uint64_t)pBB + BB, vAB = (uint64_t)pAB + AB;
      const uint32_t nA = __shfl_down(vA, 1, 64), nB = __shfl_down(vB, 1, 64);
      const uint64_t nAA = __shfl_down(vAA, 1, 64), nBB = __shfl_down(vBB, 1, 64), nAB = __shfl_down(vAB, 1, 64);
      double s = 0.0;
      if (lane < 63 && bx < J.nux) {
        const int64_t sA = (int64_t)(vA + nA), sB = (int64_t)(vB + nB);
        const int64_t sAA = (int64_t)(vAA + nAA), sBB = (int64_t)(vBB + nBB), sAB = (int64_t)(vAB + nAB);
        const int64_t cov = 64 * sAB - sA * sB;
        const int64_t va = 64 * sAA - sA * sA;
        const int64_t vb = 64 * sBB - sB * sB;
        float cs = (2.0f * (float)cov + c2n) / ((float)va + (float)vb + c2n);
        if (J.calc_l) {
          const float l = (2.0f * (float)(sA * sB) + c1n) / ((float)(sA * sA) + (float)(sB * sB) + c1n);
          cs = l * cs;
        }
        s = (double)cs;
      }
      s = wave_sum_f64(s);
      if (lane == 0) partial[(size_t)J.part_off + ((size_t)c * J.nuy + wy) * J.nchunk + chunk] = s;
    }
    pA = A; pB = B; pAA = AA; pBB = BB; pAB = AB;
  }
}

template <bool WSQ>
__global__ __launch_bounds__(256) void k_ssim(SsimJobTable T, float c1n, float c2n, double* __restrict__ partial) {
  const int j = blockIdx.z / 3, c = blockIdx.z - 3 * (blockIdx.z / 3);
  const SsimJob& J = T.job[j];
  const int lane = threadIdx.x & 63, band = blockIdx.y * 4 + (threadIdx.x >> 6);
  const int csa = c ? J.a.cs : 0, csb = c ? J.b.cs : 0;
  const int chunk = blockIdx.x;
  if (J.a.bps == 2 && J.b.bps == 2) {
    if (csa == 0 && csb == 0) ssim_wave<uint16_t, 0, uint16_t, 0, WSQ>(J, c, chunk, band, lane, c1n, c2n, partial);
    else if (csa == 1 && csb == 1) ssim_wave<uint16_t, 1, uint16_t, 1, WSQ>(J, c, chunk, band, lane, c1n, c2n, partial);
    else if (csa == 1 && csb == 0) ssim_wave<uint16_t, 1, uint16_t, 0, WSQ>(J, c, chunk, band, lane, c1n, c2n, partial);
    else ssim_wave<uint16_t, 0, uint16_t, 1, WSQ>(J, c, chunk, band, lane, c1n, c2n, partial);
  } else if (!WSQ) {
    if (J.a.bps == 1 && J.b.bps == 1) {
      if (csa == 0 && csb == 0) ssim_wave<uint8_t, 0, uint8_t, 0, false>(J, c, chunk, band, lane, c1n, c2n, partial);
      else ssim_wave<uint8_t, 1, uint8_t, 1, false>(J, c, chunk, band, lane, c1n, c2n, partial);
    } else if (J.a.bps == 1) {
      if (csa == 0) ssim_wave<uint8_t, 0, uint16_t, 0, false>(J, c, chunk, band, lane, c1n, c2n, partial);
      else ssim_wave<uint8_t, 1, uint16_t, 0, false>(J, c, chunk, band, lane, c1n, c2n, partial);
    }
  }
}

// One block per (job, component): FP64 frame totals of the window-row sums.
// IV-SSIM ERP weighting follows libs/pmbb-ivqm/src/xIVSSIM.cpp:148-165.
__global__ __launch_bounds__(256) void k_ssim_finalize(SsimJobTable T, const double* __restrict__ partial,
                                                       const double* __restrict__ erp_w, int half_wnd, int stride,
                                                       double* __restrict__ out) {
  const int j = blockIdx.x / 3, c = blockIdx.x - 3 * (blockIdx.x / 3);
  const SsimJob& J = T.job[j];
  const int tid = threadIdx.x;
  double acc = 0.0, accw = 0.0;
  for (int wy = tid; wy < J.nuy; wy += 256) {
    const double* p = partial + (size_t)J.part_off + ((size_t)c * J.nuy + wy) * J.nchunk;
    double rs = 0.0;
    for (int k = 0; k < J.nchunk; k++) rs += p[k];
    if (J.erp && erp_w) {
      const double wt = erp_w[wy * stride + half_wnd];
      rs *= wt;
      accw += wt;
    }
    acc += rs;
  }
  __shared__ double s_a[4], s_w[4];
  acc = wave_sum_f64(acc);
  accw = wave_sum_f64(accw);
  if ((tid & 63) == 0) { s_a[tid >> 6] = acc; s_w[tid >> 6] = accw; }
  __syncthreads();
  if (tid == 0) {
    out[((size_t)j * 3 + c) * 2 + 0] = (s_a[0] + s_a[1]) + (s_a[2] + s_a[3]);
    out[((size_t)j * 3 + c) * 2 + 1] = (s_w[0] + s_w[1]) + (s_w[2] + s_w[3]);
  }
}

// 2x2 rounded mean: libs/pmbb-core/src/xPixelOpsSTD.cpp:50-67.
template <typename T>
__global__ __launch_bounds__(256) void k_downsample(PicDev S, int dw, int dh, uint16_t* __restrict__ dst, int pitch) {
  const int c = blockIdx.z;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= dw || y >= dh) return;
  const int v = ldpix<T>(S, c, 2 * x, 2 * y) + ldpix<T>(S, c, 2 * x + 1, 2 * y) + ldpix<T>(S, c, 2 * x, 2 * y + 1) +
                ldpix<T>(S, c, 2 * x + 1, 2 * y + 1) + 2;
  dst[(size_t)c * pitch * dh + (size_t)y * pitch + x] = (uint16_t)(v >> 2);
}

// 4:2:0 -> 4:4:4: libs/pmbb-core/src/xPixelOpsSTD.cpp:30-49 (luma copied).
template <typename T>
__global__ __launch_bounds__(256) void k_upsample420(PicDev S, int w, int h, uint16_t* __restrict__ dst, int dst_stride) {
  const int c = blockIdx.z;
  const int x = blockIdx.x * 64 + (threadIdx.x & 63);
  const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (x >= w || y >= h) return;
  dst[(size_t)c * dst_stride * h + (size_t)y * dst_stride + x] = (uint16_t)ldpix<T>(S, c, x, y);
}

}  // namespace

// ------------------------------- launchers ---------------------------------

hipError_t launch_stats(const PicDev& p0, const PicDev& p1, int w, int h, int64_t* row_sd, uint64_t* row_ssd,
                        uint32_t* zero_buf, int n_zero, hipStream_t st) {
  const int blocks = h + 2 * (h >> p0.cs);
  if ((int64_t)blocks * 256 < n_zero) return hipErrorInvalidValue;
  if (p0.bps == 2)
    hipLaunchKernelGGL(k_stats<uint16_t>, dim3(blocks), dim3(256), 0, st, p0, p1, w, h, row_sd, row_ssd, zero_buf, n_zero);
  else
    hipLaunchKernelGGL(k_stats<uint8_t>, dim3(blocks), dim3(256), 0, st, p0, p1, w, h, row_sd, row_ssd, zero_buf, n_zero);
  return hipGetLastError();
}

hipError_t launch_gcd_finalize(const int64_t* row_sd, const uint64_t* row_ssd, int w, int h, int cs, int /*bit_depth*/,
                               const int32_t maxdiff[3], int wrap32, int32_t* gcd_out, int64_t* sd_out,
                               uint64_t* ssd_out, hipStream_t st) {
  hipLaunchKernelGGL(k_gcd_finalize, dim3(1), dim3(256), 0, st, row_sd, row_ssd, w, h, cs, maxdiff[0], maxdiff[1],
                     maxdiff[2], wrap32, gcd_out, sd_out, ssd_out);
  return hipGetLastError();
}

template <typename T, int R, int PY>
static hipError_t search_t(const PicDev& p0, const PicDev& p1, int w, int h, int maxv, const int32_t* gcd,
                           const int32_t cw[4], bool narrow, uint32_t* rows, uint16_t* scp, int scp_pitch,
                           uint8_t* best, hipStream_t st) {
  dim3 grid((w + 63) / 64, (h + 4 * PY - 1) / (4 * PY), 2);
  if (narrow)
    hipLaunchKernelGGL((k_search<T, R, PY, true>), grid, dim3(256), 0, st, p0, p1, w, h, maxv, gcd, cw[0], cw[1], cw[2],
                       rows, scp, scp_pitch, best);
  else
    hipLaunchKernelGGL((k_search<T, R, PY, false>), grid, dim3(256), 0, st, p0, p1, w, h, maxv, gcd, cw[0], cw[1],
                       cw[2], rows, scp, scp_pitch, best);
  return hipGetLastError();
}

template <typename T, int PY>
static hipError_t search_r(const PicDev& p0, const PicDev& p1, int w, int h, int maxv, const int32_t* gcd,
                           const int32_t cw[4], int R, bool narrow, uint32_t* rows, uint16_t* scp, int scp_pitch,
                           uint8_t* best, hipStream_t st) {
  switch (R) {
    case 1: return search_t<T, 1, PY>(p0, p1, w, h, maxv, gcd, cw, narrow, rows, scp, scp_pitch, best, st);
    case 2: return search_t<T, 2, PY>(p0, p1, w, h, maxv, gcd, cw, narrow, rows, scp, scp_pitch, best, st);
    case 3: return search_t<T, 3, PY>(p0, p1, w, h, maxv, gcd, cw, narrow, rows, scp, scp_pitch, best, st);
    case 4: return search_t<T, 4, PY>(p0, p1, w, h, maxv, gcd, cw, narrow, rows, scp, scp_pitch, best, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_search(const PicDev& p0, const PicDev& p1, int w, int h, int bit_depth, const int32_t* gcd,
                         const int32_t cw[4], int R, bool narrow, uint32_t* rows, uint16_t* scp, int scp_pitch,
                         uint8_t* best, int tile_py, hipStream_t st) {
  const int maxv = (1 << bit_depth) - 1;
  if (p0.bps == 2) {
    if (tile_py == 4) return search_r<uint16_t, 4>(p0, p1, w, h, maxv, gcd, cw, R, narrow, rows, scp, scp_pitch, best, st);
    return search_r<uint16_t, 8>(p0, p1, w, h, maxv, gcd, cw, R, narrow, rows, scp, scp_pitch, best, st);
  }
  if (tile_py == 4) return search_r<uint8_t, 4>(p0, p1, w, h, maxv, gcd, cw, R, narrow, rows, scp, scp_pitch, best, st);
  return search_r<uint8_t, 8>(p0, p1, w, h, maxv, gcd, cw, R, narrow, rows, scp, scp_pitch, best, st);
}

hipError_t launch_ssim(const SsimJobTable& jobs, int max_nchunk, int max_bands, float c1n, float c2n, bool wide_sq,
                       double* partial, hipStream_t st) {
  if (jobs.njobs <= 0) return hipSuccess;
  dim3 grid(max_nchunk, (max_bands + 3) / 4, jobs.njobs * 3);
  if (wide_sq) hipLaunchKernelGGL(k_ssim<true>, grid, dim3(256), 0, st, jobs, c1n, c2n, partial);
  else hipLaunchKernelGGL(k_ssim<false>, grid, dim3(256), 0, st, jobs, c1n, c2n, partial);
  return hipGetLastError();
}

hipError_t launch_ssim_finalize(const SsimJobTable& jobs, const double* partial, const double* erp_w, int wnd,
                                int stride, double* out, hipStream_t st) {
  if (jobs.njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ssim_finalize, dim3(jobs.njobs * 3), dim3(256), 0, st, jobs, partial, erp_w, wnd >> 1, stride, out);
  return hipGetLastError();
}

hipError_t launch_downsample(const PicDev& src, int w, int h, uint16_t* dst, int dst_pitch, hipStream_t st) {
  const int dw = w >> 1, dh = h >> 1;
  dim3 grid((dw + 63) / 64, (dh + 3) / 4, 3);
  if (src.bps == 2) hipLaunchKernelGGL(k_downsample<uint16_t>, grid, dim3(256), 0, st, src, dw, dh, dst, dst_pitch);
  else hipLaunchKernelGGL(k_downsample<uint8_t>, grid, dim3(256), 0, st, src, dw, dh, dst, dst_pitch);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_zero_u32(uint32_t* __restrict__ p, int n) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}
hipError_t launch_zero_u32(uint32_t* p, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_u32, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_upsample420(const PicDev& src, int w, int h, uint16_t* dst, int dst_stride, hipStream_t st) {
  dim3 grid((w + 63) / 64, (h + 3) / 4, 3);
  if (src.bps == 2) hipLaunchKernelGGL(k_upsample420<uint16_t>, grid, dim3(256), 0, st, src, w, h, dst, dst_stride);
  else hipLaunchKernelGGL(k_upsample420<uint8_t>, grid, dim3(256), 0, st, src, w, h, dst, dst_stride);
  return hipGetLastError();
}

}  // namespace ivqm
