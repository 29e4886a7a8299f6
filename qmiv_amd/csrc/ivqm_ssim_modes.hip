// gfx950 kernels of every structural-similarity configuration other than the
// default block-averaged 8 x 8 / stride 4 (which k_ssim in ivqm_kernels.hip
// runs): regular 11 x 11 windows (Gaussian float, Gaussian Q18, averaged) with
// the border extensions, block Gaussian Q18, block averaged of any window and
// stride. Window sums feed the same per-window-row partial buffer and
// k_ssim_finalize as k_ssim.
#include "ivqm_device.h"
#include "ivqm_device_util.h"

namespace ivqm {

namespace {

// ---------------------------------------------------------------------------
// k_ssim_gen: every SSIM window mode, size, stride and border extension.
// libs/pmbb-ivqm/src/xStructSimSTD.cpp:19-193 (regular 11x11: Gaussian float
// weights with FP64 sums, Q18 integer weights, plain sums), :353-419 (block
// Gaussian Q18), :420-461 (block averaged); window grid xSSIM.cpp:161-196;
// border extension xMarginOps.h:40-215 (xCommonDefCORE.h:322-331).
//
// Lane = one window column (64 per wave), wave = kSsimGenRows consecutive
// window rows; each sample row the wave's windows cover is read once and fed
// to every window containing it. Moments accumulate in FP64: with integer
// weights every product and partial sum is an integer below 2^53, so they are
// the reference's int64 sums exactly; with float weights the sums run in the
// reference's order (rows, then columns). The ratio is FP64 as written in the
// reference.
// ---------------------------------------------------------------------------
// Index of the extended plane's sample u (0 <= u < n inside): Nearest clamps,
// Reflect repeats the edge sample (c b a | a b c), Mirror does not (d c b | a b c d);
// -1 for the Constant / Zero border, whose samples hold a fixed value.
__device__ __forceinline__ int ext_index(int u, int n, int ext) {
  if ((unsigned)u < (unsigned)n) return u;
  if (ext == 1) return u < 0 ? 0 : n - 1;
  if (ext == 2) return u < 0 ? -u - 1 : 2 * n - 1 - u;
  if (ext == 3) return u < 0 ? -u : 2 * n - 2 - u;
  return -1;
}

template <typename TA, typename TB, int NW, bool WEIGHTED, bool EXT>
__device__ void ssim_gen_wave(const SsimJobDev& J, int c, int chunk, int wr0, int lane, double* __restrict__ partial) {
  constexpr int KY = kSsimGenRows;
  const int n = NW ? NW : J.wnd;
  const int s = J.stride;
  const int nk = min(KY, J.nuy - wr0);
  const int wx = chunk * 64 + lane;
  const bool active = wx < J.nux;
  const int wxc = active ? wx : J.nux - 1;  // idle lanes repeat the last window (in bounds)
  const int off = J.mode <= 2 ? 5 : 0;      // regular windows are centred
  const int xs = J.gx0 + wxc * s - off;     // first sample column of this lane's windows
  const int ys = J.gy0 + wr0 * s - off;     // first sample row of the wave's windows
  const int rows = (nk - 1) * s + n;
  typedef __attribute__((address_space(1))) const TA GTA;
  typedef __attribute__((address_space(1))) const TB GTB;
  GTA* pa = (GTA*)plane_ptr(J.a, c);
  GTB* pb = (GTB*)plane_ptr(J.b, c);
  const int sa = plane_stride(J.a, c), sb = plane_stride(J.b, c);
  const int csa = c ? J.a.cs : 0, csb = c ? J.b.cs : 0;      // chroma column shifts
  const int rsa = c ? J.a.csy : 0, rsb = c ? J.b.csy : 0;    // chroma row shifts
  const double kfix = J.ext == 4 ? 65535.0 : 0.0;  // Constant border: xPic.cpp:94-101 passes uint16 max
  double mA[KY], mB[KY], mAA[KY], mBB[KY], mAB[KY];
#pragma unroll
  for (int k = 0; k < KY; k++) mA[k] = mB[k] = mAA[k] = mBB[k] = mAB[k] = 0.0;

#pragma unroll 1
  for (int r = 0; r < rows; r++) {
    // windows k fed by this row: 0 <= r - k s < n (wave-uniform)
    const int khi = min(nk - 1, r / s);
    const int klo = r < n ? 0 : (r - n) / s + 1;
    int y = ys + r;
    bool yfix = false;
    if (EXT) {
      y = ext_index(y, J.h, J.ext);
      yfix = y < 0;
    }
    GTA* ra = pa + (size_t)(uint32_t)((yfix ? 0 : y) >> rsa) * (uint32_t)sa;
    GTB* rb = pb + (size_t)(uint32_t)((yfix ? 0 : y) >> rsb) * (uint32_t)sb;
#pragma unroll 4
    for (int dx = 0; dx < n; dx++) {
      double a, b;
      if (EXT) {
        const int x = ext_index(xs + dx, J.w, J.ext);
        if (yfix || x < 0) {
          a = kfix;
          b = kfix;
        } else {
          a = (double)ra[x >> csa];
          b = (double)rb[x >> csb];
        }
      } else {
        const int x = xs + dx;
        a = (double)ra[x >> csa];
        b = (double)rb[x >> csb];
      }
      const double aa = a * a, bb = b * b, ab = a * b;  // exact: integers below 2^32
#pragma unroll
      for (int k = 0; k < KY; k++) {
        if (k < klo || k > khi) continue;
        if (WEIGHTED) {
          const double w = J.wt[(r - k * s) * n + dx];
          mA[k] = fma(a, w, mA[k]);
          mB[k] = fma(b, w, mB[k]);
          mAA[k] = fma(aa, w, mAA[k]);
          mBB[k] = fma(bb, w, mBB[k]);
          mAB[k] = fma(ab, w, mAB[k]);
        } else {
          mA[k] += a;
          mB[k] += b;
          mAA[k] += aa;
          mBB[k] += bb;
          mAB[k] += ab;
        }
      }
    }
  }
  const double K = J.kscale;
#pragma unroll
  for (int k = 0; k < KY; k++) {
    if (k >= nk) break;
    double v = 0.0;
    if (active) {
      // xStructSimSTD.cpp:104-108 / :176-180 / :399-414: means, variances, covariance, L * CS
      const double ar = mA[k] * K, at = mB[k] * K;
      const double vr = mAA[k] * K - ar * ar, vt = mBB[k] * K - at * at, cv = mAB[k] * K - ar * at;
      v = (2 * cv + J.c2) / (vr + vt + J.c2);
      if (J.calc_l) v = ((2 * ar * at + J.c1) / (ar * ar + at * at + J.c1)) * v;
    }
    v = wave_sum_f64_dpp(v);
    if (lane == 0) partial[J.part_off + ((int64_t)c * J.nuy + wr0 + k) * J.nchunk + chunk] = v;
  }
}

// Static-row kernels (compile-time stride S and window N): the wave's 8
// windows per column span (8 - 1) S + N sample rows. The column offset dx is
// the outer loop, with the N weights of that column (one per window row)
// loaded once as wave-uniform values; the rows are unrolled, so which windows
// a row feeds and with which weight is fixed at compile time and the FMA
// stream has no branches. SSIM is symmetric in its two pictures, so a pair of
// 16-bit and 8-bit planes runs as 8-bit / 16-bit.
template <typename TA, typename TB, int S, int N, bool WEIGHTED, bool EXT>
__device__ void ssim_static_wave(const SsimJobDev& J, const PicDev& A, const PicDev& B, int c, int chunk, int wr0,
                                 int lane, double* __restrict__ partial) {
  constexpr int KY = kSsimGenRows, ROWS = (KY - 1) * S + N;
  const int nk = min(KY, J.nuy - wr0);
  const int wx = chunk * 64 + lane;
  const bool active = wx < J.nux;
  const int wxc = active ? wx : J.nux - 1;
  const int off = J.mode <= 2 ? 5 : 0;  // regular windows are centred
  const int xs = J.gx0 + wxc * S - off, ys = J.gy0 + wr0 * S - off;
  const int ylast = ys + (nk - 1) * S + N - 1;  // last row a live window reads; later rows are clamped to it
  typedef __attribute__((address_space(1))) const TA GTA;
  typedef __attribute__((address_space(1))) const TB GTB;
  GTA* pa = (GTA*)plane_ptr(A, c);
  GTB* pb = (GTB*)plane_ptr(B, c);
  const int sa = plane_stride(A, c), sb = plane_stride(B, c);
  const int csa = c ? A.cs : 0, csb = c ? B.cs : 0;      // chroma column shifts
  const int rsa = c ? A.csy : 0, rsb = c ? B.csy : 0;    // chroma row shifts
  const double kfix = J.ext == 4 ? 65535.0 : 0.0;
  const double* __restrict__ wt_t = J.wt + N * N;  // transposed weights: [dx][dy]
  double mA[KY], mB[KY], mAA[KY], mBB[KY], mAB[KY];
#pragma unroll
  for (int k = 0; k < KY; k++) mA[k] = mB[k] = mAA[k] = mBB[k] = mAB[k] = 0.0;

#pragma unroll 1
  for (int dx = 0; dx < N; dx++) {
    double w[N];
#pragma unroll
    for (int dy = 0; dy < N; dy++) w[dy] = WEIGHTED ? wt_t[dx * N + dy] : 1.0;
    int x = xs + dx;
    if (EXT) x = ext_index(x, J.w, J.ext);
    const bool xfix = EXT && x < 0;
    const uint32_t xa = (uint32_t)(xfix ? 0 : x) >> csa, xb = (uint32_t)(xfix ? 0 : x) >> csb;
#pragma unroll
    for (int rho = 0; rho < ROWS; rho++) {
      int y = min(ys + rho, ylast);
      bool fix = xfix;
      if (EXT) {
        y = ext_index(y, J.h, J.ext);
        fix = fix || y < 0;
        y = y < 0 ? 0 : y;
      }
      double a = (double)pa[(uint32_t)(y >> rsa) * (uint32_t)sa + xa];
      double b = (double)pb[(uint32_t)(y >> rsb) * (uint32_t)sb + xb];
      if (EXT && fix) {
        a = kfix;
        b = kfix;
      }
      const double aa = a * a, bb = b * b, ab = a * b;
#pragma unroll
      for (int k = 0; k < KY; k++) {
        const int dy = rho - k * S;  // compile-time
        if (dy < 0 || dy >= N) continue;
        if (WEIGHTED) {
          mA[k] = fma(a, w[dy], mA[k]);
          mB[k] = fma(b, w[dy], mB[k]);
          mAA[k] = fma(aa, w[dy], mAA[k]);
          mBB[k] = fma(bb, w[dy], mBB[k]);
          mAB[k] = fma(ab, w[dy], mAB[k]);
        } else {
          mA[k] += a;
          mB[k] += b;
          mAA[k] += aa;
          mBB[k] += bb;
          mAB[k] += ab;
        }
      }
    }
  }
  const double K = J.kscale;
#pragma unroll
  for (int k = 0; k < KY; k++) {
    if (k >= nk) break;
    double v = 0.0;
    if (active) {
      const double ar = mA[k] * K, at = mB[k] * K;
      const double vr = mAA[k] * K - ar * ar, vt = mBB[k] * K - at * at, cv = mAB[k] * K - ar * at;
      v = (2 * cv + J.c2) / (vr + vt + J.c2);
      if (J.calc_l) v = ((2 * ar * at + J.c1) / (ar * ar + at * at + J.c1)) * v;
    }
    v = wave_sum_f64_dpp(v);
    if (lane == 0) partial[J.part_off + ((int64_t)c * J.nuy + wr0 + k) * J.nchunk + chunk] = v;
  }
}

template <int S, int N, bool WEIGHTED, bool EXT>
__global__ __launch_bounds__(256) void k_ssim_static(const SsimJobDev* __restrict__ jobs, int njobs, int nblk,
                                                     double* __restrict__ partial) {
  const int per = (nblk + kXcds - 1) / kXcds;
  const int b = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (b >= nblk) return;
  const int ji = find_entry(jobs, njobs, b, [](const SsimJobDev& j) { return j.blk0; });
  const SsimJobDev J = jobs[ji];
  const int lb = b - J.blk0;
  const int per_c = J.nchunk * J.nbg;
  const int c = lb / per_c, rem = lb - c * per_c;
  const int bg = rem / J.nchunk, chunk = rem - bg * J.nchunk;
  const int lane = threadIdx.x & 63;
  const int wr0 = (bg * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * kSsimGenRows;
  if (wr0 >= J.nuy) return;
  if (J.a.bps == 2 && J.b.bps == 2)
    ssim_static_wave<uint16_t, uint16_t, S, N, WEIGHTED, EXT>(J, J.a, J.b, c, chunk, wr0, lane, partial);
  else if (J.a.bps == 1 && J.b.bps == 1)
    ssim_static_wave<uint8_t, uint8_t, S, N, WEIGHTED, EXT>(J, J.a, J.b, c, chunk, wr0, lane, partial);
  else if (J.a.bps == 1)
    ssim_static_wave<uint8_t, uint16_t, S, N, WEIGHTED, EXT>(J, J.a, J.b, c, chunk, wr0, lane, partial);
  else
    ssim_static_wave<uint8_t, uint16_t, S, N, WEIGHTED, EXT>(J, J.b, J.a, c, chunk, wr0, lane, partial);
}

// ---------------------------------------------------------------------------
// k_ssim_roll: stride-4 windows of N = 8, 11 or 16 samples without a border extension (regular 11 x 11
// centred windows from column 5, block windows from column 0: window column wx starts at sample 4 wx).
// Lane l reads the 4 samples 4 l .. 4 l + 3 of a row with one vector load and takes the next
// N - 4 samples from lanes l + 1, l + 2, ... with DPP wave shifts, so the wave's windows are 64 - (N - 1) / 4
// columns (the last lanes only supply samples; ssim_gen_shape). The wave walks kSsimRollRows window rows
// top to bottom, 4 sample rows per step: a step's rows feed the JW = ceil(N / 4) windows that overlap them,
// held as rolling accumulators (window q - j in slot j at step q), and after each step the oldest window is
// complete: its score joins its window row's partial sum. Each sample row is read once per wave.
// Moments in FP64 as in k_ssim_gen (integer weights keep every sum exact; averaged windows add with weight
// 1, and a row outside a window adds its samples with weight 0, which leaves the sums unchanged).
// ---------------------------------------------------------------------------
// Lane l + 1's value (wave_shl:1); lane 63 reads 0 (bound_ctrl: a source lane out of range reads 0).
__device__ __forceinline__ uint32_t wave_next(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x130, 0xf, 0xf, true);
}

// Samples 4 g .. 4 g + 3 of a row (luma-grid columns; chroma at x >> cs), clamped to the row's last sample.
template <typename T>
__device__ __forceinline__ void roll_load4(const __attribute__((address_space(1))) T* row, int g4, int cs, int wlast,
                                           uint32_t v[4]) {
  typedef __attribute__((address_space(1))) const T GT;
  if (g4 + 3 <= wlast) {
    if (cs == 0) {
      if (sizeof(T) == 2) {
        typedef unsigned short u16x4r __attribute__((ext_vector_type(4)));
        const u16x4r u = *(const __attribute__((address_space(1))) u16x4r*)(row + g4);
        v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
      } else {
        typedef unsigned char u8x4r __attribute__((ext_vector_type(4)));
        const u8x4r u = *(const __attribute__((address_space(1))) u8x4r*)(row + g4);
        v[0] = u.x; v[1] = u.y; v[2] = u.z; v[3] = u.w;
      }
    } else {
      const uint32_t a = row[g4 >> cs], b = row[(g4 >> cs) + 1];
      v[0] = a; v[1] = cs == 1 ? a : b; v[2] = cs == 1 ? b : row[(g4 >> cs) + 2]; v[3] = cs == 1 ? b : row[(g4 >> cs) + 3];
    }
  } else {  // the right edge: samples past the row repeat its last one (only lanes past the windows read them)
#pragma unroll
    for (int t = 0; t < 4; t++) v[t] = row[min(g4 + t, wlast) >> cs];
  }
}

template <typename TA, typename TB, int N, bool WEIGHTED>
__device__ void ssim_roll_wave(const SsimJobDev& J, const PicDev& A, const PicDev& B, int c, int chunk, int wr0,
                               int lane, const double* wl, double* __restrict__ partial) {
  constexpr int JW = (N + 3) / 4, NG = JW, WPW = 64 - (N - 1) / 4, KR = kSsimRollRows;
  static_assert(N == 8 || N == 11 || N == 16, "window");
  const int nk = min(KR, J.nuy - wr0);
  const int wx = chunk * WPW + lane;
  const bool active = lane < WPW && wx < J.nux;
  const int g4 = 4 * wx;  // this lane's first sample column (windows start at 4 wx: see the launch rule)
  const int ys = J.gy0 + wr0 * 4 - (J.mode <= 2 ? 5 : 0);
  const int ylast = ys + (nk - 1) * 4 + N - 1;  // last row a live window reads; later rows are clamped to it
  typedef __attribute__((address_space(1))) const TA GTA;
  typedef __attribute__((address_space(1))) const TB GTB;
  GTA* pa = (GTA*)plane_ptr(A, c);
  GTB* pb = (GTB*)plane_ptr(B, c);
  const int sa = plane_stride(A, c), sb = plane_stride(B, c);
  const int csa = c ? A.cs : 0, csb = c ? B.cs : 0;
  const int rsa = c ? A.csy : 0, rsb = c ? B.csy : 0;
  const int wlast = J.w - 1;
  const double K = J.kscale;
  double mA[JW], mB[JW], mAA[JW], mBB[JW], mAB[JW];
#pragma unroll
  for (int j = 0; j < JW; j++) mA[j] = mB[j] = mAA[j] = mBB[j] = mAB[j] = 0.0;
  const int nq = nk + JW - 1;  // steps: window nk - 1 completes in step nk - 1 + JW - 1
#pragma unroll 1
  for (int q = 0; q < nq; q++) {
#pragma unroll 1
    for (int r = 0; r < 4; r++) {
      // one sample row at a time: the weights stay in LDS (read where used, not hoisted into registers)
      asm volatile("" ::: "memory");
      const int y = min(ys + 4 * q + r, ylast);
      uint32_t va[4 * NG], vb[4 * NG];
      roll_load4<TA>(pa + (size_t)(uint32_t)(y >> rsa) * (uint32_t)sa, g4, csa, wlast, va);
      roll_load4<TB>(pb + (size_t)(uint32_t)(y >> rsb) * (uint32_t)sb, g4, csb, wlast, vb);
#pragma unroll
      for (int m = 1; m < NG; m++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
          va[4 * m + t] = wave_next(va[4 * (m - 1) + t]);
          vb[4 * m + t] = wave_next(vb[4 * (m - 1) + t]);
        }
#pragma unroll
      for (int dx = 0; dx < N; dx++) {
        const double a = (double)va[dx], b = (double)vb[dx];
        const double aa = a * a, bb = b * b, ab = a * b;  // exact: integers below 2^32
#pragma unroll
        for (int j = 0; j < JW; j++) {
          // slot j holds window q - j, whose row dy = r + 4 j this is; rows dy >= N carry zero weights
          const double w = wl[(r + 4 * j) * N + dx];
          mA[j] = fma(a, w, mA[j]);
          mB[j] = fma(b, w, mB[j]);
          mAA[j] = fma(aa, w, mAA[j]);
          mBB[j] = fma(bb, w, mBB[j]);
          mAB[j] = fma(ab, w, mAB[j]);
        }
        // one column at a time (15 independent accumulations per column keep the FP64 pipe busy; letting
        // the scheduler hoist the conversions and products of every column raises the live registers)
        __builtin_amdgcn_sched_barrier(0);
      }
    }
    // window k = q - (JW - 1) is complete
    const int k = q - (JW - 1);  // wave-uniform
    if (k >= 0) {
      double v = 0.0;
      if (active) {
        // xStructSimSTD.cpp:104-108 / :176-180 / :399-414: means, variances, covariance, L * CS
        const double ar = mA[JW - 1] * K, at = mB[JW - 1] * K;
        const double vr = mAA[JW - 1] * K - ar * ar, vt = mBB[JW - 1] * K - at * at, cv = mAB[JW - 1] * K - ar * at;
        v = (2 * cv + J.c2) / (vr + vt + J.c2);
        if (J.calc_l) v = ((2 * ar * at + J.c1) / (ar * ar + at * at + J.c1)) * v;
      }
      v = wave_sum_f64_dpp(v);
      if (lane == 0) partial[J.part_off + ((int64_t)c * J.nuy + wr0 + k) * J.nchunk + chunk] = v;
    }
#pragma unroll
    for (int j = JW - 1; j > 0; j--) {
      mA[j] = mA[j - 1];
      mB[j] = mB[j - 1];
      mAA[j] = mAA[j - 1];
      mBB[j] = mBB[j - 1];
      mAB[j] = mAB[j - 1];
    }
    mA[0] = mB[0] = mAA[0] = mBB[0] = mAB[0] = 0.0;
  }
}

template <int N, bool WEIGHTED>
__global__ __launch_bounds__(256) void k_ssim_roll(const SsimJobDev* __restrict__ jobs, int njobs, int nblk,
                                                   double* __restrict__ partial) {
  const int per = (nblk + kXcds - 1) / kXcds;  // XCD-aware order, as in k_ssim
  const int b = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (b >= nblk) return;
  const int ji = find_entry(jobs, njobs, b, [](const SsimJobDev& j) { return j.blk0; });
  const SsimJobDev J = jobs[ji];
  const int lb = b - J.blk0;
  const int per_c = J.nchunk * J.nbg;
  const int c = lb / per_c, rem = lb - c * per_c;
  const int bg = rem / J.nchunk, chunk = rem - bg * J.nchunk;
  const int lane = threadIdx.x & 63;
  // The window weights [dy][dx] in LDS (the same table for every job; 1 for averaged windows), padded with
  // zero rows to 4 ceil(N / 4): a step's row r feeds slot j with weight row r + 4 j, zero past the window.
  constexpr int NR = 4 * ((N + 3) / 4);
  __shared__ double wl[NR * N];
  for (int i = (int)threadIdx.x; i < NR * N; i += 256) wl[i] = i < N * N ? (WEIGHTED ? J.wt[i] : 1.0) : 0.0;
  __syncthreads();
  const int wr0 = (bg * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * kSsimRollRows;
  if (wr0 >= J.nuy) return;  // whole wave
  if (J.a.bps == 2 && J.b.bps == 2)
    ssim_roll_wave<uint16_t, uint16_t, N, WEIGHTED>(J, J.a, J.b, c, chunk, wr0, lane, wl, partial);
  else if (J.a.bps == 1 && J.b.bps == 1)
    ssim_roll_wave<uint8_t, uint8_t, N, WEIGHTED>(J, J.a, J.b, c, chunk, wr0, lane, wl, partial);
  else if (J.a.bps == 1)
    ssim_roll_wave<uint8_t, uint16_t, N, WEIGHTED>(J, J.a, J.b, c, chunk, wr0, lane, wl, partial);
  else  // SSIM is symmetric in its two pictures
    ssim_roll_wave<uint8_t, uint16_t, N, WEIGHTED>(J, J.b, J.a, c, chunk, wr0, lane, wl, partial);
}

template <int NW, bool WEIGHTED, bool EXT>
__global__ __launch_bounds__(256) void k_ssim_gen(const SsimJobDev* __restrict__ jobs, int njobs, int nblk,
                                                  double* __restrict__ partial) {
  const int per = (nblk + kXcds - 1) / kXcds;  // XCD-aware order, as in k_ssim
  const int b = (int)(blockIdx.x % kXcds) * per + (int)(blockIdx.x / kXcds);
  if (b >= nblk) return;
  const int ji = find_entry(jobs, njobs, b, [](const SsimJobDev& j) { return j.blk0; });
  const SsimJobDev J = jobs[ji];
  const int lb = b - J.blk0;
  const int per_c = J.nchunk * J.nbg;
  const int c = lb / per_c, rem = lb - c * per_c;
  const int bg = rem / J.nchunk, chunk = rem - bg * J.nchunk;
  const int lane = threadIdx.x & 63;
  const int wr0 = (bg * 4 + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * kSsimGenRows;
  if (wr0 >= J.nuy) return;  // whole wave
  if (J.a.bps == 2 && J.b.bps == 2)
    ssim_gen_wave<uint16_t, uint16_t, NW, WEIGHTED, EXT>(J, c, chunk, wr0, lane, partial);
  else if (J.a.bps == 1 && J.b.bps == 1)
    ssim_gen_wave<uint8_t, uint8_t, NW, WEIGHTED, EXT>(J, c, chunk, wr0, lane, partial);
  else if (J.a.bps == 1)
    ssim_gen_wave<uint8_t, uint16_t, NW, WEIGHTED, EXT>(J, c, chunk, wr0, lane, partial);
  else
    ssim_gen_wave<uint16_t, uint8_t, NW, WEIGHTED, EXT>(J, c, chunk, wr0, lane, partial);
}

}  // namespace

namespace {
template <int S, int N, bool W, bool E>
void launch_static(dim3 g, dim3 t, hipStream_t st, const SsimJobDev* jobs, int njobs, int nblk, double* partial) {
  hipLaunchKernelGGL((k_ssim_static<S, N, W, E>), g, t, 0, st, jobs, njobs, nblk, partial);
}
template <int S, int N, bool W>
bool launch_static_ext(bool ext, dim3 g, dim3 t, hipStream_t st, const SsimJobDev* jobs, int njobs, int nblk,
                       double* partial) {
  if (ext) launch_static<S, N, W, true>(g, t, st, jobs, njobs, nblk, partial);
  else launch_static<S, N, W, false>(g, t, st, jobs, njobs, nblk, partial);
  return true;
}
// Static-row kernel for (stride, window) when one is instantiated; false otherwise.
bool launch_static_any(int s, int n, bool w, bool ext, dim3 g, dim3 t, hipStream_t st, const SsimJobDev* jobs,
                       int njobs, int nblk, double* partial) {
#define IVQM_REG(S)                                                                                   \
  if (n == 11 && s == S)                                                                              \
    return w ? launch_static_ext<S, 11, true>(ext, g, t, st, jobs, njobs, nblk, partial)              \
             : launch_static_ext<S, 11, false>(ext, g, t, st, jobs, njobs, nblk, partial);
#define IVQM_BLK(S, N)                                                                                \
  if (n == N && s == S) {                                                                             \
    if (w) launch_static<S, N, true, false>(g, t, st, jobs, njobs, nblk, partial);                    \
    else launch_static<S, N, false, false>(g, t, st, jobs, njobs, nblk, partial);                     \
    return true;                                                                                      \
  }
  IVQM_REG(1) IVQM_REG(2) IVQM_REG(3) IVQM_REG(4)
  IVQM_BLK(1, 8) IVQM_BLK(2, 8) IVQM_BLK(4, 8) IVQM_BLK(8, 8)
  IVQM_BLK(1, 16) IVQM_BLK(2, 16) IVQM_BLK(4, 16) IVQM_BLK(8, 16) IVQM_BLK(16, 16)
#undef IVQM_REG
#undef IVQM_BLK
  return false;
}
}  // namespace

hipError_t launch_ssim_gen(const SsimJobDev* jobs, int njobs, int nblk, int mode, int wnd, int ext, int stride,
                           double* partial, hipStream_t st) {
  if (njobs <= 0 || nblk <= 0) return hipSuccess;
  const dim3 g(kXcds * ((nblk + kXcds - 1) / kXcds)), t(256);
  const bool regular = mode <= 2, weighted = mode == 0 || mode == 1 || mode == 3, ext_on = regular && ext != 0;
  const int n = regular ? 11 : wnd;
  if (ssim_gen_shape(mode, wnd, ext, stride).roll) {
#define IVQM_ROLL(N)                                                                                         \
  if (n == N) {                                                                                              \
    if (weighted) hipLaunchKernelGGL((k_ssim_roll<N, true>), g, t, 0, st, jobs, njobs, nblk, partial);        \
    else hipLaunchKernelGGL((k_ssim_roll<N, false>), g, t, 0, st, jobs, njobs, nblk, partial);               \
    return hipGetLastError();                                                                                \
  }
    IVQM_ROLL(8) IVQM_ROLL(11) IVQM_ROLL(16)
#undef IVQM_ROLL
    return hipErrorInvalidValue;
  }
  // static rows pay at stride 1 and 2, and for weighted windows (measured at 4096 x 2048: averaged 16 / 4
  // runs faster in the general kernel, whose lanes share fewer cache lines per load at larger strides)
  if ((stride <= 2 || weighted) && launch_static_any(stride, n, weighted, ext_on, g, t, st, jobs, njobs, nblk, partial))
    return hipGetLastError();
  if (regular) {
    if (weighted && ext_on) hipLaunchKernelGGL((k_ssim_gen<11, true, true>), g, t, 0, st, jobs, njobs, nblk, partial);
    else if (weighted) hipLaunchKernelGGL((k_ssim_gen<11, true, false>), g, t, 0, st, jobs, njobs, nblk, partial);
    else if (ext_on) hipLaunchKernelGGL((k_ssim_gen<11, false, true>), g, t, 0, st, jobs, njobs, nblk, partial);
    else hipLaunchKernelGGL((k_ssim_gen<11, false, false>), g, t, 0, st, jobs, njobs, nblk, partial);
  } else if (mode == 3) {
    if (wnd == 8) hipLaunchKernelGGL((k_ssim_gen<8, true, false>), g, t, 0, st, jobs, njobs, nblk, partial);
    else if (wnd == 16) hipLaunchKernelGGL((k_ssim_gen<16, true, false>), g, t, 0, st, jobs, njobs, nblk, partial);
    else return hipErrorInvalidValue;
  } else {
    hipLaunchKernelGGL((k_ssim_gen<0, false, false>), g, t, 0, st, jobs, njobs, nblk, partial);
  }
  return hipGetLastError();
}

}  // namespace ivqm
