// Seeded synthetic frame-pair generator for the QMIV metric workloads.
//
// Host-only C++ (no HIP). Produces deterministic planar YCbCr pictures so that
// the GPU path, the CPU oracle and the reference build all see identical input.
//
// Specification (frozen; a fixture under tests/golden pins the first and last
// samples of every plane of a small pair so that drift is detected):
//   PRNG       splitmix64. Stream seed = splitmix64(0x514D4956 ^ cfg<<40 ^
//              frame<<16 ^ picture<<8 ^ plane). Per-row noise streams use
//              splitmix64(stream_seed ^ (row+1)*0x9E3779B97F4A7C15).
//   Uniform    top 53 bits of one draw, scaled by 2^-53.
//   Gaussian   Box-Muller in FP64 (one normal per uniform pair, cosine branch).
//   Rounding   half away from zero, then clip to [0, 2^bitdepth - 1].
//   Reference  per plane: gradient a0 + ax*x/Wp + ay*y/Hp plus 8 sinusoids
//   picture    A*sin(2*pi*f*(x*cos(t) + y*sin(t)) + phi), plus N(0, sigma_r^2).
//              Luma: a0 in [0.2,0.4]*Max, ax, ay in [0.1,0.3]*Max,
//              A in [0.01,0.05]*Max, f in [1/256,1/16] cycles/px, t, phi in [0,2pi).
//              Chroma (at chroma resolution): a0 in [0.4,0.6]*Max; gradient and
//              sinusoid amplitudes halved.
//   Test       per plane: ref(clamp(x+dx+gx), clamp(y+dy+gy)) + offset + N(0, sigma_t^2).
//   picture    (dx, dy) is drawn uniformly from [-S, S]^2 per 32x32 luma block in
//              raster order; a chroma sample uses the block of its co-sited luma
//              sample and the same integer shift at chroma resolution. (gx, gy)
//              is a per-frame global shift drawn from [-G, G]^2 before the block
//              shifts. Offsets are fixed per plane, or drawn from U{-K..K} per
//              plane when offset_rand > 0.
//   Picture 0 is the test picture (QMIV InputFile0), picture 1 the reference
//   (QMIV InputFile1).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

inline uint64_t splitmix64_next(uint64_t& s) {
  s += 0x9E3779B97F4A7C15ull;
  uint64_t z = s;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
inline uint64_t mix64(uint64_t v) { uint64_t s = v; return splitmix64_next(s); }

struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  double uni() { return (double)(splitmix64_next(s) >> 11) * (1.0 / 9007199254740992.0); }
  double uni(double a, double b) { return a + (b - a) * uni(); }
  int32_t uni_int(int32_t lo, int32_t hi) {  // inclusive range
    const double u = uni();
    int32_t v = lo + (int32_t)std::floor(u * (double)(hi - lo + 1));
    return v > hi ? hi : v;
  }
  double gauss() {
    const double u1 = 1.0 - uni();  // (0,1]
    const double u2 = uni();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
};

inline uint64_t stream_seed(int cfg, int frame, int picture, int plane) {
  return mix64(0x514D4956ull ^ ((uint64_t)(uint32_t)cfg << 40) ^ ((uint64_t)(uint32_t)frame << 16) ^
               ((uint64_t)(uint32_t)picture << 8) ^ (uint64_t)(uint32_t)plane);
}

inline double round_half_away(double v) { return v >= 0.0 ? std::floor(v + 0.5) : -std::floor(-v + 0.5); }
inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return v < lo ? lo : (v > hi ? hi : v); }

}  // namespace

extern "C" {

typedef struct ivqm_synth_spec {
  int32_t cfg;          // workload id, enters the seed
  int32_t width;        // luma width
  int32_t height;       // luma height
  int32_t bit_depth;    // 8..16
  int32_t chroma_420;   // 1 = 4:2:0, 0 = 4:4:4
  double  sigma_r;      // reference noise sigma (LSB)
  double  sigma_t;      // test noise sigma (LSB)
  int32_t block_shift;  // S: block shifts drawn from [-S, S]
  int32_t global_shift; // G: per-frame global shift drawn from [-G, G]
  int32_t offsets[3];   // fixed per-plane offsets (used when offset_rand == 0)
  int32_t offset_rand;  // K > 0: offsets drawn from U{-K..K}
} ivqm_synth_spec;

// Generates frame pair `frame` of the workload. Planes are tightly packed
// (stride = plane width). out0/out1 receive the test and reference pictures.
// Samples are written as uint16 (u16 != nullptr) or uint8 (u8 != nullptr, bit
// depth 8 only); pass exactly one of the two families of pointers.
int ivqm_synth_pair(const ivqm_synth_spec* sp, int32_t frame, uint16_t* const out0_u16[3],
                    uint16_t* const out1_u16[3], uint8_t* const out0_u8[3], uint8_t* const out1_u8[3]) {
  if (!sp || sp->width <= 0 || sp->height <= 0 || sp->bit_depth < 8 || sp->bit_depth > 16) return -22;
  if (sp->chroma_420 && ((sp->width & 1) || (sp->height & 1))) return -22;
  const bool as_u8 = out0_u8 != nullptr;
  if (as_u8 && sp->bit_depth != 8) return -22;
  const int32_t W = sp->width, H = sp->height, Max = (1 << sp->bit_depth) - 1;
  const int cs = sp->chroma_420 ? 1 : 0;

  // Shift field: per-frame global shift, then one (dx, dy) per 32x32 luma block.
  Rng srng(stream_seed(sp->cfg, frame, 1, 3));
  int32_t gx = 0, gy = 0;
  if (sp->global_shift > 0) {
    gx = srng.uni_int(-sp->global_shift, sp->global_shift);
    gy = srng.uni_int(-sp->global_shift, sp->global_shift);
  }
  const int32_t BW = (W + 31) / 32, BH = (H + 31) / 32;
  std::vector<int32_t> bdx((size_t)BW * BH), bdy((size_t)BW * BH);
  for (int32_t i = 0; i < BW * BH; i++) {
    bdx[i] = sp->block_shift > 0 ? srng.uni_int(-sp->block_shift, sp->block_shift) : 0;
    bdy[i] = sp->block_shift > 0 ? srng.uni_int(-sp->block_shift, sp->block_shift) : 0;
  }

  for (int p = 0; p < 3; p++) {
    const int sx = p ? cs : 0, sy = p ? cs : 0;
    const int32_t Wp = W >> sx, Hp = H >> sy;
    // ---------- reference picture ----------
    Rng prng(stream_seed(sp->cfg, frame, 0, p));
    const double amp = p ? 0.5 : 1.0;
    const double a0 = p ? prng.uni(0.4, 0.6) * Max : prng.uni(0.2, 0.4) * Max;
    const double ax = amp * prng.uni(0.1, 0.3) * Max, ay = amp * prng.uni(0.1, 0.3) * Max;
    double A[8], fx[8], fy[8], ph[8];
    for (int k = 0; k < 8; k++) {
      A[k] = amp * prng.uni(0.01, 0.05) * Max;
      const double f = prng.uni(1.0 / 256.0, 1.0 / 16.0);
      const double t = prng.uni(0.0, 2.0 * M_PI);
      ph[k] = prng.uni(0.0, 2.0 * M_PI);
      fx[k] = 2.0 * M_PI * f * std::cos(t);
      fy[k] = 2.0 * M_PI * f * std::sin(t);
    }
    // separable evaluation: sin(a + b) = sin(a)cos(b) + cos(a)sin(b)
    std::vector<double> sxv((size_t)8 * Wp), cxv((size_t)8 * Wp);
    for (int k = 0; k < 8; k++)
      for (int32_t x = 0; x < Wp; x++) {
        sxv[(size_t)k * Wp + x] = std::sin(fx[k] * x);
        cxv[(size_t)k * Wp + x] = std::cos(fx[k] * x);
      }
    const uint64_t ref_seed = stream_seed(sp->cfg, frame, 0, p);
    std::vector<uint16_t> ref((size_t)Wp * Hp);
#pragma omp parallel for schedule(static)
    for (int32_t y = 0; y < Hp; y++) {
      Rng nrng(mix64(ref_seed ^ ((uint64_t)(y + 1) * 0x9E3779B97F4A7C15ull)));
      double sy_[8], cy_[8];
      for (int k = 0; k < 8; k++) { sy_[k] = std::sin(fy[k] * y + ph[k]); cy_[k] = std::cos(fy[k] * y + ph[k]); }
      for (int32_t x = 0; x < Wp; x++) {
        double v = a0 + ax * (double)x / (double)Wp + ay * (double)y / (double)Hp;
        for (int k = 0; k < 8; k++)
          v += A[k] * (sxv[(size_t)k * Wp + x] * cy_[k] + cxv[(size_t)k * Wp + x] * sy_[k]);
        v += sp->sigma_r * nrng.gauss();
        ref[(size_t)y * Wp + x] = (uint16_t)clampi((int32_t)round_half_away(v), 0, Max);
      }
    }
    // ---------- test picture ----------
    Rng trng(stream_seed(sp->cfg, frame, 1, p));
    const int32_t off = sp->offset_rand > 0 ? trng.uni_int(-sp->offset_rand, sp->offset_rand) : sp->offsets[p];
    const uint64_t tst_seed = stream_seed(sp->cfg, frame, 1, p);
    std::vector<uint16_t> tst((size_t)Wp * Hp);
#pragma omp parallel for schedule(static)
    for (int32_t y = 0; y < Hp; y++) {
      Rng nrng(mix64(tst_seed ^ ((uint64_t)(y + 1) * 0x9E3779B97F4A7C15ull) ^ 0x5A5A5A5A5A5A5A5Aull));
      const int32_t by = (y << sy) >> 5;
      for (int32_t x = 0; x < Wp; x++) {
        const int32_t bx = (x << sx) >> 5;
        const int32_t dx = bdx[(size_t)by * BW + bx] + gx, dy = bdy[(size_t)by * BW + bx] + gy;
        const int32_t rx = clampi(x + dx, 0, Wp - 1), ry = clampi(y + dy, 0, Hp - 1);
        double v = (double)ref[(size_t)ry * Wp + rx] + (double)off + sp->sigma_t * nrng.gauss();
        tst[(size_t)y * Wp + x] = (uint16_t)clampi((int32_t)round_half_away(v), 0, Max);
      }
    }
    const size_t n = (size_t)Wp * Hp;
    if (as_u8) {
      for (size_t i = 0; i < n; i++) { out0_u8[p][i] = (uint8_t)tst[i]; out1_u8[p][i] = (uint8_t)ref[i]; }
    } else {
      std::memcpy(out0_u16[p], tst.data(), n * sizeof(uint16_t));
      std::memcpy(out1_u16[p], ref.data(), n * sizeof(uint16_t));
    }
  }
  return 0;
}

}  // extern "C"
