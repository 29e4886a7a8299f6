// Host-only check of the batch engine's arena bounds rule (ivqm::resolve_picture): no HIP call is
// made, so it runs without a GPU. Prints one line per case and exits non-zero on the first mismatch.
#include <cstdio>
#include <cstdlib>

#include "../../qmiv_amd/csrc/ivqm_engine.h"

using ivqm::PicDev;
using ivqm::PicRef;

static int fails = 0;

static void expect(const char* name, const PicRef& r, uint16_t* scp, int64_t cap_scp, uint16_t* pyr, int64_t cap_pyr,
                   int want) {
  PicDev out{};
  const char* why = "";
  const int rc = ivqm::resolve_picture(r, scp, cap_scp, pyr, cap_pyr, out, &why);
  std::printf("%-44s rc=%d %s\n", name, rc, rc ? why : "");
  if (rc != want) fails++;
  if (rc == IVQM_OK && (r.arena == 1 || r.arena == 2)) {
    uint16_t* base = r.arena == 1 ? scp : pyr;
    const int64_t cap = r.arena == 1 ? cap_scp : cap_pyr;
    for (int k = 0; k < 3; k++) {
      const uint16_t* p = static_cast<const uint16_t*>(out.p[k]);
      if (k && r.chroma_from_user) continue;
      // the last element of every plane lies inside the arena
      if (p < base || p + (int64_t)out.stride[k] * r.h > base + cap) {
        std::printf("  plane %d outside the arena\n", k);
        fails++;
      }
    }
  }
}

int main() {
  static uint16_t scp[3 * 64 * 16], pyr[3 * 64 * 8 + 100];
  const int64_t cap_scp = sizeof(scp) / 2, cap_pyr = sizeof(pyr) / 2;
  PicRef r;
  r.arena = 1; r.pitch = 64; r.w = 60; r.h = 16; r.off = 0;
  expect("SCP picture filling the arena", r, scp, cap_scp, pyr, cap_pyr, IVQM_OK);
  r.off = 1;
  expect("SCP picture one element past the end", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  r.off = -64;
  expect("negative offset", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  r.off = 0; r.w = 65;
  expect("width above the pitch", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  r.w = 60;
  expect("arena not allocated", r, nullptr, 0, pyr, cap_pyr, IVQM_EINVAL);
  r.arena = 2; r.h = 8; r.off = 100;
  expect("pyramid picture at the end of its arena", r, scp, cap_scp, pyr, cap_pyr, IVQM_OK);
  r.off = 101;
  expect("pyramid picture past its arena", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  r.off = 0; r.h = 16;
  expect("pyramid picture of SCP height (grown plan)", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  r.h = 8;
  r.chroma_from_user = true;
  r.user = PicDev{};
  r.user.cs = 1; r.user.csy = 1; r.user.bps = 2;
  expect("level 1 with the caller's chroma", r, scp, cap_scp, pyr, cap_pyr, IVQM_OK);
  r.user.csy = 0;
  expect("caller's chroma taken from a 4:2:2 picture", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  r.user.cs = 0;
  expect("caller's chroma taken from a 4:4:4 picture", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  r.arena = 7;
  expect("unknown arena", r, scp, cap_scp, pyr, cap_pyr, IVQM_EINVAL);
  std::printf(fails ? "FAIL %d\n" : "OK\n", fails);
  return fails ? 1 : 0;
}
