"""Writes tests/golden/golden_vectors.json from the reference library itself.

The reference pmbb libraries are compiled from their own sources by
oracle/Makefile into oracle/_ref/ (this script needs them; run
`make -C oracle` first). Inputs are regenerated from the seeded generator in
qmiv_amd/csrc/synth.cpp, so the fixture stores only the generator spec, input
checksums (plus the first and last 16 samples of every plane, to detect
generator drift) and the reference's outputs. Large integer arrays are stored
as SHA-1 digests; floating-point results as exact repr strings.

Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import dataclasses
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle.oracle import Oracle, Reference  # noqa: E402
from qmiv_amd import synth  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden_vectors.json")

# (name, workload, frame, R, erp) -- small shapes, all bit depths, both chroma
# formats, an ERP case large enough for 5-scale MS-SSIM, and an adversarial
# 16-bit case whose squared differences wrap int32.
CASES = [
    ("w64h48_10b_420", synth.Workload(cfg=100, width=64, height=48), 0, 2, False),
    ("w96h64_8b_420", synth.Workload(cfg=101, width=96, height=64, bit_depth=8), 1, 2, False),
    ("w128h96_14b_420", synth.Workload(cfg=102, width=128, height=96, bit_depth=14, sigma_r=40, sigma_t=30), 0, 2, False),
    ("w80h72_16b_444_r4", synth.Workload(cfg=103, width=80, height=72, bit_depth=16, chroma_420=False, sigma_r=12,
                                         sigma_t=12, block_shift=3, offset_rand=8), 0, 4, False),
    ("w72h64_10b_444_r1", synth.Workload(cfg=104, width=72, height=64, chroma_420=False), 2, 1, False),
    ("w64h64_16b_444_wrap", synth.Workload(cfg=105, width=64, height=64, bit_depth=16, chroma_420=False,
                                           sigma_r=30000, sigma_t=30000, block_shift=2, offset_rand=600), 0, 2, False),
    ("w96h64_10b_420_r3", synth.Workload(cfg=106, width=96, height=64, block_shift=3, sigma_t=5), 0, 3, True),
    ("w512h512_10b_420_erp", synth.Workload(cfg=107, width=512, height=512, global_shift=2), 0, 2, True),
    ("w1920h1080_10b_420", synth.CPU_1080P, 0, 2, False),
]


def sha1(a) -> str:
    return hashlib.sha1(np.ascontiguousarray(a).tobytes()).hexdigest()


def f(x) -> str:
    return repr(float(x))


def make_case(name, wl, frame, R, erp, ref: Reference, orc: Oracle):
    p0n, p1n = synth.make_pair(wl, frame)
    p0 = orc.to444(p0n, wl.chroma_420)
    p1 = orc.to444(p1n, wl.chroma_420)
    bd = wl.bit_depth
    case = {
        "name": name, "workload": dataclasses.asdict(wl), "frame": frame, "R": R, "erp": erp,
        "inputs": {
            "p0": [sha1(p) for p in p0n], "p1": [sha1(p) for p in p1n],
            "p0_head": [p.ravel()[:16].tolist() for p in p0n], "p0_tail": [p.ravel()[-16:].tolist() for p in p0n],
            "p1_head": [p.ravel()[:16].tolist() for p in p1n], "p1_tail": [p.ravel()[-16:].tolist() for p in p1n],
        },
    }
    up = [ref.upsample(p1n[1])] if wl.chroma_420 else []
    case["upsample_cb1_sha1"] = sha1(up[0]) if up else None
    case["downsample_y0_sha1"] = sha1(ref.downsample(p0[0]))
    g = ref.gcd(p0, p1, bd, R=R)
    case["gcd_r2t"] = g
    case["sd"] = [int(ref.sd(p1[c], p0[c], bd)) for c in range(3)]
    case["ssd"] = [int(ref.ssd(p0[c], p1[c], bd)) for c in range(3)]
    rows_r2t = ref.cps_rows(p0, p1, bd, g, R=R)
    rows_t2r = ref.cps_rows(p1, p0, bd, [-x for x in g], R=R)
    h = wl.height
    if h <= 96:
        case["rows_r2t"] = rows_r2t.astype(np.uint64).tolist()
        case["rows_t2r"] = rows_t2r.astype(np.uint64).tolist()
    case["rows_r2t_sha1"] = sha1(rows_r2t.astype(np.uint64))
    case["rows_t2r_sha1"] = sha1(rows_t2r.astype(np.uint64))
    case["rows_r2t_total"] = [int(x) for x in rows_r2t.sum(axis=0, dtype=np.uint64)]
    case["rows_t2r_total"] = [int(x) for x in rows_t2r.sum(axis=0, dtype=np.uint64)]
    s0, s1 = ref.scp(p0, p1, bd, g, R=R)
    case["scp0_sha1"] = [sha1(p) for p in s0]
    case["scp1_sha1"] = [sha1(p) for p in s1]
    case["psnr"] = [f(x) for x in ref.psnr(p0, p1, bd)]
    case["wspsnr_erp"] = [f(x) for x in ref.wspsnr(p0, p1, bd, erp=True)]
    v, a, b = ref.ivpsnr(p0, p1, bd, g, R=R, erp=False)
    case["ivpsnr"] = [f(v), f(a), f(b)]
    v, a, b = ref.ivpsnr(p0, p1, bd, g, R=R, erp=True)
    case["ivpsnr_erp"] = [f(v), f(a), f(b)]
    if bd <= 14:  # the reference's SIMD SSIM kernels use signed 16-bit madd (valid up to 15-bit samples)
        case["ssim"] = [f(x) for x in ref.ssim(p0, p1, bd)]
        case["msssim"] = [f(x) for x in ref.ssim(p0, p1, bd, ms=True)]
        case["ivssim"] = f(ref.ivssim(p0, p1, s0, s1, bd, erp=False))
        case["ivssim_erp"] = f(ref.ivssim(p0, p1, s0, s1, bd, erp=True))
        case["ivmsssim"] = f(ref.ivssim(p0, p1, s0, s1, bd, ms=True))
    return case


# Structural-similarity modes (xSSIM::eMode: 0 RegularGaussianFlt, 1 RegularGaussianInt,
# 2 RegularAveraged, 3 BlockGaussianInt, 4 BlockAveraged) and border extensions (eMrgExt:
# 0 None, 1 Nearest, 2 Reflect, 3 Mirror, 4 Constant, 5 Zero), from the x86-64-v3 build:
# on x86-64-v4 plain SSIM in BlockGaussianInt 8/4 and 8/8 runs the multi-block averaging
# kernels (xSSIM.cpp:217-226 selects them by window and stride only). Pictures carry a
# margin of 8 so that the 11x11 windows of the extension modes stay inside it.
MODE_WORKLOADS = {
    "w112h80_10b_444": synth.Workload(cfg=130, width=112, height=80, chroma_420=False),
    "w128h96_8b_420": synth.Workload(cfg=131, width=128, height=96, bit_depth=8, global_shift=1),
}
MODE_CASES = (
    [(0, e, 11, s) for e in (0, 1, 2, 3, 4, 5) for s in (1, 3)]
    + [(1, e, 11, s) for e in (0, 1, 2, 3, 4, 5) for s in (1, 4)]
    # RegularAveraged with Constant extension is left out: the reference squares the
    # Constant value 65535 in int32 (xStructSimSTD.cpp:172-174), which wraps
    + [(2, e, 11, s) for e in (0, 1, 2, 3, 5) for s in (1, 11)]
    + [(3, 0, 8, s) for s in (1, 4, 8)] + [(3, 0, 16, s) for s in (3, 16)]
    + [(4, 0, 8, s) for s in (1, 2, 8)] + [(4, 0, 16, s) for s in (4, 8)] + [(4, 0, 32, 8)]
)


def make_mode_cases(orc: Oracle):
    ref = Reference("v3")
    out = []
    for wname, wl in MODE_WORKLOADS.items():
        p0n, p1n = synth.make_pair(wl, 0)
        p0 = orc.to444(p0n, wl.chroma_420)
        p1 = orc.to444(p1n, wl.chroma_420)
        bd = wl.bit_depth
        g = ref.gcd(p0, p1, bd)
        s0, s1 = ref.scp(p0, p1, bd, g)
        for mode, ext, wnd, stride in MODE_CASES:
            if wname.startswith("w128") and stride == 1 and mode != 2:
                continue  # keep the fixture small: stride 1 on the first picture only (plus averaged)
            print("mode case", wname, mode, ext, wnd, stride, flush=True)
            case = {"workload": wname, "mode": mode, "ext": ext, "wnd": wnd, "stride": stride,
                    "ssim": [f(x) for x in ref.ssim_ex(p0, p1, bd, mode, ext, wnd, stride)]}
            iv = ref.ivssim_ex(p0, p1, s0, s1, bd, mode, ext, wnd, stride, erp=True)
            case["ivssim_erp"] = [f(x) for x in iv]
            if stride > 1:
                case["msssim"] = [f(x) for x in ref.ssim_ex(p0, p1, bd, mode, ext, wnd, stride, ms=True)]
                case["ivmsssim"] = f(ref.ivssim_ex(p0, p1, s0, s1, bd, mode, ext, wnd, stride, ms=True)[0])
            out.append(case)
    return out


# Mask mode (apps/QMIV/src/xAppQMIV.cpp:902-906, :930-1029): GCD, MSE / PSNR / WS-PSNR and
# IV-PSNR with a weight mask (qmiv_amd/synth.py make_mask), bit-exact doubles from the reference.
MASK_CASES = [
    ("m96x72_10b_420_m10", synth.Workload(cfg=140, width=96, height=72, global_shift=2), 10, 2),
    ("m80x64_8b_420_m8_r3", synth.Workload(cfg=141, width=80, height=64, bit_depth=8), 8, 3),
    ("m72x64_16b_444_m14_wrap", synth.Workload(cfg=142, width=72, height=64, bit_depth=16, chroma_420=False,
                                               sigma_r=30000, sigma_t=30000, block_shift=2), 14, 2),
    ("m96x64_10b_420_m12_r4", synth.Workload(cfg=143, width=96, height=64), 12, 4),
    ("m128x96_12b_444_m8_r1", synth.Workload(cfg=144, width=128, height=96, bit_depth=12, chroma_420=False), 8, 1),
]


def make_mask_cases(ref: Reference, orc: Oracle):
    out = []
    for name, wl, bdm, R in MASK_CASES:
        print("mask case", name, flush=True)
        p0n, p1n = synth.make_pair(wl, 0)
        p0, p1 = orc.to444(p0n, wl.chroma_420), orc.to444(p1n, wl.chroma_420)
        bd = wl.bit_depth
        m = synth.make_mask(wl.width, wl.height, bdm, seed=wl.cfg)
        g = ref.gcd_masked(p0, p1, m, bd, bdm, R=R)
        case = {"name": name, "workload": dataclasses.asdict(wl), "bit_depth_m": bdm, "R": R,
                "mask_sha1": sha1(m), "nnm": int(np.count_nonzero(m)), "gcd": g}
        for erp in (False, True):
            mse, psnr, ws = ref.psnr_family_masked(p0, p1, m, bd, bdm, erp=erp)
            iv = ref.ivpsnr_masked(p0, p1, m, bd, bdm, g, R=R, erp=erp)
            sfx = "_erp" if erp else ""
            case["mse" + sfx] = [f(x) for x in mse]
            case["psnr" + sfx] = [f(x) for x in psnr]
            case["wspsnr" + sfx] = [f(x) for x in ws]
            case["ivpsnr" + sfx] = [f(x) for x in iv]
        # GenShftCompPicsM with the masked GCD; masked pixels keep the destination's 0
        s0, s1 = ref.scp_masked(p0, p1, m, bd, g, R=R, fill=0)
        case["scp0_sha1"] = [sha1(p) for p in s0]
        case["scp1_sha1"] = [sha1(p) for p in s1]
        out.append(case)
    return out


# Colour-space conversion (xColorSpace::ConvertYCbCr2RGB / ConvertRGB2YCbCr, fixed point): the
# planes of a 4:4:4 synthetic picture at 8..14 bits, each direction and colour space; SHA-1 of the
# converted planes. 16-bit conversions overflow int32 in the reference and its builds disagree.
COLOUR_WORKLOADS = [synth.Workload(cfg=190 + bd, width=96, height=40, bit_depth=bd, chroma_420=False,
                                   sigma_r=(1 << bd) / 8, sigma_t=(1 << bd) / 16) for bd in (8, 10, 12, 14)]


def make_colour_cases(ref: Reference):
    out = []
    for wl in COLOUR_WORKLOADS:
        p0, _ = synth.make_pair(wl, 0)
        for direction in (0, 1):
            for space in range(4):
                conv = ref.convert_colour(p0, wl.bit_depth, direction, space)
                out.append({"workload": dataclasses.asdict(wl), "direction": direction, "space": space,
                            "sha1": [sha1(c) for c in conv], "head": [c.ravel()[:8].tolist() for c in conv]})
    return out


def main():
    ref = Reference()
    orc = Oracle()
    cases = []
    for c in CASES:
        print("case", c[0], flush=True)
        cases.append(make_case(*c, ref=ref, orc=orc))
    doc = {
        "generator": "qmiv_amd/csrc/synth.cpp",
        "reference_build": f"oracle/Makefile (x86-64-{ref.level}) over /root/reference/libs",
        "cases": cases,
        "ssim_mode_workloads": {k: dataclasses.asdict(v) for k, v in MODE_WORKLOADS.items()},
        "ssim_mode_reference_build": "oracle/Makefile (x86-64-v3), pictures with a margin of 8",
        "ssim_mode_cases": make_mode_cases(orc),
        "mask_cases": make_mask_cases(ref, orc),
        "colour_cases": make_colour_cases(ref),
    }
    with open(OUT, "w") as fh:
        json.dump(doc, fh, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
