"""Every structural-similarity mode, window, stride and border extension on the GPU
(k_ssim_gen, through the C ABI) against the FP64 oracle: SSIM, MS-SSIM, IV-SSIM with ERP
weights (both directions) and IV-MS-SSIM.

Modes (xSSIM::eMode, libs/pmbb-ivqm/src/xSSIM.h:24-31): 0 RegularGaussianFlt,
1 RegularGaussianInt, 2 RegularAveraged, 3 BlockGaussianInt, 4 BlockAveraged. Border
extensions (eMrgExt): 0 None, 1 Nearest, 2 Reflect, 3 Mirror, 4 Constant, 5 Zero.

Tolerance 1e-10: the kernel keeps every moment in FP64 (exact integers for integer
weights) and evaluates the ratio in FP64, so GPU and oracle differ only in the order of
the FP64 window sums."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth

pytestmark = pytest.mark.gpu

TOL = 1e-10
M = Q.IVQM_M_SSIM | Q.IVQM_M_MSSSIM | Q.IVQM_M_IVSSIM | Q.IVQM_M_IVMSSSIM

# (mode, ext, window, stride)
CONFIGS = [
    (0, 0, 11, 1), (0, 1, 11, 4), (0, 3, 11, 2),
    (1, 0, 11, 1), (1, 2, 11, 3), (1, 4, 11, 4),
    (2, 0, 11, 4), (2, 5, 11, 1), (2, 1, 11, 11),
    (3, 0, 8, 4), (3, 0, 16, 1), (3, 0, 8, 8),
    (4, 0, 16, 4), (4, 0, 32, 8), (4, 0, 8, 1), (4, 0, 8, 2), (4, 0, 5, 3),
    (0, 0, 11, 4), (1, 0, 11, 4),
]
WORKLOADS = {
    "10b420": synth.Workload(cfg=700, width=200, height=136, global_shift=1),
    "8b420": synth.Workload(cfg=701, width=160, height=128, bit_depth=8),
    "16b444": synth.Workload(cfg=702, width=144, height=112, bit_depth=16, chroma_420=False, sigma_r=12,
                             sigma_t=12, block_shift=2),
    # wider than one wave of stride-4 windows (62 columns of 11-wide windows) and taller than one wave's rows
    "10b420w": synth.Workload(cfg=704, width=532, height=290, global_shift=1, block_shift=2),
}


def _check(oracle, wl, mode, ext, wnd, stride):
    p0n, p1n = synth.make_pair(wl, 0)
    p0, p1 = oracle.to444(p0n, wl.chroma_420), oracle.to444(p1n, wl.chroma_420)
    bd = wl.bit_depth
    print(f"case {wl.width}x{wl.height} {bd}-bit mode {mode} ext {ext} wnd {wnd} stride {stride}", flush=True)
    with Q.IvqmContext(wl.width, wl.height, bd, erp=1, ssim_mode=mode, ssim_border=ext, ssim_window=wnd,
                       ssim_stride=stride) as ctx:
        ctx.upload_pair(p0n, p1n, wl.chroma_420)
        r = ctx.frame(M)
    g = oracle.gcd(p0, p1, bd)
    s0, s1 = oracle.scp(p0, p1, bd, g)
    args = (mode, ext, wnd, stride)
    errs = {
        "ssim": np.abs(np.array(r.ssim[:3]) - oracle.ssim_ex(p0, p1, bd, *args)).max(),
        "msssim": np.abs(np.array(r.msssim[:3]) - oracle.ssim_ex(p0, p1, bd, *args, ms=True)).max(),
    }
    v, r2t, t2r = oracle.ivssim_ex(p0, p1, s0, s1, bd, *args, erp=True)
    errs["ivssim"] = max(abs(r.ivssim - v), abs(r.ivssim_r2t - r2t), abs(r.ivssim_t2r - t2r))
    errs["ivmsssim"] = abs(r.ivmsssim - oracle.ivssim_ex(p0, p1, s0, s1, bd, *args, ms=True)[0])
    print("   ", {k: f"{e:.1e}" for k, e in errs.items()}, "ssim", list(r.ssim[:3]), flush=True)
    for k, e in errs.items():
        assert e <= TOL, (k, e)


@pytest.mark.parametrize("cfg", CONFIGS, ids=[f"m{m}-e{e}-w{w}-s{s}" for m, e, w, s in CONFIGS])
def test_ssim_modes_10bit_420(oracle, cfg):
    _check(oracle, WORKLOADS["10b420"], *cfg)


@pytest.mark.parametrize("cfg", [(0, 2, 11, 3), (1, 1, 11, 1), (3, 0, 16, 4), (4, 0, 16, 16)])
def test_ssim_modes_8bit(oracle, cfg):
    _check(oracle, WORKLOADS["8b420"], *cfg)


@pytest.mark.parametrize("cfg", [(1, 4, 11, 2), (2, 3, 11, 4), (3, 0, 8, 2), (4, 0, 32, 4), (1, 0, 11, 4),
                                 (3, 0, 16, 4)])
def test_ssim_modes_16bit_444(oracle, cfg):
    _check(oracle, WORKLOADS["16b444"], *cfg)


@pytest.mark.parametrize("cfg", [(1, 0, 11, 4), (0, 0, 11, 4), (2, 0, 11, 4), (3, 0, 8, 4), (3, 0, 16, 4),
                                 (4, 0, 16, 4)])
def test_ssim_modes_stride4_several_waves(oracle, cfg):
    """Stride-4 windows without a border extension run the rolling-row kernel (one vector load per sample
    row and lane, neighbours by DPP, window rows walked with rolling accumulators): a picture several
    waves wide and tall, with a ragged right edge."""
    _check(oracle, WORKLOADS["10b420w"], *cfg)


def test_picture_smaller_than_window_is_rejected():
    wl = synth.Workload(cfg=703, width=16, height=16)
    p0n, p1n = synth.make_pair(wl, 0)
    with Q.IvqmContext(16, 16, 10, ssim_mode=4, ssim_window=32, ssim_stride=8) as ctx:
        ctx.upload_pair(p0n, p1n, True)
        with pytest.raises(Q.IvqmError):
            ctx.frame(Q.IVQM_M_SSIM)
