"""SSIM kernel variants against the FP64 oracle: 32-bit vs 64-bit squared sums at
10-bit, and 16-bit content at small and full amplitude, 4:4:4 and 4:2:0."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth

pytestmark = pytest.mark.gpu


def _ssim_err(oracle, wl, **kw):
    p0n, p1n = synth.make_pair(wl, 0)
    p0, p1 = oracle.to444(p0n, wl.chroma_420), oracle.to444(p1n, wl.chroma_420)
    with Q.IvqmContext(wl.width, wl.height, wl.bit_depth, **kw) as ctx:
        ctx.upload_pair(p0n, p1n, wl.chroma_420)
        r = ctx.frame(Q.IVQM_M_SSIM)
    want = oracle.ssim(p0, p1, wl.bit_depth)
    err = np.abs(np.array(r.ssim[:3]) - want)
    print(wl, kw, "ssim gpu", list(r.ssim[:3]), "oracle", want, "err", err, flush=True)
    return err


@pytest.mark.parametrize("wide", [0, 1])
@pytest.mark.parametrize("c420", [True, False])
def test_ssim_sum_width_10bit(oracle, wide, c420):
    wl = synth.Workload(cfg=600, width=160, height=96, chroma_420=c420)
    assert np.max(_ssim_err(oracle, wl, ssim_wide_sums=wide)) <= 1e-6


@pytest.mark.parametrize("sigma,c420", [(4.0, False), (12.0, False), (12.0, True)])
def test_ssim_16bit(oracle, sigma, c420):
    wl = synth.Workload(cfg=601, width=96, height=80, bit_depth=16, chroma_420=c420, sigma_r=sigma, sigma_t=sigma,
                        block_shift=2)
    assert np.max(_ssim_err(oracle, wl)) <= 1e-6
