"""Large-shift stress workload at full size: 8192x4096, 16-bit samples over
[0, 65535], blocks shifted by up to 3 px, offsets up to 8 LSB, noise sigma 12,
search range 4 (9x9 candidates), 4:4:4 and 4:2:0. IV-PSNR (and the GCD and row
sums behind it) must equal the oracle bit for bit; IV-SSIM and MS-SSIM use exact
integer moments with FP32 ratios and must stay within 1e-6 of the FP64 oracle
(the error is printed per metric)."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("wl", [synth.STRESS_8K_444, synth.STRESS_8K_420], ids=["444", "420"])
def test_stress_8k_16bit_r4(oracle, wl):
    R = 4
    p0n, p1n = synth.make_pair(wl, 0)
    m = Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM | Q.IVQM_M_MSSSIM
    with Q.IvqmContext(wl.width, wl.height, 16, search_range=R, erp=1) as ctx:
        ctx.upload_pair(p0n, p1n, wl.chroma_420)
        r = ctx.frame(m)
        rows = [ctx.rows(0), ctx.rows(1)]
    p0 = oracle.to444(p0n, wl.chroma_420)
    p1 = oracle.to444(p1n, wl.chroma_420)
    g = oracle.gcd(p0, p1, 16)
    assert list(r.gcd_r2t) == g
    erp_w = True
    rows_a, _, _ = oracle.cps(p0, p1, 16, g, R=R)
    rows_b, _, _ = oracle.cps(p1, p0, 16, [-x for x in g], R=R)
    assert np.array_equal(rows[0][:, :3], rows_a[:, :3])
    assert np.array_equal(rows[1][:, :3], rows_b[:, :3])
    want = oracle.ivpsnr(p0, p1, 16, g, R=R, erp=erp_w)
    assert (repr(r.ivpsnr), repr(r.ivpsnr_r2t), repr(r.ivpsnr_t2r)) == tuple(repr(x) for x in want)
    s0, s1 = oracle.scp(p0, p1, 16, g, R=R)
    iv = oracle.ivssim(p0, p1, s0, s1, 16, erp=True)
    ms = oracle.msssim(p0, p1, 16)
    errs = {"ivssim": abs(r.ivssim - iv[0]), "ivssim_r2t": abs(r.ivssim_r2t - iv[1]),
            "ivssim_t2r": abs(r.ivssim_t2r - iv[2]),
            "msssim": float(np.max(np.abs(np.array(r.msssim[:3]) - np.array(ms))))}
    print(wl.chroma_420 and "420" or "444", "SSIM-family |err| vs FP64 oracle:",
          {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    assert max(errs.values()) <= 1e-6
