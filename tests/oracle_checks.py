"""Checks one frame-pair result of the HIP path against the CPU oracle.

PSNR, WS-PSNR and IV-PSNR must be identical doubles (``repr``-equal), the GCD
identical integers; SSIM, MS-SSIM, IV-SSIM and IV-MS-SSIM within ``tol``
absolute of the FP64 oracle. Returns the SSIM-family errors so the caller can
report them per metric.
"""
from __future__ import annotations

import numpy as np

import qmiv_amd as Q

SSIM_TOL = 1e-6  # SSIM family vs the FP64 oracle (integer moments, FP32 ratios, FP64 totals)


def to444(oracle, planes, chroma_420):
    return oracle.to444([np.ascontiguousarray(p).astype(np.uint16) for p in planes], chroma_420)


def check_pair(oracle, r, p0n, p1n, bit_depth, chroma_420=True, metrics=Q.IVQM_M_ALL, erp=False, R=2,
               wrap32=False, tol=SSIM_TOL, label=""):
    """r: FrameResult (or its as_dict()); p0n / p1n: the pair's planes as delivered (4:2:0 or 4:4:4)."""
    d = r if isinstance(r, dict) else r.as_dict()
    p0 = to444(oracle, p0n, chroma_420)
    p1 = to444(oracle, p1n, chroma_420)
    bd = bit_depth
    errs = {}
    g = oracle.gcd(p0, p1, bd, wrap32=wrap32)
    if metrics & (Q.IVQM_M_PSNR | Q.IVQM_M_WSPSNR | Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM | Q.IVQM_M_IVMSSSIM):
        assert d["gcd_r2t"] == g, (label, d["gcd_r2t"], g)
    if metrics & Q.IVQM_M_PSNR:
        assert not any(d["exact"][:3]), label  # synthetic pairs are never identical
        assert [repr(x) for x in d["psnr"][:3]] == [repr(x) for x in oracle.psnr(p0, p1, bd)], label
    if metrics & Q.IVQM_M_WSPSNR:
        assert [repr(x) for x in d["wspsnr"][:3]] == [repr(x) for x in oracle.wspsnr(p0, p1, bd, erp=erp)], label
    if metrics & Q.IVQM_M_IVPSNR:
        want = oracle.ivpsnr(p0, p1, bd, g, R=R, erp=erp)
        got = (d["ivpsnr"], d["ivpsnr_r2t"], d["ivpsnr_t2r"])
        assert tuple(repr(x) for x in got) == tuple(repr(x) for x in want), (label, got, want)
    if metrics & Q.IVQM_M_SSIM:
        errs["ssim"] = float(np.max(np.abs(np.array(d["ssim"][:3]) - oracle.ssim(p0, p1, bd))))
    if metrics & Q.IVQM_M_MSSSIM:
        errs["msssim"] = float(np.max(np.abs(np.array(d["msssim"][:3]) - oracle.msssim(p0, p1, bd))))
    if metrics & (Q.IVQM_M_IVSSIM | Q.IVQM_M_IVMSSSIM):
        s0, s1 = oracle.scp(p0, p1, bd, g, R=R)
        if metrics & Q.IVQM_M_IVSSIM:
            v, a, b = oracle.ivssim(p0, p1, s0, s1, bd, erp=erp)
            errs["ivssim"] = max(abs(d["ivssim"] - v), abs(d["ivssim_r2t"] - a), abs(d["ivssim_t2r"] - b))
        if metrics & Q.IVQM_M_IVMSSSIM:
            errs["ivmsssim"] = abs(d["ivmsssim"] - oracle.ivmsssim(p0, p1, s0, s1, bd))
    for k, e in errs.items():
        assert e <= tol, (label, k, e)
    return errs


def merge_worst(worst, errs):
    for k, e in errs.items():
        worst[k] = max(worst.get(k, 0.0), e)
    return worst
