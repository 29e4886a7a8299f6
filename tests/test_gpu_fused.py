"""Fused IV-SSIM: in batches of two or more pairs the search writes exact 4x4
sub-block moments of (test picture, SCP) instead of the SCP pictures, and the
block-averaged 8 / 4 SSIM pass reads those. The integers are the ones the SSIM
pass sums from the pictures, so every result must equal the single-pair (SCP)
path's bit for bit, and IV-SSIM stays within 1e-6 of the FP64 oracle."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth
from qmiv_amd.context import IvqmBatch

pytestmark = pytest.mark.gpu

IV = Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM
SSIM_TOL = 1e-6

# sizes that leave partial search tiles and sub-block rows / columns past the last window
PAIRS = [
    synth.Workload(cfg=600, width=128, height=64, bit_depth=8),
    synth.Workload(cfg=601, width=200, height=96),
    synth.Workload(cfg=602, width=72, height=66, chroma_420=False),
    synth.Workload(cfg=603, width=332, height=150, global_shift=1),
    synth.Workload(cfg=604, width=1028, height=518, block_shift=2),
    synth.Workload(cfg=605, width=96, height=80, bit_depth=12, sigma_r=20, sigma_t=15),  # reference form: SCP path
]


def _to_dev(wl, p0, p1):
    import torch
    dev = torch.device("cuda:0")
    u8 = wl.bit_depth == 8
    cv = (lambda a: torch.from_numpy(a).to(dev)) if u8 else (lambda a: torch.from_numpy(a.view(np.int16)).to(dev))
    return ([cv(p) for p in p0], [cv(p) for p in p1], wl.width, wl.height, wl.bit_depth, wl.chroma_420)


def _host(wls, frame=0):
    out = []
    for wl in wls:
        p0, p1 = synth.make_pair(wl, frame, as_u8=wl.bit_depth == 8)
        out.append((wl, p0, p1))
    return out


def _single(wl, p0, p1, metrics, **kw):
    with Q.IvqmContext(wl.width, wl.height, wl.bit_depth, **kw) as ctx:
        ctx.upload_pair(p0, p1, wl.chroma_420)
        return ctx.frame(metrics).as_dict()


def _batch(host, metrics, use_graph=False, reps=1, **kw):
    import torch
    items = [_to_dev(wl, p0, p1) for wl, p0, p1 in host]
    torch.cuda.synchronize()
    res = []
    with IvqmBatch(**kw) as b:
        b.set_tuning(0, 0, use_graph)
        pairs = b.make_pairs(items)
        for _ in range(reps):
            res.append([r.as_dict() for r in b.run(pairs, metrics)])
    return res


@pytest.mark.parametrize("use_graph", [False, True])
def test_fused_batch_equals_single_pair_path(use_graph):
    host = _host(PAIRS)
    want = [_single(wl, p0, p1, IV) for wl, p0, p1 in host]
    for rep, got in enumerate(_batch(host, IV, use_graph=use_graph, reps=3)):
        for i, (g, w) in enumerate(zip(got, want)):
            print(PAIRS[i].width, PAIRS[i].height, PAIRS[i].bit_depth, "IV-SSIM", g["ivssim"], flush=True)
            assert g == w, (rep, i, PAIRS[i])


@pytest.mark.parametrize("R,cap", [(1, 0), (2, 2), (3, 0), (4, 0)], ids=["R1", "R2-narrow", "R3", "R4"])
def test_fused_search_ranges_and_variants(R, cap):
    host = _host(PAIRS[:4])
    kw = dict(search_range=R, search_mode_cap=cap)
    want = [_single(wl, p0, p1, IV, **kw) for wl, p0, p1 in host]
    got = _batch(host, IV, **kw)[0]
    for i, (g, w) in enumerate(zip(got, want)):
        assert g == w, (R, cap, i)


def test_fused_erp_weights():
    host = _host(PAIRS[1:5])
    want = [_single(wl, p0, p1, IV, erp=1) for wl, p0, p1 in host]
    got = _batch(host, IV, erp=1)[0]
    assert got == want


def test_fused_ivssim_against_oracle(oracle):
    host = _host(PAIRS[:4])
    got = _batch(host, IV)[0]
    for (wl, p0n, p1n), r in zip(host, got):
        bd = wl.bit_depth
        p0 = oracle.to444([np.asarray(p, np.uint16) for p in p0n], wl.chroma_420)
        p1 = oracle.to444([np.asarray(p, np.uint16) for p in p1n], wl.chroma_420)
        g = oracle.gcd(p0, p1, bd, wrap32=False)
        _, scp1 = oracle.cps(p0, p1, bd, g, R=2, want_scp=True)[:2]
        _, scp0 = oracle.cps(p1, p0, bd, [-x for x in g], R=2, want_scp=True)[:2]
        iv, ir, it = oracle.ivssim(p0, p1, scp0, scp1, bd)
        err = max(abs(r["ivssim"] - iv), abs(r["ivssim_r2t"] - ir), abs(r["ivssim_t2r"] - it))
        print(wl.width, wl.height, bd, "fused IV-SSIM |err| vs FP64 oracle", f"{err:.2e}", flush=True)
        assert err <= SSIM_TOL


# IV-MS-SSIM in fused batches (k_search420 pairs: 4:2:0 up to 12-bit, keyed): the search writes the moments
# and MS-SSIM level 1 of both SCP pictures; level 0 is the IV-SSIM jobs' twin. Odd sizes from level 1 on
# truncate (xSSIM.cpp:65).
PAIRS_MS = PAIRS + [synth.Workload(cfg=606, width=204, height=130, block_shift=2),
                    synth.Workload(cfg=607, width=588, height=530, bit_depth=8, global_shift=1)]
IVMS = IV | Q.IVQM_M_IVMSSSIM


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("metrics", [IVMS, Q.IVQM_M_ALL], ids=["ivms", "all"])
def test_fused_ivmsssim_equals_single_pair_path(metrics, use_graph):
    host = _host(PAIRS_MS)
    want = [_single(wl, p0, p1, metrics) for wl, p0, p1 in host]
    for rep, got in enumerate(_batch(host, metrics, use_graph=use_graph, reps=2)):
        for i, (g, w) in enumerate(zip(got, want)):
            print(PAIRS_MS[i].width, PAIRS_MS[i].height, PAIRS_MS[i].bit_depth, "IV-MS-SSIM", g["ivmsssim"],
                  flush=True)
            assert g == w, (rep, i, PAIRS_MS[i])


def test_fused_ivmsssim_against_oracle(oracle):
    host = _host([PAIRS_MS[1], PAIRS_MS[3], PAIRS_MS[6], PAIRS_MS[7]])
    got = _batch(host, IVMS, erp=1)[0]
    for (wl, p0n, p1n), r in zip(host, got):
        bd = wl.bit_depth
        p0 = oracle.to444([np.asarray(p, np.uint16) for p in p0n], wl.chroma_420)
        p1 = oracle.to444([np.asarray(p, np.uint16) for p in p1n], wl.chroma_420)
        g = oracle.gcd(p0, p1, bd, wrap32=False)
        _, scp1 = oracle.cps(p0, p1, bd, g, R=2, want_scp=True)[:2]
        _, scp0 = oracle.cps(p1, p0, bd, [-x for x in g], R=2, want_scp=True)[:2]
        iv = oracle.ivssim(p0, p1, scp0, scp1, bd, erp=True)[0]
        ivms = oracle.ivmsssim(p0, p1, scp0, scp1, bd)
        errs = (abs(r["ivssim"] - iv), abs(r["ivmsssim"] - ivms))
        print(wl.width, wl.height, bd, "fused IV-SSIM / IV-MS-SSIM |err| vs FP64 oracle",
              " ".join(f"{e:.2e}" for e in errs), flush=True)
        assert max(errs) <= SSIM_TOL
