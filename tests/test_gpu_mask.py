"""Mask mode on the GPU through the C ABI (apps/QMIV/src/xAppQMIV.cpp:930-1029): GCD,
MSE / PSNR / WS-PSNR and IV-PSNR with a weight mask, bit-exact against the reference's golden
values and the oracle; shift-compensated pictures with a mask (GenShftCompPicsM); masked and
unmasked pairs in one batch; the combinations the app rejects."""
import json
import os

import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
M4 = Q.IVQM_M_MSE | Q.IVQM_M_PSNR | Q.IVQM_M_WSPSNR | Q.IVQM_M_IVPSNR


def _golden_mask_cases():
    with open(os.path.join(ROOT, "tests", "golden", "golden_vectors.json")) as fh:
        return json.load(fh)["mask_cases"]


def _wl(case):
    w = dict(case["workload"])
    w["offsets"] = tuple(w["offsets"])
    return synth.Workload(**w)


def _r(x):
    return repr(float(x))


@pytest.mark.parametrize("case", _golden_mask_cases(), ids=lambda c: c["name"])
@pytest.mark.parametrize("erp", [False, True])
def test_mask_matches_reference_golden(case, erp):
    wl = _wl(case)
    bd, bdm = wl.bit_depth, case["bit_depth_m"]
    print("case", case["name"], "erp", erp, flush=True)
    p0n, p1n = synth.make_pair(wl, 0)
    m = synth.make_mask(wl.width, wl.height, bdm, seed=wl.cfg)
    with Q.IvqmContext(wl.width, wl.height, bd, erp=int(erp), search_range=case["R"]) as ctx:
        ctx.upload_pair(p0n, p1n, wl.chroma_420)
        ctx.upload_mask(m if bdm > 8 else m.astype(np.uint8), bdm)
        r = ctx.frame(M4)
    sfx = "_erp" if erp else ""
    assert list(r.gcd_r2t) == case["gcd"]
    assert [_r(x) for x in r.mse[:3]] == case["mse" + sfx]
    assert [_r(x) for x in r.psnr[:3]] == case["psnr" + sfx]
    assert [_r(x) for x in r.wspsnr[:3]] == case["wspsnr" + sfx]
    assert [_r(r.ivpsnr), _r(r.ivpsnr_r2t), _r(r.ivpsnr_t2r)] == case["ivpsnr" + sfx]


@pytest.mark.parametrize("c420,bdm,R", [(True, 10, 2), (False, 12, 3)])
def test_mask_against_oracle_640x360(oracle, c420, bdm, R):
    wl = synth.Workload(cfg=160, width=640, height=360, chroma_420=c420, global_shift=1)
    bd = wl.bit_depth
    p0n, p1n = synth.make_pair(wl, 0)
    p0, p1 = oracle.to444(p0n, c420), oracle.to444(p1n, c420)
    m = synth.make_mask(wl.width, wl.height, bdm, seed=7)
    with Q.IvqmContext(wl.width, wl.height, bd, erp=1, search_range=R) as ctx:
        ctx.upload_pair(p0n, p1n, c420)
        ctx.upload_mask(m, bdm)
        r = ctx.frame(M4)
    g = oracle.gcd_masked(p0, p1, m, bd, bdm)
    assert list(r.gcd_r2t) == g
    mse, psnr, ws = oracle.psnr_family_masked(p0, p1, m, bd, bdm, erp=True)
    assert list(r.mse[:3]) == mse and list(r.psnr[:3]) == psnr and list(r.wspsnr[:3]) == ws
    iv = oracle.ivpsnr_masked(p0, p1, m, bd, bdm, g, R=R, erp=True)
    assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == iv
    assert list(r.ssd[:3]) == [oracle.weighted_ssd(p0[c], p1[c], m) for c in range(3)]


@pytest.mark.parametrize("bd,R,c420", [(10, 1, True), (10, 2, True), (8, 3, False), (8, 2, True), (10, 2, False),
                                       (8, 3, True), (10, 3, True), (10, 4, True)])
def test_masked_search_keyed_form_equals_reference_form(oracle, bd, R, c420):
    """The keyed masked search (search_mode_cap 0: the 4:2:0 form for 4:2:0 pairs, else the packed dot
    key, where the ranges allow; narrow key, cap 2; masked candidates keyed above every considered one)
    and the reference-form masked search (cap 1) give the same row sums, and they match the oracle."""
    wl = synth.Workload(cfg=170 + R, width=328, height=200, bit_depth=bd, chroma_420=c420, global_shift=1)
    p0n, p1n = synth.make_pair(wl, 0, as_u8=bd == 8)
    m = synth.make_mask(wl.width, wl.height, 10, seed=11 + R)
    out = []
    for cap in (0, 1, 2):
        with Q.IvqmContext(wl.width, wl.height, bd, search_range=R, search_mode_cap=cap) as ctx:
            ctx.upload_pair(p0n, p1n, c420)
            ctx.upload_mask(m, 10)
            r = ctx.frame(Q.IVQM_M_IVPSNR)
            out.append((r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r, list(r.gcd_r2t)))
    print("keyed", out[0], "reference form", out[1], "narrow key", out[2], flush=True)
    assert out[0] == out[1] == out[2]
    p0 = oracle.to444([np.asarray(p, np.uint16) for p in p0n], c420)
    p1 = oracle.to444([np.asarray(p, np.uint16) for p in p1n], c420)
    g = oracle.gcd_masked(p0, p1, m, bd, 10)
    assert out[0][3] == g
    assert out[0][:3] == oracle.ivpsnr_masked(p0, p1, m, bd, 10, g, R=R)


def _sha1(a):
    import hashlib
    return hashlib.sha1(np.ascontiguousarray(a, np.uint16).tobytes()).hexdigest()


@pytest.mark.parametrize("case", _golden_mask_cases(), ids=lambda c: c["name"])
def test_masked_scp_matches_reference_golden(case):
    """ivqm_gen_scp with a mask bound (xShftCompPicProc::GenShftCompPicsM, xShftCompPic.h:48-50) and
    the masked GCD: both pictures hash to the reference's, whose masked pixels keep 0."""
    wl = _wl(case)
    bd, bdm = wl.bit_depth, case["bit_depth_m"]
    print("case", case["name"], flush=True)
    p0n, p1n = synth.make_pair(wl, 0)
    m = synth.make_mask(wl.width, wl.height, bdm, seed=wl.cfg)
    with Q.IvqmContext(wl.width, wl.height, bd, search_range=case["R"]) as ctx:
        ctx.upload_pair(p0n, p1n, wl.chroma_420)
        ctx.upload_mask(m if bdm > 8 else m.astype(np.uint8), bdm)
        ctx.gen_scp(case["gcd"])
        s0, s1 = ctx.scp(0), ctx.scp(1)
    assert [_sha1(p) for p in s0] == case["scp0_sha1"]
    assert [_sha1(p) for p in s1] == case["scp1_sha1"]


@pytest.mark.parametrize("bd,R,c420", [(10, 2, True), (8, 3, False), (10, 4, True), (8, 3, True), (10, 1, True)])
def test_masked_scp_every_search_form_against_oracle(oracle, bd, R, c420):
    """Masked SCP generation in each masked search form (4:2:0 or packed dot key, reference form, narrow key)
    at 328x200: identical pictures, equal to the oracle's GenShftCompPicsM with 0 at masked pixels;
    a masked IV-PSNR call on the same context afterwards is unchanged."""
    wl = synth.Workload(cfg=190 + R, width=328, height=200, bit_depth=bd, chroma_420=c420, global_shift=1)
    p0n, p1n = synth.make_pair(wl, 0, as_u8=bd == 8)
    m = synth.make_mask(wl.width, wl.height, 10, seed=21 + R)
    p0 = oracle.to444([np.asarray(p, np.uint16) for p in p0n], c420)
    p1 = oracle.to444([np.asarray(p, np.uint16) for p in p1n], c420)
    g = oracle.gcd_masked(p0, p1, m, bd, 10)
    w0, w1 = oracle.scp_masked(p0, p1, m, bd, g, R=R, fill=0)
    iv = oracle.ivpsnr_masked(p0, p1, m, bd, 10, g, R=R)
    for cap in (0, 1, 2):
        print("cap", cap, flush=True)
        with Q.IvqmContext(wl.width, wl.height, bd, search_range=R, search_mode_cap=cap) as ctx:
            ctx.upload_pair(p0n, p1n, c420)
            ctx.upload_mask(m, 10)
            ctx.gen_scp(g)
            s0, s1 = ctx.scp(0), ctx.scp(1)
            assert all(np.array_equal(x, y) for x, y in zip(s0 + s1, w0 + w1)), cap
            assert ctx.ivpsnr(g) == iv


def test_masked_and_unmasked_pairs_in_one_batch():
    """A batch of masked (16-bit and 8-bit masks) and unmasked pairs of different sizes gives each
    pair the result of its own single-pair context."""
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    wls = [synth.Workload(cfg=170, width=256, height=128), synth.Workload(cfg=171, width=192, height=96),
           synth.Workload(cfg=172, width=128, height=128, bit_depth=8)]
    bdms = [10, None, 8]
    items, masks, want = [], [], []
    keep = []
    for wl, bdm in zip(wls, bdms):
        p0n, p1n = synth.make_pair(wl, 0)
        t0 = [torch.from_numpy(p.astype(np.uint16).view(np.int16)).to(dev) for p in p0n]
        t1 = [torch.from_numpy(p.astype(np.uint16).view(np.int16)).to(dev) for p in p1n]
        items.append((t0, t1, wl.width, wl.height, wl.bit_depth, wl.chroma_420))
        m = synth.make_mask(wl.width, wl.height, bdm, seed=wl.cfg) if bdm else None
        if m is not None:
            tm = torch.from_numpy(m.view(np.int16) if bdm > 8 else m.astype(np.uint8)).to(dev)
            masks.append((tm, bdm))
        else:
            masks.append((None, 0))
        keep += [t0, t1, masks[-1][0]]
        with Q.IvqmContext(wl.width, wl.height, wl.bit_depth, erp=1) as ctx:
            ctx.upload_pair([p.astype(np.uint16) for p in p0n], [p.astype(np.uint16) for p in p1n], wl.chroma_420)
            if m is not None:
                ctx.upload_mask(m, bdm)
            want.append(ctx.frame(M4))
    with Q.IvqmBatch(erp=1) as bt:
        arr = Q.IvqmBatch.make_pairs(items)
        got = bt.run_masked(arr, Q.IvqmBatch.make_masks(masks), M4)
    for g, w in zip(got, want):
        assert list(g.gcd_r2t) == list(w.gcd_r2t)
        for f in ("mse", "psnr", "wspsnr"):
            assert list(getattr(g, f)[:3]) == list(getattr(w, f)[:3]), f
        assert (g.ivpsnr, g.ivpsnr_r2t, g.ivpsnr_t2r) == (w.ivpsnr, w.ivpsnr_r2t, w.ivpsnr_t2r)


def test_mask_rejections():
    wl = synth.Workload(cfg=180, width=64, height=64)
    p0n, p1n = synth.make_pair(wl, 0)
    with Q.IvqmContext(64, 64, 10) as ctx:
        ctx.upload_pair(p0n, p1n, True)
        ctx.upload_mask(synth.make_mask(64, 64, 10, seed=1), 10)
        for m in (Q.IVQM_M_SSIM, Q.IVQM_M_IVSSIM, Q.IVQM_M_MSSSIM):  # xAppQMIV.cpp:360
            with pytest.raises(Q.IvqmError):
                ctx.frame(Q.IVQM_M_PSNR | m)
        ctx.upload_mask(np.zeros((64, 64), np.uint16), 10)  # no weighted sample
        with pytest.raises(Q.IvqmError):
            ctx.frame(Q.IVQM_M_PSNR)
        ctx.bind_mask(None)  # unbound: the unmasked metrics again
        r = ctx.frame(Q.IVQM_M_PSNR | Q.IVQM_M_SSIM)
        assert r.psnr[0] > 0 and r.ssim[0] > 0
