"""GPU parity: the HIP path (through the C ABI) against the oracle on the same seeded inputs.

Bar: GCD, per-row search distances, per-pixel winners, SCP pictures, SSDs bit-exact;
PSNR / WS-PSNR / IV-PSNR identical FP64 values (repr equality); SSIM family within
1e-6 absolute of the FP64 oracle (the error is printed for every metric).
"""
import numpy as np
import pytest

import qmiv_amd as Q
from conftest import golden_cases
from qmiv_amd import synth

pytestmark = pytest.mark.gpu

SSIM_TOL = 1e-6
CASES = golden_cases()


def wl_of(case):
    w = dict(case["workload"])
    w["offsets"] = tuple(w["offsets"])
    return synth.Workload(**w)


def _ctx(wl, R=2, erp=False, **kw):
    return Q.IvqmContext(wl.width, wl.height, wl.bit_depth, search_range=R, erp=int(erp), **kw)


def _check_frame(ctx, oracle, wl, p0n, p1n, R, erp, label):
    bd = wl.bit_depth
    p0 = oracle.to444(p0n, wl.chroma_420)
    p1 = oracle.to444(p1n, wl.chroma_420)
    ctx.enable_best_map(True)
    r = ctx.frame(Q.IVQM_M_ALL)
    g = oracle.gcd(p0, p1, bd, wrap32=False)
    assert list(r.gcd_r2t) == g, label
    # rows of both directions, exact
    rows_a, scp1, best_a = oracle.cps(p0, p1, bd, g, R=R, want_scp=True, want_best=True)
    rows_b, scp0, best_b = oracle.cps(p1, p0, bd, [-x for x in g], R=R, want_scp=True, want_best=True)
    assert np.array_equal(ctx.rows(0)[:, :3], rows_a[:, :3]), label
    assert np.array_equal(ctx.rows(1)[:, :3], rows_b[:, :3]), label
    assert np.array_equal(ctx.best_map(0), best_a), label
    assert np.array_equal(ctx.best_map(1), best_b), label
    for c in range(3):
        assert np.array_equal(ctx.scp(0)[c], scp0[c]), (label, "scp0", c)
        assert np.array_equal(ctx.scp(1)[c], scp1[c]), (label, "scp1", c)
    # integer SSDs and the FP64 dB values, bit-exact
    assert list(r.ssd[:3]) == [oracle.ssd(p0[c], p1[c]) for c in range(3)], label
    assert [repr(x) for x in r.psnr[:3]] == [repr(x) for x in oracle.psnr(p0, p1, bd)], label
    assert [repr(x) for x in r.wspsnr[:3]] == [repr(x) for x in oracle.wspsnr(p0, p1, bd, erp=erp)], label
    v, a, b = oracle.ivpsnr(p0, p1, bd, g, R=R, erp=erp)
    assert (repr(r.ivpsnr), repr(r.ivpsnr_r2t), repr(r.ivpsnr_t2r)) == (repr(v), repr(a), repr(b)), label
    # SSIM family vs the FP64 oracle
    errs = {}
    errs["ssim"] = np.max(np.abs(np.array(r.ssim[:3]) - oracle.ssim(p0, p1, bd)))
    errs["msssim"] = np.max(np.abs(np.array(r.msssim[:3]) - oracle.msssim(p0, p1, bd)))
    iv, ir, it = oracle.ivssim(p0, p1, scp0, scp1, bd, erp=erp)
    errs["ivssim"] = max(abs(r.ivssim - iv), abs(r.ivssim_r2t - ir), abs(r.ivssim_t2r - it))
    errs["ivmsssim"] = abs(r.ivmsssim - oracle.ivmsssim(p0, p1, scp0, scp1, bd))
    print(label, "SSIM-family |err| vs FP64 oracle:", {k: f"{v:.2e}" for k, v in errs.items()}, flush=True)
    for k, e in errs.items():
        assert e <= SSIM_TOL, (label, k, e)
    return r


@pytest.mark.parametrize("case", CASES, ids=[c["name"] for c in CASES])
def test_gpu_matches_oracle_on_golden_inputs(case, oracle):
    wl = wl_of(case)
    p0n, p1n = synth.make_pair(wl, case["frame"])
    with _ctx(wl, R=case["R"], erp=case["erp"]) as ctx:
        ctx.upload_pair(p0n, p1n, wl.chroma_420)
        r = _check_frame(ctx, oracle, wl, p0n, p1n, case["R"], case["erp"], case["name"])
        # and the golden vectors of the reference library themselves
        assert list(r.gcd_r2t) == case["gcd_r2t"]
        assert [repr(x) for x in r.psnr[:3]] == case["psnr"]
        want = case["ivpsnr_erp"] if case["erp"] else case["ivpsnr"]
        assert [repr(r.ivpsnr), repr(r.ivpsnr_r2t), repr(r.ivpsnr_t2r)] == want
        if "ssim" in case:
            assert np.max(np.abs(np.array(r.ssim[:3]) - [float(x) for x in case["ssim"]])) <= SSIM_TOL


def test_gpu_8bit_samples(oracle):
    wl = synth.Workload(cfg=200, width=160, height=96, bit_depth=8)
    p0n, p1n = synth.make_pair(wl, 0, as_u8=True)
    with _ctx(wl) as ctx:
        ctx.upload_pair(p0n, p1n, True)
        _check_frame(ctx, oracle, wl, [p.astype(np.uint16) for p in p0n], [p.astype(np.uint16) for p in p1n], 2,
                     False, "u8")


@pytest.mark.parametrize("w,h", [(66, 34), (130, 72), (200, 40)])
def test_gpu_ragged_sizes(oracle, w, h):
    """Sizes that are not multiples of the tile (64 x 32) or of the SSIM window stride."""
    wl = synth.Workload(cfg=201, width=w, height=h, block_shift=2)
    p0n, p1n = synth.make_pair(wl, 3)
    with _ctx(wl, erp=True) as ctx:
        ctx.upload_pair(p0n, p1n, True)
        _check_frame(ctx, oracle, wl, p0n, p1n, 2, True, f"{w}x{h}")


def test_gpu_wide_path_large_weights(oracle):
    """Weights that overflow the packed-key range force the exact (wide) search."""
    wl = synth.Workload(cfg=202, width=96, height=64)
    p0n, p1n = synth.make_pair(wl, 0)
    p0 = oracle.to444(p0n, True)
    p1 = oracle.to444(p1n, True)
    cw = (1000, 3, 2, 0)
    with Q.IvqmContext(wl.width, wl.height, 10, cmp_weights_search=cw) as ctx:
        ctx.upload_pair(p0n, p1n, True)
        v, a, b = ctx.ivpsnr()
        g = ctx.gcd()
        ov = oracle.ivpsnr(p0, p1, 10, g, R=2, cws=cw)
        assert (repr(v), repr(a), repr(b)) == tuple(repr(x) for x in ov)


@pytest.mark.parametrize("bd,R,u8", [(10, 2, False), (8, 2, True), (8, 4, True), (9, 3, False), (10, 1, False), (10, 3, False),
                                      (12, 2, False)])
def test_gpu_search_variants_identical(oracle, bd, R, u8):
    """The packed-int16 dot, packed-key and reference-form searches give the same
    rows, SCP pictures, winner maps and dB values; the default one also matches
    the oracle."""
    wl = synth.Workload(cfg=207, width=200, height=72, bit_depth=bd, global_shift=1, block_shift=R)
    p0n, p1n = synth.make_pair(wl, 0, as_u8=u8)
    out = []
    for cap in (0, 1, 2):
        with _ctx(wl, R=R, search_mode_cap=cap) as ctx:
            ctx.upload_pair(p0n, p1n, True)
            ctx.enable_best_map(True)
            r = ctx.frame(Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM)
            out.append((r.as_dict(), ctx.rows(0), ctx.rows(1), ctx.best_map(0), ctx.best_map(1),
                        [ctx.scp(0)[c].copy() for c in range(3)], [ctx.scp(1)[c].copy() for c in range(3)]))
    for o in out[1:]:
        assert o[0] == out[0][0]
        for a, b in zip(o[1:5], out[0][1:5]):
            assert np.array_equal(a, b)
        for c in range(3):
            assert np.array_equal(o[5][c], out[0][5][c]) and np.array_equal(o[6][c], out[0][6][c])
    p0 = oracle.to444([p.astype(np.uint16) for p in p0n], True)
    p1 = oracle.to444([p.astype(np.uint16) for p in p1n], True)
    g = out[0][0]["gcd_r2t"]
    ov = oracle.ivpsnr(p0, p1, bd, g, R=R)
    assert (repr(out[0][0]["ivpsnr"]), repr(out[0][0]["ivpsnr_r2t"]), repr(out[0][0]["ivpsnr_t2r"])) == \
        tuple(repr(x) for x in ov)


def test_gpu_per_metric_calls_match_fused_frame(oracle):
    wl = synth.Workload(cfg=203, width=256, height=128, global_shift=1)
    p0n, p1n = synth.make_pair(wl, 1)
    with _ctx(wl, erp=True) as ctx:
        ctx.upload_pair(p0n, p1n, True)
        r = ctx.frame(Q.IVQM_M_ALL)
        assert ctx.gcd() == list(r.gcd_r2t)
        assert ctx.psnr()[:3] == list(r.psnr[:3])
        assert ctx.wspsnr()[:3] == list(r.wspsnr[:3])
        assert ctx.ssim()[:3] == list(r.ssim[:3])
        assert ctx.msssim()[:3] == list(r.msssim[:3])
        assert ctx.ivpsnr() == (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r)
        # explicit GCD path (xIVPSNR.h:36 with the caller's GCD)
        assert ctx.ivpsnr(list(r.gcd_r2t)) == (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r)
        ctx.gen_scp()
        assert ctx.ivssim() == (r.ivssim, r.ivssim_r2t, r.ivssim_t2r)
        assert ctx.ivmsssim() == (r.ivmsssim, r.ivmsssim_r2t, r.ivmsssim_t2r)


def test_gpu_ring_async_equals_sync():
    wl = synth.Workload(cfg=204, width=192, height=128)
    pairs = [synth.make_pair(wl, f) for f in range(4)]
    with _ctx(wl) as ctx:
        sync = []
        for p0n, p1n in pairs:
            ctx.upload_pair(p0n, p1n, True)
            sync.append(ctx.frame(Q.IVQM_M_ALL).as_dict())
        for f, (p0n, p1n) in enumerate(pairs):
            ctx.upload_pair(p0n, p1n, True)
            ctx.enqueue(Q.IVQM_M_ALL, f)
            # uploads wait for the stream, so each enqueued pass sees its own pair
        for f in range(len(pairs)):
            assert ctx.collect(f).as_dict() == sync[f]


def test_gpu_exact_components_fake_psnr(oracle):
    """An identical component: SSD 0 -> flt64 max, or PSNR(SSD=1) with fake values (xPSNR.cpp:146-147)."""
    wl = synth.Workload(cfg=205, width=64, height=64)
    p0n, p1n = synth.make_pair(wl, 0)
    p1n = [p0n[0].copy(), p1n[1], p1n[2]]
    for fake in (0, 1):
        with Q.IvqmContext(64, 64, 10, fake_vals_for_exact=fake) as ctx:
            ctx.upload_pair(p0n, p1n, True)
            r = ctx.frame(Q.IVQM_M_PSNR | Q.IVQM_M_WSPSNR)
            assert r.exact[0] == 1 and r.ssd[0] == 0
            want = oracle.psnr_from_ssd(1.0, 64 * 64, 10) if fake else np.finfo(np.float64).max
            assert r.psnr[0] == want
            assert r.wspsnr[0] == want


def test_gpu_device_resident_bind_and_pixel_ops(oracle):
    import torch
    wl = synth.Workload(cfg=206, width=128, height=64)
    p0n, p1n = synth.make_pair(wl, 0)
    dev = torch.device("cuda:0")
    t0 = [torch.from_numpy(p.astype(np.int16).view(np.int16)).to(dev) for p in p0n]
    t1 = [torch.from_numpy(p.astype(np.int16).view(np.int16)).to(dev) for p in p1n]
    with _ctx(wl) as a, _ctx(wl) as b:
        a.bind(0, t0, True)
        a.bind(1, t1, True)
        b.upload_pair(p0n, p1n, True)
        torch.cuda.synchronize()
        assert a.frame(Q.IVQM_M_ALL).as_dict() == b.frame(Q.IVQM_M_ALL).as_dict()
    # standalone upsample / downsample kernels
    import ctypes as C
    from qmiv_amd._lib import Picture, check, load
    L = load()
    pic = Picture()
    for k in range(3):
        pic.plane[k] = t0[k].data_ptr()
        pic.stride[k] = t0[k].shape[1]
    pic.chroma_420, pic.bytes_per_sample = 1, 2
    up = torch.zeros((3, 64, 128), dtype=torch.int16, device=dev)
    check(L.ivqm_upsample420(C.byref(pic), 128, 64, C.c_void_p(up.data_ptr()), 128, None))
    dn = torch.zeros((3, 32, 64), dtype=torch.int16, device=dev)
    check(L.ivqm_downsample(C.byref(pic), 128, 64, C.c_void_p(dn.data_ptr()), None))
    torch.cuda.synchronize()
    up_h = up.cpu().numpy().view(np.uint16)
    dn_h = dn.cpu().numpy().view(np.uint16)
    p444 = oracle.to444(p0n, True)
    for c in range(3):
        assert np.array_equal(up_h[c], p444[c])
        assert np.array_equal(dn_h[c], oracle.downsample(p444[c]))


def test_gpu_full_size_4096x2048_erp(oracle):
    """One full 4096x2048 ERP pair of the SSIM-family workload against the oracle."""
    wl = synth.ERP_4K2K
    p0n, p1n = synth.make_pair(wl, 0)
    with _ctx(wl, erp=True) as ctx:
        ctx.upload_pair(p0n, p1n, True)
        _check_frame(ctx, oracle, wl, p0n, p1n, 2, True, "4096x2048")


def test_gpu_errors_are_reported():
    with pytest.raises(Q.IvqmError):
        Q.IvqmContext(64, 64, 17)
    with pytest.raises(Q.IvqmError):
        Q.IvqmContext(64, 64, 10, search_range=9)
    with Q.IvqmContext(64, 64, 10) as ctx:
        with pytest.raises(Q.IvqmError):
            ctx.frame()  # no pictures bound
