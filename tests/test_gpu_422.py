"""4:2:2 input on the GPU (the reference reads it with UpsampleH, libs/pmbb-core/src/xSeq.cpp:266-280;
the app's -cf 422, apps/QMIV/src/xAppQMIV.cpp:55): chroma planes of (width/2) x height, read at
(x >> 1, y). Every metric against the oracle on the 4:4:4 replication (IV-PSNR / PSNR / WS-PSNR / GCD
bit for bit, SSIM family within 1e-6), 8- and 16-bit samples, ERP, an odd height, mask mode, a batch
mixing the three chroma formats, the serving ABI and the standalone upsampling utility."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth

from oracle_checks import check_pair, merge_worst

pytestmark = pytest.mark.gpu
CF422 = 2


def _pair422(wl, frame, u8=False):
    """A 4:2:2 pair: the generator's 4:4:4 planes with every other chroma column."""
    p0, p1 = synth.make_pair(wl, frame, as_u8=u8)
    sub = lambda p: [np.ascontiguousarray(p[0])] + [np.ascontiguousarray(c[:, ::2]) for c in p[1:]]
    return sub(p0), sub(p1)


def _wl(cfg, w, h, bd=10, **kw):
    return synth.Workload(cfg=cfg, width=w, height=h, bit_depth=bd, chroma_420=False, global_shift=1, **kw)


@pytest.mark.parametrize("w,h,bd,erp", [(544, 528, 10, 0), (544, 528, 10, 1), (386, 257, 8, 1), (320, 200, 16, 0)])
def test_context_all_metrics_422(oracle, w, h, bd, erp):
    wl = _wl(610 + bd, w, h, bd)
    u8 = bd == 8
    p0, p1 = _pair422(wl, 0, u8)
    print("4:2:2", w, h, bd, "erp", erp, flush=True)
    with Q.IvqmContext(w, h, bd, erp=erp) as ctx:
        ctx.upload_pair(p0, p1, CF422)
        r = ctx.frame(Q.IVQM_M_ALL)
    errs = check_pair(oracle, r, p0, p1, bd, CF422, Q.IVQM_M_ALL, erp=bool(erp), label=f"{w}x{h} {bd}-bit")
    print("max |err| vs FP64 oracle:", errs, flush=True)


def test_422_search_forms_identical(oracle):
    """The keyed (packed dot), narrow-key and reference-form searches of a 4:2:2 pair agree, and
    their row sums and IV-PSNR equal the oracle's."""
    wl = _wl(620, 328, 200, 10)
    p0, p1 = _pair422(wl, 0)
    a, b = oracle.to444(p0, CF422), oracle.to444(p1, CF422)
    g = oracle.gcd(a, b, 10)
    v, r2t, t2r, ra, rb = oracle.ivpsnr(a, b, 10, g, want_rows=True)
    for cap in (0, 1, 2):
        with Q.IvqmContext(wl.width, wl.height, 10, search_mode_cap=cap) as ctx:
            ctx.upload_pair(p0, p1, CF422)
            r = ctx.frame(Q.IVQM_M_IVPSNR)
            assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == (v, r2t, t2r), cap
            assert np.array_equal(ctx.rows(0)[:, :3], ra[:, :3]), cap
            assert np.array_equal(ctx.rows(1)[:, :3], rb[:, :3]), cap


def test_422_mask_mode(oracle):
    wl = _wl(630, 256, 144, 10)
    p0, p1 = _pair422(wl, 0)
    m = synth.make_mask(wl.width, wl.height, 10, seed=5)
    with Q.IvqmContext(wl.width, wl.height, 10, erp=1) as ctx:
        ctx.upload_pair(p0, p1, CF422)
        ctx.upload_mask(m, 10)
        r = ctx.frame(Q.IVQM_M_MSE | Q.IVQM_M_PSNR | Q.IVQM_M_WSPSNR | Q.IVQM_M_IVPSNR)
        g = oracle.gcd_masked(oracle.to444(p0, CF422), oracle.to444(p1, CF422), m, 10, 10)
        ctx.gen_scp(g)
        s0, s1 = ctx.scp(0), ctx.scp(1)
    a, b = oracle.to444(p0, CF422), oracle.to444(p1, CF422)
    assert list(r.gcd_r2t) == g
    mse, psnr, ws = oracle.psnr_family_masked(a, b, m, 10, 10, erp=True)
    assert list(r.mse[:3]) == mse and list(r.psnr[:3]) == psnr and list(r.wspsnr[:3]) == ws
    assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == oracle.ivpsnr_masked(a, b, m, 10, 10, g, erp=True)
    w0, w1 = oracle.scp_masked(a, b, m, 10, g, fill=0)
    assert all(np.array_equal(x, y) for x, y in zip(s0 + s1, w0 + w1))


def test_batch_mixing_chroma_formats(oracle):
    """4:2:0, 4:2:2 and 4:4:4 pairs (8- and 10-bit) in one batch, all metrics, hipGraph capture and
    replay: each pair equals its own single-pair context and the oracle."""
    import torch
    dev = torch.device("cuda:0")
    specs = [(640, 0, 10, 1), (641, 1, 10, CF422), (642, 2, 8, CF422), (643, 3, 10, 0), (644, 4, 8, 1)]
    sizes = [(512, 256), (384, 200), (264, 136), (256, 128), (200, 104)]
    host, items, keep = [], [], []
    for (cfg, k, bd, cf), (w, h) in zip(specs, sizes):
        if cf == CF422:
            p0, p1 = _pair422(_wl(cfg, w, h, bd), 0, u8=bd == 8)
        else:
            wl = synth.Workload(cfg=cfg, width=w, height=h, bit_depth=bd, chroma_420=cf == 1, global_shift=1)
            p0, p1 = synth.make_pair(wl, 0, as_u8=bd == 8)
        def dv(planes):
            if planes[0].dtype == np.uint8:
                return [torch.from_numpy(np.ascontiguousarray(p)).to(dev) for p in planes]
            return [torch.from_numpy(np.ascontiguousarray(p).view(np.int16)).to(dev) for p in planes]
        t0, t1 = dv(p0), dv(p1)
        keep += [t0, t1]
        host.append((p0, p1, w, h, bd, cf))
        items.append((t0, t1, w, h, bd, cf))
    torch.cuda.synchronize()
    m = Q.IVQM_M_ALL
    with Q.IvqmBatch() as b:
        singles = [b.run(b.make_pairs([it]), m)[0].as_dict() for it in items]
        pairs = b.make_pairs(items)
        b.set_tuning(0, 0, True)
        first = [r.as_dict() for r in b.run(pairs, m)]
        second = [r.as_dict() for r in b.run(pairs, m)]
        # IV-PSNR + IV-SSIM: the batch takes the fused form (sub-block moments instead of SCP pictures)
        ivs = Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM
        fused = [r.as_dict() for r in b.run(pairs, ivs)]
        fused_singles = [b.run(b.make_pairs([it]), ivs)[0].as_dict() for it in items]
    assert first == second == singles
    assert fused == fused_singles
    worst = {}
    for (p0, p1, w, h, bd, cf), r, rf in zip(host, first, fused):
        label = f"{w}x{h} {bd}-bit chroma format {cf}"
        print(label, flush=True)
        merge_worst(worst, check_pair(oracle, r, p0, p1, bd, cf, m, label=label))
        merge_worst(worst, check_pair(oracle, rf, p0, p1, bd, cf, ivs, label=label + " fused"))
    print("max |err| vs FP64 oracle:", worst, flush=True)


def test_server_422_sequence(oracle):
    wl = _wl(650, 384, 216, 10)
    frames = [_pair422(wl, f) for f in range(2)]
    with Q.IvqmServer(memory_budget=1 << 30, max_batch_pairs=4, max_inflight=2) as srv:
        sid = srv.add_sequence([p1 for p0, p1 in frames], wl.width, wl.height, 10, CF422)
        srv.submit(1, sid, 0, [p0 for p0, p1 in frames], Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM | Q.IVQM_M_PSNR)
        srv.drain()
        res = srv.poll(8)
    assert len(res) == 1 and res[0].status == 0 and res[0].num_frames == 2
    for f, (p0, p1) in enumerate(frames):
        check_pair(oracle, res[0].frames[f], p0, p1, 10, CF422,
                   Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM | Q.IVQM_M_PSNR, label=f"served frame {f}")


def test_upsample_and_downsample_utilities_422(oracle):
    """ivqm_upsample420 and ivqm_downsample take 4:2:2 pictures (odd height): the 4:4:4 replication
    and its 2x2 mean."""
    import ctypes as C
    import torch
    from qmiv_amd._lib import Picture, check, load
    L = load()
    wl = _wl(660, 130, 67, 10)
    p0, _ = _pair422(wl, 0)
    t = [torch.from_numpy(np.ascontiguousarray(p).view(np.int16)).cuda() for p in p0]
    pic = Picture()
    for k in range(3):
        pic.plane[k] = t[k].data_ptr()
        pic.stride[k] = t[k].shape[1]
    pic.chroma_420, pic.bytes_per_sample = CF422, 2
    W, H = wl.width, wl.height
    up = torch.zeros((3, H, W), dtype=torch.int16, device="cuda")
    check(L.ivqm_upsample420(C.byref(pic), W, H, C.c_void_p(up.data_ptr()), W, None))
    dn = torch.zeros((3, H // 2, W // 2), dtype=torch.int16, device="cuda")
    check(L.ivqm_downsample(C.byref(pic), W, H, C.c_void_p(dn.data_ptr()), None))
    torch.cuda.synchronize()
    up_h, dn_h = up.cpu().numpy().view(np.uint16), dn.cpu().numpy().view(np.uint16)
    want = oracle.to444(p0, CF422)
    for c in range(3):
        assert np.array_equal(up_h[c], want[c])
        assert np.array_equal(dn_h[c], oracle.downsample(want[c][: H // 2 * 2]))
