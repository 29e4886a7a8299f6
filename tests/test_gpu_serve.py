"""The serving loop (include/ivqm_serve.h): every job's per-frame results equal
the per-pair context calls on the same pictures, alone and in a mix of
concurrent jobs of other sizes, bit depths and metric sets; admission respects
the memory budget; invalid requests are rejected."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth

pytestmark = pytest.mark.gpu

IV = Q.IVQM_M_IVPSNR
IVS = Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM
# (workload, 8-bit samples)
SEQS = [
    (synth.Workload(cfg=700, width=256, height=128, global_shift=1), False),
    (synth.Workload(cfg=701, width=192, height=96, bit_depth=8), True),
    (synth.Workload(cfg=702, width=320, height=176, global_shift=2), False),
]
NFR = 4  # frames per sequence


def _frames(wl, u8):
    """Reference frames (P1) and test frames (P0) of a sequence."""
    ref, tst = [], []
    for f in range(NFR):
        p0, p1 = synth.make_pair(wl, f, as_u8=u8)
        ref.append(p1)
        tst.append(p0)
    return ref, tst


@pytest.fixture(scope="module")
def data():
    return [(wl, u8, *_frames(wl, u8)) for wl, u8 in SEQS]


def _single(wl, p0, p1, metrics):
    with Q.IvqmContext(wl.width, wl.height, wl.bit_depth) as ctx:
        ctx.upload_pair(p0, p1, True)
        return ctx.frame(metrics).as_dict()


def _fields(metrics):
    f = ["gcd_r2t", "ivpsnr", "ivpsnr_r2t", "ivpsnr_t2r"]
    if metrics & Q.IVQM_M_IVSSIM:
        f += ["ivssim", "ivssim_r2t", "ivssim_t2r"]
    return f


def _run(srv, jobs, tests):
    """jobs: list of (tag, seq index, seq_id, first, n, metrics); returns {tag: JobResult}."""
    for tag, si, sid, first, n, m in jobs:
        srv.submit(tag, sid, first, tests[si][first:first + n], m)
    srv.drain()
    out = {}
    while True:
        r = srv.poll(64)
        if not r:
            break
        for x in r:
            out[x.tag] = x
    return out


def test_serve_matches_single_pair_calls_alone_and_mixed(data):
    with Q.IvqmServer(memory_budget=256 << 20, max_batch_pairs=6) as srv:
        sids = [srv.add_sequence(ref, wl.width, wl.height, wl.bit_depth) for wl, u8, ref, tst in data]
        tests = [d[3] for d in data]
        probe = (1, 0, sids[0], 1, 3, IVS)
        alone = _run(srv, [probe], tests)[1]
        assert alone.status == 0 and alone.num_frames == 3
        mix = [probe]
        tag = 100
        rng = np.random.default_rng(5)
        for _ in range(24):
            si = int(rng.integers(0, len(SEQS)))
            n = int(rng.integers(1, NFR + 1))
            first = int(rng.integers(0, NFR - n + 1))
            mix.append((tag, si, sids[si], first, n, IV if rng.random() < 0.5 else IVS))
            tag += 1
        rng.shuffle(mix)
        res = _run(srv, mix, tests)
        assert set(res) == {t[0] for t in mix}
        st = srv.stats()
        assert st["jobs_completed"] == st["jobs_submitted"] == 1 + len(mix)
        assert st["max_batch_fill"] > 1  # pairs of several jobs shared launches
        # the probe: identical alone and in the mix
        mixed = res[1]
        for f in range(3):
            a, b = alone.frames[f].as_dict(), mixed.frames[f].as_dict()
            for k in _fields(IVS):
                assert a[k] == b[k], (f, k)
        # every job equals the per-pair context calls on the same pictures
        for tag, si, sid, first, n, m in mix:
            r = res[tag]
            assert r.status == 0 and r.num_frames == n
            assert r.latency_us >= r.queue_us >= 0
            wl, u8, ref, tst = data[si]
            for f in range(n):
                want = _single(wl, tst[first + f], ref[first + f], m)
                got = r.frames[f].as_dict()
                for k in _fields(m):
                    assert repr(got[k]) == repr(want[k]), (tag, f, k)


def test_serve_matches_oracle(data, oracle):
    wl, u8, ref, tst = data[0]
    with Q.IvqmServer() as srv:
        sid = srv.add_sequence(ref, wl.width, wl.height, wl.bit_depth)
        srv.submit(7, sid, 0, tst[:2], IVS)
        srv.drain()
        (r,) = srv.poll(4)
    for f in range(2):
        p0 = oracle.to444(tst[f], True)
        p1 = oracle.to444(ref[f], True)
        g = oracle.gcd(p0, p1, wl.bit_depth)
        got = r.frames[f]
        assert list(got.gcd_r2t) == g
        assert (repr(got.ivpsnr), repr(got.ivpsnr_r2t), repr(got.ivpsnr_t2r)) == \
            tuple(repr(x) for x in oracle.ivpsnr(p0, p1, wl.bit_depth, g))
        s0, s1 = oracle.scp(p0, p1, wl.bit_depth, g)
        v, a, b = oracle.ivssim(p0, p1, s0, s1, wl.bit_depth)
        assert abs(got.ivssim - v) <= 1e-6 and abs(got.ivssim_r2t - a) <= 1e-6 and abs(got.ivssim_t2r - b) <= 1e-6


def test_serve_memory_budget_and_device_planes(data):
    import torch
    wl, u8, ref, tst = data[2]
    # staged bytes of one 320x176 10-bit 4:2:0 frame with 64-sample pitches
    frame_bytes = (320 * 176 + 2 * 192 * 88) * 2
    budget = 3 * frame_bytes + 4096
    with Q.IvqmServer(memory_budget=budget, max_batch_pairs=2, max_inflight=2) as srv:
        sid = srv.add_sequence(ref, wl.width, wl.height, wl.bit_depth)
        for t in range(10):
            srv.submit(t, sid, 0, tst[:2], IV)  # two frames per job: at most one job staged at a time
        srv.submit(50, sid, 0, tst[:4], IV)  # larger than the budget: fails, others unaffected
        dev = [[torch.from_numpy(p.view(np.int16)).cuda() for p in fr] for fr in tst]
        torch.cuda.synchronize()
        srv.submit(60, sid, 0, dev[:4], IV)  # device planes are read in place, outside the budget
        srv.drain()
        res = {}
        while True:
            r = srv.poll(64)
            if not r:
                break
            res.update({x.tag: x for x in r})
        st = srv.stats()
    assert st["budget_peak"] <= budget
    assert res[50].status == -12
    assert all(res[t].status == 0 for t in range(10)) and res[60].status == 0
    for f in range(2):
        assert res[0].frames[f].ivpsnr == res[9].frames[f].ivpsnr == res[60].frames[f].ivpsnr


def test_serve_rejects_invalid_requests(data):
    wl, u8, ref, tst = data[0]
    with Q.IvqmServer() as srv:
        sid = srv.add_sequence(ref, wl.width, wl.height, wl.bit_depth)
        with pytest.raises(Q.IvqmError):
            srv.submit(1, sid + 99, 0, tst[:1], IV)  # unknown sequence
        with pytest.raises(Q.IvqmError):
            srv.submit(2, sid, NFR - 1, tst[:2], IV)  # frame range past the end
        with pytest.raises(Q.IvqmError):
            srv.submit(3, sid, 0, tst[:1], 0)  # no metrics
        srv.submit(4, sid, 0, tst[:1], IV)
        try:
            srv.remove_sequence(sid)  # busy while job 4 is unfinished, or removed once it is done
            removed = True
        except Q.IvqmError as e:
            assert e.status == -16
            removed = False
        srv.drain()
        assert srv.poll(4)[0].status == 0
        if not removed:
            srv.remove_sequence(sid)


def test_serve_mask_mode_alone_mixed_and_oracle(data, oracle):
    """Jobs with per-frame weight masks (host-staged 16-bit masks and device-resident 8-bit masks)
    mixed with unmasked IV-PSNR / IV-SSIM jobs: each masked frame equals the masked context call on
    the same pictures and mask, and the oracle; unmasked jobs are unchanged by the masked ones;
    structural-similarity metrics with a mask are rejected."""
    import torch
    M4 = Q.IVQM_M_MSE | Q.IVQM_M_PSNR | Q.IVQM_M_WSPSNR | Q.IVQM_M_IVPSNR
    with Q.IvqmServer(memory_budget=256 << 20, max_batch_pairs=8) as srv:
        sids = [srv.add_sequence(ref, wl.width, wl.height, wl.bit_depth) for wl, u8, ref, tst in data]
        jobs = []  # (tag, si, first, n, metrics, masks, mask bit depth, masks on the device)
        for tag, si in enumerate([0, 1, 2, 0, 2]):
            wl = data[si][0]
            bdm = 10 if tag % 2 == 0 else 8
            masks = [synth.make_mask(wl.width, wl.height, bdm, seed=40 + 4 * tag + f) for f in range(2)]
            jobs.append((tag, si, 1, 2, M4, masks, bdm, tag == 3))
        unmasked = [(50 + k, si, 0, 2, IVS if k % 2 else IV, None, 0, False) for k, si in enumerate([0, 1, 2])]
        keep = []
        for tag, si, first, n, m, masks, bdm, on_dev in jobs + unmasked:
            tst = data[si][3][first:first + n]
            if masks is not None and on_dev:  # device planes: test frames and masks both on the device
                tst = [[torch.from_numpy(np.ascontiguousarray(p).astype(np.uint16).view(np.int16)).cuda()
                        if p.dtype != np.uint8 else torch.from_numpy(p).cuda() for p in fr] for fr in tst]
                masks = [torch.from_numpy(mk.astype(np.uint8)).cuda() if bdm == 8 else
                         torch.from_numpy(mk.view(np.int16)).cuda() for mk in masks]
                torch.cuda.synchronize()
                keep += [tst, masks]
            srv.submit(tag, sids[si], first, tst, m, masks=masks, mask_bit_depth=bdm)
        with pytest.raises(Q.IvqmError):  # SSIM family with a mask
            srv.submit(99, sids[0], 0, data[0][3][:1], IVS, masks=jobs[0][5][:1], mask_bit_depth=10)
        srv.drain()
        res = {}
        while True:
            r = srv.poll(64)
            if not r:
                break
            res.update({x.tag: x for x in r})
        st = srv.stats()
    assert st["max_batch_fill"] > 1
    for tag, si, first, n, m, masks, bdm, on_dev in jobs:
        wl, u8, ref, tst = data[si]
        r = res[tag]
        assert r.status == 0 and r.num_frames == n
        for f in range(n):
            p0n, p1n, mk = tst[first + f], ref[first + f], masks[f]
            with Q.IvqmContext(wl.width, wl.height, wl.bit_depth) as ctx:
                ctx.upload_pair(p0n, p1n, True)
                ctx.upload_mask(mk if bdm > 8 else mk.astype(np.uint8), bdm)
                want = ctx.frame(m).as_dict()
            got = r.frames[f].as_dict()
            for k in ("gcd_r2t", "mse", "psnr", "wspsnr", "ivpsnr", "ivpsnr_r2t", "ivpsnr_t2r"):
                assert repr(got[k]) == repr(want[k]), (tag, f, k)
            p0 = oracle.to444([np.asarray(p, np.uint16) for p in p0n], True)
            p1 = oracle.to444([np.asarray(p, np.uint16) for p in p1n], True)
            g = oracle.gcd_masked(p0, p1, mk, wl.bit_depth, bdm)
            assert list(got["gcd_r2t"]) == g
            mse, psnr, ws = oracle.psnr_family_masked(p0, p1, mk, wl.bit_depth, bdm)
            assert list(got["mse"][:3]) == mse and list(got["psnr"][:3]) == psnr
            assert (got["ivpsnr"], got["ivpsnr_r2t"], got["ivpsnr_t2r"]) == \
                oracle.ivpsnr_masked(p0, p1, mk, wl.bit_depth, bdm, g)
    for tag, si, first, n, m, masks, bdm, on_dev in unmasked:
        wl, u8, ref, tst = data[si]
        for f in range(n):
            want = _single(wl, tst[first + f], ref[first + f], m)
            got = res[tag].frames[f].as_dict()
            for k in _fields(m):
                assert repr(got[k]) == repr(want[k]), (tag, f, k)
