"""Parity at the sizes of the measured workloads, through the C ABI.

* The mixed batch: an 8-bit and a 10-bit pair of each size (1280x720,
  1920x1080, 2048x2048, 4096x2048), all metrics, hipGraph capture and replay,
  autotuned, every pair against the oracle; and the whole 64-pair batch equal,
  pair by pair, to one pair per launch.
* The fused IV-SSIM batch (two or more pairs: the search writes sub-block
  moments instead of SCP pictures) at 4096x2048 ERP against the oracle.
* Rows wider than one pass of the row-statistics kernel (widths 4104 and 8200,
  10-bit 4:2:0) against the oracle.
* The GCD of a pair whose component sums pass 2^31 (8192x4096, 14-bit, both
  chroma formats), with the exact sum (default) and with the int32 sum of the
  x86-64-v4 reference build (gcd_sd_wrap32, xDistortionAVX512.cpp:14-73).
* The served stream at the mixed-batch sizes: IV-PSNR and IV-PSNR + IV-SSIM
  jobs, each job identical alone and in the mix, sampled frames against the
  oracle.

IV-PSNR / PSNR / WS-PSNR and GCD are compared bit for bit; the SSIM family
within 1e-6 of the FP64 oracle (tests/oracle_checks.py), and the largest
error per metric is printed.
"""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth
from qmiv_amd.context import IvqmBatch

from oracle_checks import check_pair, merge_worst

pytestmark = pytest.mark.gpu

IV = Q.IVQM_M_IVPSNR
IVS = Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM


def _dev(torch, planes):
    if planes[0].dtype == np.uint8:
        return [torch.from_numpy(p).cuda() for p in planes]
    return [torch.from_numpy(p.view(np.int16)).cuda() for p in planes]


def _mixed_pairs(torch, indices):
    host, items = [], []
    for i in indices:
        wl = synth.batch_workload(i)
        u8 = wl.bit_depth == 8
        p0, p1 = synth.make_pair(wl, i, as_u8=u8)
        host.append((i, wl, p0, p1))
        items.append((_dev(torch, p0), _dev(torch, p1), wl.width, wl.height, wl.bit_depth, True))
    torch.cuda.synchronize()
    return host, items


def test_mixed_batch_sizes_match_oracle(oracle):
    """An 8-bit and a 10-bit pair of each mixed-batch size in one batch: all metrics, autotuned,
    hipGraph captured and replayed; every pair against the oracle."""
    import torch
    idx = [0, 1, 16, 17, 32, 33, 48, 49]
    host, items = _mixed_pairs(torch, idx)
    m = Q.IVQM_M_ALL
    with IvqmBatch() as b:
        pairs = b.make_pairs(items)
        py, nb = b.autotune(pairs, m)
        b.set_tuning(py, nb, True)
        first = [r.as_dict() for r in b.run(pairs, m)]  # capture
        got = [r.as_dict() for r in b.run(pairs, m)]    # replay
    assert got == first
    worst = {}
    for (i, wl, p0, p1), r in zip(host, got):
        label = f"pair {i}: {wl.width}x{wl.height} {wl.bit_depth}-bit"
        print(label, flush=True)
        merge_worst(worst, check_pair(oracle, r, p0, p1, wl.bit_depth, True, m, label=label))
    print("tuning", (py, nb), "SSIM-family max |err| vs FP64 oracle:", worst, flush=True)


def test_mixed_batch_64_pairs_equals_one_pair_per_launch():
    """The whole 64-pair mixed batch (hipGraph, autotuned) equals, pair by pair, the batch of one pair."""
    import torch
    host, items = _mixed_pairs(torch, range(64))
    m = Q.IVQM_M_ALL
    with IvqmBatch() as b:
        singles = [b.run(b.make_pairs([it]), m)[0].as_dict() for it in items]
        pairs = b.make_pairs(items)
        py, nb = b.autotune(pairs, m)
        b.set_tuning(py, nb, True)
        for rep in range(2):
            got = [r.as_dict() for r in b.run(pairs, m)]
            for k, (g, s) in enumerate(zip(got, singles)):
                assert g == s, (rep, host[k][0])


def test_fused_ivssim_batch_erp_4096x2048_matches_oracle(oracle):
    """Three 4096x2048 ERP 10-bit pairs with IV-PSNR + IV-SSIM: the batch takes the fused form (the
    search writes sub-block moments, no SCP pictures); every pair against the oracle with ERP weights,
    and equal to the single-pair (SCP picture) form."""
    import torch
    wl = synth.ERP_4K2K
    host = [synth.make_pair(wl, f) for f in range(3)]
    items = [(_dev(torch, p0), _dev(torch, p1), wl.width, wl.height, wl.bit_depth, True) for p0, p1 in host]
    torch.cuda.synchronize()
    with IvqmBatch(erp=1) as b:
        got = [r.as_dict() for r in b.run(b.make_pairs(items), IVS)]
        singles = [b.run(b.make_pairs([it]), IVS)[0].as_dict() for it in items]
    assert got == singles
    worst = {}
    for f, ((p0, p1), r) in enumerate(zip(host, got)):
        merge_worst(worst, check_pair(oracle, r, p0, p1, wl.bit_depth, True, IVS, erp=True, label=f"frame {f}"))
    print("fused IV-SSIM 4096x2048 ERP: max |err| vs FP64 oracle:", worst, flush=True)


def test_rows_wider_than_one_stats_pass_match_oracle(oracle):
    """10-bit 4:2:0 pairs whose luma rows take more than one pass of the row-statistics kernel's
    narrow form (4096 samples of both pictures per wave and pass): widths 4104 (a second pass of one
    vector) and 8200 (three passes; its chroma rows of 4100 samples are not 16-byte aligned on odd
    rows, so they take the scalar path), in one batch; GCD, PSNR, WS-PSNR and IV-PSNR bit for bit."""
    import torch
    m = Q.IVQM_M_PSNR | Q.IVQM_M_WSPSNR | Q.IVQM_M_IVPSNR
    host, items = [], []
    for k, w in enumerate((4104, 8200)):
        wl = synth.Workload(cfg=760 + k, width=w, height=24, bit_depth=10)
        p0, p1 = synth.make_pair(wl, k)
        host.append((wl, p0, p1))
        items.append((_dev(torch, p0), _dev(torch, p1), w, wl.height, 10, True))
    torch.cuda.synchronize()
    with IvqmBatch() as b:
        got = [r.as_dict() for r in b.run(b.make_pairs(items), m)]
    for (wl, p0, p1), r in zip(host, got):
        label = f"{wl.width}x{wl.height} 10-bit"
        print(label, flush=True)
        check_pair(oracle, r, p0, p1, 10, True, m, label=label)


@pytest.mark.parametrize("chroma_420", [True, False])
def test_gcd_sum_past_2_31_exact_and_wrap32(oracle, chroma_420):
    """|sum(P1 - P0)| >= 2^31 in every component (8192x4096, 14-bit, offsets +100 / -90 / +70):
    the default exact sum gives the offsets, gcd_sd_wrap32 the x86-64-v4 build's int32 sums."""
    w, h, bd = 8192, 4096, 14
    rng = np.random.default_rng(41)
    cw, ch = (w // 2, h // 2) if chroma_420 else (w, h)
    p0 = [rng.integers(2000, 14000, (h, w), dtype=np.uint16)] + \
         [rng.integers(2000, 14000, (ch, cw), dtype=np.uint16) for _ in range(2)]
    off = (100, -90, 70)
    p1 = [(p0[c].astype(np.int32) + off[c]).astype(np.uint16) for c in range(3)]
    a = [x.astype(np.uint16) for x in oracle.to444(p0, chroma_420)]
    b = [x.astype(np.uint16) for x in oracle.to444(p1, chroma_420)]
    want_exact = oracle.gcd(a, b, bd, wrap32=False)
    want_wrap = oracle.gcd(a, b, bd, wrap32=True)
    assert want_exact[:3] == list(off)
    assert all(want_exact[c] != want_wrap[c] for c in range(3))  # every component wraps
    for wrap, want in ((0, want_exact), (1, want_wrap)):
        with Q.IvqmContext(w, h, bd, gcd_sd_wrap32=wrap) as ctx:
            ctx.upload_pair(p0, p1, chroma_420)
            assert ctx.gcd() == want, (wrap, ctx.gcd(), want)
            r = ctx.frame(Q.IVQM_M_PSNR)
            assert list(r.gcd_r2t) == want
    print("GCD exact", want_exact, "int32 sums", want_wrap, flush=True)


def test_served_stream_at_mixed_batch_sizes(oracle):
    """Sequences of every mixed-batch size and depth resident on the server; jobs of 1-3 frames asking
    IV-PSNR or IV-PSNR + IV-SSIM arrive together. Each job's results equal the same job run alone on
    an idle server, and sampled frames match the oracle."""
    nfr = 3
    seqs = []
    for k, (w, h) in enumerate(synth.BATCH_SIZES):
        for bd in (8, 10):
            wl = synth.Workload(cfg=730 + 2 * k + (bd == 10), width=w, height=h, bit_depth=bd)
            ref, tst = [], []
            for f in range(nfr):
                p0, p1 = synth.make_pair(wl, f, as_u8=bd == 8)
                ref.append(p1)
                tst.append(p0)
            seqs.append((wl, ref, tst))
    rng = np.random.default_rng(9)
    jobs = []
    for tag in range(24):
        si = int(rng.integers(0, len(seqs)))
        n = int(rng.integers(1, nfr + 1))
        first = int(rng.integers(0, nfr - n + 1))
        jobs.append((tag, si, first, n, IV if tag % 2 else IVS))

    def run(srv, sids, batch):
        for tag, si, first, n, m in batch:
            srv.submit(tag, sids[si], first, seqs[si][2][first:first + n], m)
        srv.drain()
        out = {}
        while True:
            r = srv.poll(64)
            if not r:
                break
            out.update({x.tag: x for x in r})
        return out

    fields = ["gcd_r2t", "ivpsnr", "ivpsnr_r2t", "ivpsnr_t2r", "ivssim", "ivssim_r2t", "ivssim_t2r"]
    with Q.IvqmServer(memory_budget=8 << 30, max_batch_pairs=16, max_inflight=3) as srv:
        sids = [srv.add_sequence(ref, wl.width, wl.height, wl.bit_depth) for wl, ref, tst in seqs]
        alone = {}
        for j in jobs:
            alone.update(run(srv, sids, [j]))
        mixed = run(srv, sids, jobs)
        st = srv.stats()
    assert st["max_batch_fill"] > 1
    for tag, si, first, n, m in jobs:
        a, b = alone[tag], mixed[tag]
        assert a.status == 0 and b.status == 0 and a.num_frames == b.num_frames == n
        for f in range(n):
            fa, fb = a.frames[f].as_dict(), b.frames[f].as_dict()
            for k in fields:
                assert fa[k] == fb[k], (tag, f, k)
    worst = {}
    for tag, si, first, n, m in jobs[:8]:  # one job per sequence size and depth, roughly
        wl, ref, tst = seqs[si]
        r = mixed[tag]
        f = n - 1
        label = f"job {tag}: {wl.width}x{wl.height} {wl.bit_depth}-bit frame {first + f}"
        merge_worst(worst, check_pair(oracle, r.frames[f], tst[first + f], ref[first + f], wl.bit_depth, True, m,
                                      label=label))
    print("served stream: max |err| vs FP64 oracle:", worst, flush=True)
