"""Batched form (include/ivqm_batch.h): mixed sizes, bit depths and chroma formats in
one launch per phase. Each pair's result must equal the single-pair result (the
batch composition must not change any value), with and without hipGraph replay and
after autotuning."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth
from qmiv_amd.context import IvqmBatch

pytestmark = pytest.mark.gpu

MIXED = [
    synth.Workload(cfg=500, width=128, height=64, bit_depth=8),
    synth.Workload(cfg=501, width=200, height=96),
    synth.Workload(cfg=502, width=72, height=64, chroma_420=False),
    synth.Workload(cfg=503, width=544, height=520, global_shift=1),
    synth.Workload(cfg=504, width=96, height=80, bit_depth=8),
    synth.Workload(cfg=505, width=136, height=72, bit_depth=12, sigma_r=20, sigma_t=15),
]


def _dev_pairs():
    import torch
    dev = torch.device("cuda:0")
    items, host = [], []
    for wl in MIXED:
        u8 = wl.bit_depth == 8
        p0, p1 = synth.make_pair(wl, 0, as_u8=u8)
        host.append((wl, p0, p1))
        cv = (lambda a: torch.from_numpy(a).to(dev)) if u8 else (lambda a: torch.from_numpy(a.view(np.int16)).to(dev))
        items.append(([cv(p) for p in p0], [cv(p) for p in p1], wl.width, wl.height, wl.bit_depth, wl.chroma_420))
    torch.cuda.synchronize()
    return items, host


def _single(wl, p0, p1, metrics):
    with Q.IvqmContext(wl.width, wl.height, wl.bit_depth) as ctx:
        ctx.upload_pair(p0, p1, wl.chroma_420)
        return ctx.frame(metrics).as_dict()


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_single_pair_results(use_graph):
    items, host = _dev_pairs()
    want = [_single(wl, p0, p1, Q.IVQM_M_ALL) for wl, p0, p1 in host]
    with IvqmBatch() as b:
        b.set_tuning(0, 0, use_graph)
        pairs = b.make_pairs(items)
        for rep in range(3):  # graph capture on the first pass, replays after
            got = [r.as_dict() for r in b.run(pairs, Q.IVQM_M_ALL)]
            for i, (g, w) in enumerate(zip(got, want)):
                assert g == w, (rep, i, MIXED[i])


def test_batch_async_ring_and_subsets():
    items, host = _dev_pairs()
    with IvqmBatch() as b:
        pairs_all = b.make_pairs(items)
        pairs_a = b.make_pairs(items[:3])
        pairs_b = b.make_pairs(items[3:])
        ref = [r.as_dict() for r in b.run(pairs_all, Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM)]
        b.enqueue(pairs_a, Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM, 1)
        b.enqueue(pairs_b, Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM, 2)
        got = [r.as_dict() for r in b.collect(1, 3)] + [r.as_dict() for r in b.collect(2, 3)]
        assert got == ref


def test_batch_autotune_keeps_results():
    items, host = _dev_pairs()
    with IvqmBatch() as b:
        pairs = b.make_pairs(items)
        ref = [r.as_dict() for r in b.run(pairs, Q.IVQM_M_ALL)]
        py, nb = b.autotune(pairs, Q.IVQM_M_ALL)
        assert py in (16, 32) and nb in (4, 8, 16)
        print("autotuned: search tile rows", py, "ssim band rows", nb)
        for py2, nb2 in ((16, 4), (32, 16), (4, 8), (py, nb)):  # 4 and 8 are taken as 16 and 32
            b.set_tuning(py2, nb2, False)
            assert [r.as_dict() for r in b.run(pairs, Q.IVQM_M_ALL)] == ref


def test_batch_matches_oracle_on_a_mixed_pair(oracle):
    items, host = _dev_pairs()
    with IvqmBatch() as b:
        res = b.run(b.make_pairs(items), Q.IVQM_M_ALL)
    for (wl, p0n, p1n), r in zip(host, res):
        p0 = oracle.to444([p.astype(np.uint16) for p in p0n], wl.chroma_420)
        p1 = oracle.to444([p.astype(np.uint16) for p in p1n], wl.chroma_420)
        g = oracle.gcd(p0, p1, wl.bit_depth)
        assert list(r.gcd_r2t) == g
        v = oracle.ivpsnr(p0, p1, wl.bit_depth, g)
        assert (r.ivpsnr, r.ivpsnr_r2t, r.ivpsnr_t2r) == v
        s0, s1 = oracle.scp(p0, p1, wl.bit_depth, g)
        assert abs(r.ivssim - oracle.ivssim(p0, p1, s0, s1, wl.bit_depth)[0]) <= 1e-6


def test_batch_ring_results_survive_buffer_growth():
    """Batches waiting in ring slots keep their results while later, larger
    batches grow the engine's buffers (one pair per enqueue, growing sizes)."""
    items, host = _dev_pairs()
    order = sorted(range(len(items)), key=lambda i: MIXED[i].width * MIXED[i].height)
    want = {i: _single(*host[i], Q.IVQM_M_ALL) for i in order}
    with IvqmBatch() as b:
        for use_graph in (False, True):
            b.set_tuning(0, 0, use_graph)
            for ring, i in enumerate(order):
                b.enqueue(b.make_pairs([items[i]]), Q.IVQM_M_ALL, ring)
            for ring, i in enumerate(order):
                (got,) = b.collect(ring, 1)
                assert got.as_dict() == want[i], (use_graph, MIXED[i])


def _chroma_layout(torch, plane, layout):
    """A device copy of a 2-D uint16 plane: 'aligned' (contiguous), 'stride' (row stride not a multiple of
    4 samples) or 'offset' (base 2 bytes past an 8-byte boundary)."""
    h, w = plane.shape
    src = torch.from_numpy(plane.view(np.int16)).cuda()
    if layout == "aligned":
        return src.clone()
    if layout == "stride":
        buf = torch.zeros((h, w + 2), dtype=torch.int16, device="cuda")
        v = buf[:, :w]
    else:
        buf = torch.zeros(h * w + 8, dtype=torch.int16, device="cuda")
        v = buf[1:1 + h * w].view(h, w)
    v.copy_(src)
    return v


def test_batch_graph_replay_across_chroma_alignment(oracle):
    """MS-SSIM level 1 of a 4:2:0 uint16 picture takes the caller's chroma planes when they are aligned
    (luma-only pyramid job) and copies them otherwise. One captured graph must not be replayed across
    that difference: batches of the same shapes alternate aligned and unaligned chroma with hipGraph on,
    and every result equals the single-pair result (pictures uploaded into aligned slots)."""
    import torch
    wls = [synth.Workload(cfg=510, width=136, height=136), synth.Workload(cfg=511, width=136, height=136,
                                                                          global_shift=1)]
    host = [(wl, *synth.make_pair(wl, 0)) for wl in wls]
    m = Q.IVQM_M_SSIM | Q.IVQM_M_MSSSIM | Q.IVQM_M_IVSSIM | Q.IVQM_M_IVMSSSIM
    want = [_single(wl, p0, p1, m) for wl, p0, p1 in host]

    def items(layout):
        out = []
        for wl, p0, p1 in host:
            pics = []
            for p in (p0, p1):
                luma = torch.from_numpy(p[0].view(np.int16)).cuda()
                pics.append([luma] + [_chroma_layout(torch, p[c], layout) for c in (1, 2)])
            out.append((pics[0], pics[1], wl.width, wl.height, wl.bit_depth, True))
        torch.cuda.synchronize()
        return out

    with IvqmBatch() as b:
        b.set_tuning(0, 0, True)
        for layout in ("aligned", "stride", "aligned", "offset", "stride", "aligned"):
            it = items(layout)
            got = [r.as_dict() for r in b.run(b.make_pairs(it), m)]
            for i, (g, w) in enumerate(zip(got, want)):
                assert g == w, (layout, i)
    from oracle_checks import check_pair
    for (wl, p0, p1), r in zip(host, want):
        check_pair(oracle, r, p0, p1, wl.bit_depth, True, m, label=str(wl))
