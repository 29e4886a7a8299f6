"""The search's expanded reference score (mode 6, for samples whose range rules out the keyed forms:
9- to 16-bit 4:4:4 and 16-bit 4:2:0 at the default weights) against the reference-form arithmetic
(search_mode_cap 1) and the oracle: rows, winner maps and SCP pictures bit for bit, IV-PSNR repr-equal.
xCorrespPixelShiftSSE.cpp:42-70 and :96-127 define the semantics: the int32 score wraps, the first
strict minimum in raster order wins, and when no candidate scores below INT_MAX none is taken (zero
distances, zero winner sample)."""
import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import synth

pytestmark = pytest.mark.gpu


def _run(p0n, p1n, w, h, bd, cf, R, cap, **kw):
    with Q.IvqmContext(w, h, bd, search_range=R, search_mode_cap=cap, **kw) as ctx:
        ctx.upload_pair(p0n, p1n, cf)
        ctx.enable_best_map(True)
        r = ctx.frame(Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM)
        return (r.as_dict(), ctx.rows(0).copy(), ctx.rows(1).copy(), ctx.best_map(0).copy(), ctx.best_map(1).copy(),
                [ctx.scp(0)[c].copy() for c in range(3)], [ctx.scp(1)[c].copy() for c in range(3)])


def _check(oracle, p0n, p1n, w, h, bd, cf, R, cw=(4, 1, 1, 0), **kw):
    print(f"quad case {w}x{h} {bd}-bit cf{cf} R={R} cw={cw}", flush=True)
    got = _run(p0n, p1n, w, h, bd, cf, R, 0, cmp_weights_search=cw, **kw)
    ref = _run(p0n, p1n, w, h, bd, cf, R, 1, cmp_weights_search=cw, **kw)
    assert got[0] == ref[0]
    for a, b in zip(got[1:5], ref[1:5]):
        assert np.array_equal(a, b)
    for c in range(3):
        assert np.array_equal(got[5][c], ref[5][c]) and np.array_equal(got[6][c], ref[6][c])
    p0 = oracle.to444([p.astype(np.uint16) for p in p0n], cf)
    p1 = oracle.to444([p.astype(np.uint16) for p in p1n], cf)
    g = list(got[0]["gcd_r2t"])
    rows_a, scp1, best_a = oracle.cps(p0, p1, bd, g, R=R, cw=cw, want_scp=True, want_best=True)
    rows_b, scp0, best_b = oracle.cps(p1, p0, bd, [-x for x in g], R=R, cw=cw, want_scp=True, want_best=True)
    assert np.array_equal(got[1][:, :3], rows_a[:, :3]) and np.array_equal(got[2][:, :3], rows_b[:, :3])
    assert np.array_equal(got[3], best_a) and np.array_equal(got[4], best_b)
    for c in range(3):
        assert np.array_equal(got[5][c], scp0[c]) and np.array_equal(got[6][c], scp1[c])
    ov = oracle.ivpsnr(p0, p1, bd, g, R=R, cws=cw)
    assert (repr(got[0]["ivpsnr"]), repr(got[0]["ivpsnr_r2t"]), repr(got[0]["ivpsnr_t2r"])) == \
        tuple(repr(x) for x in ov)
    return got


@pytest.mark.parametrize("R", [1, 2, 3, 4])
@pytest.mark.parametrize("cf", [0, 1])
def test_quad_16bit_stress_content(oracle, R, cf):
    """Stress-workload content (16-bit over the full range, noise sigma 12, blocks shifted by up to 3 px) at
    a ragged size: scores wrap for far candidates, and the expanded form must keep every wrapped value."""
    wl = synth.Workload(cfg=4, width=200, height=84, bit_depth=16, chroma_420=bool(cf), sigma_r=12.0,
                        sigma_t=12.0, block_shift=3, offset_rand=8)
    p0n, p1n = synth.make_pair(wl, 1)
    _check(oracle, p0n, p1n, wl.width, wl.height, 16, cf, R)


@pytest.mark.parametrize("bd", [11, 13, 14])
def test_quad_444_other_depths(oracle, bd):
    wl = synth.Workload(cfg=711, width=136, height=72, bit_depth=bd, chroma_420=False, global_shift=1, block_shift=2)
    p0n, p1n = synth.make_pair(wl, 0)
    _check(oracle, p0n, p1n, wl.width, wl.height, bd, 0, 2)


@pytest.mark.parametrize("cw", [(3, 2, 5, 0), (4, -1, 2, 0), (1, 1, 1, 0)])
def test_quad_other_weights(oracle, cw):
    wl = synth.Workload(cfg=712, width=96, height=48, bit_depth=16, chroma_420=False, sigma_r=12.0, sigma_t=12.0,
                        block_shift=2)
    p0n, p1n = synth.make_pair(wl, 0)
    _check(oracle, p0n, p1n, wl.width, wl.height, 16, 0, 2, cw=cw)


def test_quad_no_candidate_below_int_max(oracle):
    """Uniform pictures whose every candidate scores 2 a^2 + b^2 + c^2 = 2^31 - 1 (weights 2:1:1, no GCD):
    the reference takes no candidate, so the row sums are zero and both SCP pictures are zero."""
    w, h = 72, 40
    a, b, c = 30013, 1390, 18547
    assert (2 * a * a + b * b + c * c) % 2**32 == 2**31 - 1
    p0n = [np.full((h, w), v, np.uint16) for v in (a, b, c)]
    p1n = [np.zeros((h, w), np.uint16) for _ in range(3)]
    got = _check(oracle, p0n, p1n, w, h, 16, 0, 2, cw=(2, 1, 1, 0), unnoticeable_coef=(0.0, 0.0, 0.0, 0.0))
    assert not got[1].any() and not got[2].any()
    assert all(not got[5][k].any() and not got[6][k].any() for k in range(3))
