"""ORACLE - test infrastructure only.

ctypes bindings for the CPU checkers:
  * ``Oracle``: the C restatement of QMIV's metric hot path (oracle/ivqm_oracle.c)
  * ``Reference``: the reference pmbb libraries compiled from their own sources
    plus oracle/ref_driver.cpp (oracle/_ref/libqmiv_ref_v{3,4}.so)

Only tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of bench.py
import this module; the product path (qmiv_amd) never does.

All picture arguments are sequences of three 2-D numpy uint16 arrays (4:4:4
planar, C-contiguous).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_u16p = C.POINTER(C.c_uint16)
_u64p = C.POINTER(C.c_uint64)
_i32p = C.POINTER(C.c_int32)
_f64p = C.POINTER(C.c_double)
_f32p = C.POINTER(C.c_float)


def build() -> None:
    """Builds liboracle.so (and oracle/_ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "-j8"], check=True)


def _planes(p):
    arrs = [np.ascontiguousarray(a, dtype=np.uint16) for a in p]
    ptrs = (_u16p * 3)(*[a.ctypes.data_as(_u16p) for a in arrs])
    return arrs, ptrs


def _i32x4(v):
    return (C.c_int32 * 4)(*[int(x) for x in v])


def _f32x4(v):
    return (C.c_float * 4)(*[float(x) for x in v])


def host_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            flags = next((ln for ln in f if ln.startswith("flags")), "")
    except OSError:
        return False
    need = ("avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl")
    return all(n in flags.split() for n in need)


class Oracle:
    """The C restatement (see ivqm_oracle.c for the reference lines each follows)."""

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        self.L = L
        L.or_kbns_sum.restype = C.c_double
        L.or_kbns_sum.argtypes = [_f64p, C.c_int64]
        L.or_round_f32.restype = C.c_int32
        L.or_round_f32.argtypes = [C.c_float]
        L.or_round_f64.restype = C.c_int32
        L.or_round_f64.argtypes = [C.c_double]
        L.or_sd.restype = C.c_int64
        L.or_sd.argtypes = [_u16p, _u16p, C.c_int64, C.c_int32]
        L.or_ssd.restype = C.c_uint64
        L.or_ssd.argtypes = [_u16p, _u16p, C.c_int64]
        L.or_psnr_from_ssd.restype = C.c_double
        L.or_psnr_from_ssd.argtypes = [C.c_double, C.c_int64, C.c_int32]
        L.or_ivpsnr.restype = C.c_double
        L.or_ivpsnr_quality.restype = C.c_double
        L.or_ssim_window.restype = C.c_double
        L.or_ssim_plane.restype = C.c_double
        L.or_ivssim.restype = C.c_double
        L.or_ivmsssim.restype = C.c_double
        L.or_pic_value.restype = C.c_double
        L.or_ssim_num_blocks.restype = C.c_int32
        L.or_ms_num_scales.restype = C.c_int32
        L.or_ssim_plane_ex.restype = C.c_double
        L.or_ivpsnr_masked.restype = C.c_double
        L.or_psnr_masked.restype = C.c_double
        L.or_count_nonzero.restype = C.c_int64
        L.or_ivssim_ex.restype = C.c_double
        L.or_ivmsssim_ex.restype = C.c_double

    # -- pixel ops ---------------------------------------------------------
    def upsample(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint16)
        h, w = src.shape
        dst = np.empty((2 * h, 2 * w), np.uint16)
        self.L.or_upsample_hv(dst.ctypes.data_as(_u16p), src.ctypes.data_as(_u16p), 2 * w, 2 * h)
        return dst

    def downsample(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint16)
        h, w = src.shape
        dst = np.empty((h >> 1, w >> 1), np.uint16)
        self.L.or_downsample_hv(dst.ctypes.data_as(_u16p), src.ctypes.data_as(_u16p), w, h)
        return dst

    def upsample_h(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint16)
        h, w = src.shape
        dst = np.empty((h, 2 * w), np.uint16)
        self.L.or_upsample_h(dst.ctypes.data_as(_u16p), src.ctypes.data_as(_u16p), 2 * w, h)
        return dst

    def to444(self, planes, chroma420):
        """chroma420: the chroma format code of the planes, 0 (or False) 4:4:4, 1 (or True) 4:2:0, 2 4:2:2."""
        cf = int(chroma420)
        if cf == 0:
            return [np.ascontiguousarray(p, np.uint16) for p in planes]
        up = self.upsample if cf == 1 else self.upsample_h
        return [np.ascontiguousarray(planes[0], np.uint16), up(planes[1]), up(planes[2])]

    # -- scalar helpers ------------------------------------------------------
    def kbns(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return self.L.or_kbns_sum(x.ctypes.data_as(_f64p), len(x))

    def round_f32(self, v):
        return self.L.or_round_f32(float(v))

    def round_f64(self, v):
        return self.L.or_round_f64(float(v))

    def psnr_from_ssd(self, ssd, area, bd):
        return self.L.or_psnr_from_ssd(float(ssd), int(area), int(bd))

    def sd(self, a, b, wrap32=False):
        a = np.ascontiguousarray(a, np.uint16)
        b = np.ascontiguousarray(b, np.uint16)
        return self.L.or_sd(a.ctypes.data_as(_u16p), b.ctypes.data_as(_u16p), a.size, int(wrap32))

    def ssd(self, a, b):
        a = np.ascontiguousarray(a, np.uint16)
        b = np.ascontiguousarray(b, np.uint16)
        return self.L.or_ssd(a.ctypes.data_as(_u16p), b.ctypes.data_as(_u16p), a.size)

    def ws_weights(self, h, lat=180):
        w = np.zeros(h, np.float64)
        corr = C.c_double(0)
        self.L.or_ws_weights(h, lat, w.ctypes.data_as(_f64p), C.byref(corr))
        return w, corr.value

    # -- metrics ---------------------------------------------------------------
    def gcd(self, p0, p1, bd, coef=(0.01, 0.01, 0.01, 0.0), wrap32=None):
        """App direction: GCD_R2T = round(mean(P1 - P0)) (CalcGlobalColorDiff(Tst=P1, Ref=P0))."""
        if wrap32 is None:
            wrap32 = bd <= 14
        a, pa = _planes(p1)
        b, pb = _planes(p0)
        h, w = a[0].shape
        out = (C.c_int32 * 4)()
        self.L.or_gcd(pa, pb, w, h, bd, _f32x4(coef), int(wrap32), out)
        return [out[i] for i in range(4)]

    def cps(self, tst, ref, bd, gcs, R=2, cw=(4, 1, 1, 0), want_scp=False, want_best=False):
        a, pa = _planes(tst)
        b, pb = _planes(ref)
        h, w = a[0].shape
        rows = np.zeros((h, 4), np.uint64)
        scp = [np.zeros((h, w), np.uint16) for _ in range(3)] if want_scp else None
        scp_p = (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in scp]) if want_scp else None
        best = np.zeros((h, w), np.uint8) if want_best else None
        self.L.or_cps(pa, pb, w, h, bd, _i32x4(gcs), R, _i32x4(cw), rows.ctypes.data_as(_u64p), scp_p,
                      best.ctypes.data_as(C.POINTER(C.c_uint8)) if want_best else None)
        return rows, scp, best

    def ivpsnr_quality(self, rows, w, h, bd, erp_w=None, cw_avg=(4, 1, 1, 0)):
        rows = np.ascontiguousarray(rows, np.uint64)
        ew = np.ascontiguousarray(erp_w, np.float64) if erp_w is not None else None
        return self.L.or_ivpsnr_quality(rows.ctypes.data_as(_u64p), w, h, bd,
                                        ew.ctypes.data_as(_f64p) if ew is not None else None, _i32x4(cw_avg))

    def ivpsnr(self, p0, p1, bd, gcd, R=2, cws=(4, 1, 1, 0), cwa=(4, 1, 1, 0), erp=False, lat=180, want_rows=False):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        ew = self.ws_weights(h, lat)[0] if erp else None
        r2t, t2r = C.c_double(0), C.c_double(0)
        ra = np.zeros((h, 4), np.uint64)
        rb = np.zeros((h, 4), np.uint64)
        v = self.L.or_ivpsnr(pa, pb, w, h, bd, _i32x4(gcd), R, _i32x4(cws), _i32x4(cwa),
                             ew.ctypes.data_as(_f64p) if ew is not None else None, C.byref(r2t), C.byref(t2r),
                             ra.ctypes.data_as(_u64p), rb.ctypes.data_as(_u64p))
        if want_rows:
            return v, r2t.value, t2r.value, ra, rb
        return v, r2t.value, t2r.value

    def scp(self, p0, p1, bd, gcd, R=2, cws=(4, 1, 1, 0)):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        s0 = [np.zeros((h, w), np.uint16) for _ in range(3)]
        s1 = [np.zeros((h, w), np.uint16) for _ in range(3)]
        self.L.or_gen_scp(pa, pb, w, h, bd, _i32x4(gcd), R, _i32x4(cws),
                          (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s0]),
                          (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s1]))
        return s0, s1

    def psnr(self, p0, p1, bd):
        return [self.psnr_from_ssd(self.ssd(p1[c], p0[c]), p0[0].size, bd) for c in range(3)]

    def wspsnr(self, p0, p1, bd, erp=True, lat=180):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 3)()
        self.L.or_wspsnr(pa, pb, w, h, bd, int(erp), lat, out)
        return [out[i] for i in range(3)]

    def ssim(self, p0, p1, bd, wnd=8, stride=4):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 3)()
        self.L.or_ssim(pa, pb, w, h, bd, wnd, stride, out)
        return [out[i] for i in range(3)]

    def msssim(self, p0, p1, bd, wnd=8, stride=4):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 3)()
        self.L.or_msssim(pa, pb, w, h, bd, wnd, stride, out)
        return [out[i] for i in range(3)]

    def ssim_plane(self, t, r, bd, wnd=8, stride=4, calc_l=True, erp_w=None):
        t = np.ascontiguousarray(t, np.uint16)
        r = np.ascontiguousarray(r, np.uint16)
        h, w = t.shape
        maxv = float((1 << bd) - 1)
        c1 = (0.01 * maxv) * (0.01 * maxv)
        c2 = (0.03 * maxv) * (0.03 * maxv)
        ew = np.ascontiguousarray(erp_w, np.float64) if erp_w is not None else None
        return self.L.or_ssim_plane(t.ctypes.data_as(_u16p), r.ctypes.data_as(_u16p), w, h, wnd, stride,
                                    C.c_double(c1), C.c_double(c2), int(calc_l),
                                    ew.ctypes.data_as(_f64p) if ew is not None else None)

    def ivssim(self, p0, p1, scp0, scp1, bd, wnd=8, stride=4, cwa=(4, 1, 1, 0), erp=False, lat=180):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        s0, ps0 = _planes(scp0)
        s1, ps1 = _planes(scp1)
        h, w = a[0].shape
        ew = self.ws_weights(h, lat)[0] if erp else None
        r2t, t2r = C.c_double(0), C.c_double(0)
        v = self.L.or_ivssim(pa, pb, ps0, ps1, w, h, bd, wnd, stride, _i32x4(cwa),
                             ew.ctypes.data_as(_f64p) if ew is not None else None, C.byref(r2t), C.byref(t2r))
        return v, r2t.value, t2r.value

    def ivmsssim(self, p0, p1, scp0, scp1, bd, wnd=8, stride=4, cwa=(4, 1, 1, 0)):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        s0, ps0 = _planes(scp0)
        s1, ps1 = _planes(scp1)
        h, w = a[0].shape
        return self.L.or_ivmsssim(pa, pb, ps0, ps1, w, h, bd, wnd, stride, _i32x4(cwa))

    def pic_value(self, v3, cw=(4, 1, 1, 0)):
        return self.L.or_pic_value((C.c_double * 3)(*v3), _i32x4(cw))

    # -- mask mode (the mask: one uint16 plane at luma resolution, bit depth bdm) --
    def count_nonzero(self, m):
        m = np.ascontiguousarray(m, np.uint16)
        return self.L.or_count_nonzero(m.ctypes.data_as(_u16p), C.c_int64(m.size))

    def weighted_ssd(self, a, b, msk):
        a = np.ascontiguousarray(a, np.uint16)
        b = np.ascontiguousarray(b, np.uint16)
        m = np.ascontiguousarray(msk, np.uint16)
        self.L.or_weighted_ssd.restype = C.c_uint64
        return self.L.or_weighted_ssd(a.ctypes.data_as(_u16p), b.ctypes.data_as(_u16p), m.ctypes.data_as(_u16p),
                                      C.c_int64(m.size))

    def gcd_masked(self, p0, p1, msk, bd, bdm, coef=(0.01, 0.01, 0.01, 0.0)):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        out = (C.c_int32 * 4)()
        self.L.or_gcd_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, bdm, _f32x4(coef), out)
        return [out[i] for i in range(4)]

    def psnr_family_masked(self, p0, p1, msk, bd, bdm, erp=False, lat=180):
        """(mse[3], psnr[3], wspsnr[3]) as the app computes them with a mask."""
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        ew, corr = self.ws_weights(h, lat) if erp else (None, 1.0)
        mse, psnr, ws = (C.c_double * 3)(), (C.c_double * 3)(), (C.c_double * 3)()
        self.L.or_psnr_family_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, bdm,
                                     ew.ctypes.data_as(_f64p) if ew is not None else None, C.c_double(corr),
                                     mse, psnr, ws)
        return list(mse), list(psnr), list(ws)

    def cps_masked(self, tst, ref, msk, gcs, R=2, cw=(4, 1, 1, 0)):
        a, pa = _planes(tst)
        b, pb = _planes(ref)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        rows = np.zeros((h, 4), np.uint64)
        self.L.or_cps_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, _i32x4(gcs), R, _i32x4(cw),
                             rows.ctypes.data_as(_u64p))
        return rows

    def scp_masked(self, p0, p1, msk, bd, gcd, R=2, cws=(4, 1, 1, 0), fill=0):
        """(SCP0, SCP1) with a mask; pixels whose mask sample is 0 keep `fill`."""
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        s0 = [np.full((h, w), fill, np.uint16) for _ in range(3)]
        s1 = [np.full((h, w), fill, np.uint16) for _ in range(3)]
        self.L.or_gen_scp_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, _i32x4(gcd), R, _i32x4(cws),
                                 (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s0]),
                                 (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s1]))
        return s0, s1

    def ivpsnr_masked(self, p0, p1, msk, bd, bdm, gcd, R=2, cws=(4, 1, 1, 0), cwa=(4, 1, 1, 0), erp=False, lat=180):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        ew = self.ws_weights(h, lat)[0] if erp else None
        r2t, t2r = C.c_double(0), C.c_double(0)
        v = self.L.or_ivpsnr_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, bdm, _i32x4(gcd), R, _i32x4(cws),
                                    _i32x4(cwa), ew.ctypes.data_as(_f64p) if ew is not None else None,
                                    C.byref(r2t), C.byref(t2r))
        return v, r2t.value, t2r.value

    # -- colour space (dir 0: YCbCr -> RGB, 1: RGB -> YCbCr; space 0 BT.601, 1 BT.709, 2 SMPTE240M, 3 BT.2020) --
    def convert_colour(self, planes, bd, direction, space):
        out = [np.ascontiguousarray(p, np.uint16).copy() for p in planes]
        fn = self.L.or_ycbcr2rgb if direction == 0 else self.L.or_rgb2ycbcr
        fn(*[o.ctypes.data_as(_u16p) for o in out], C.c_int64(out[0].size), bd, space)
        return out

    # -- structural-similarity modes (mode: xSSIM::eMode 0..4, ext: eMrgExt 0..5) --
    def ssim_weights(self, which):
        n = {0: 121, 1: 121, 2: 64, 3: 256}[which]
        out = np.zeros(n, np.float64)
        self.L.or_ssim_weights(which, out.ctypes.data_as(_f64p))
        k = int(round(n ** 0.5))
        return out.reshape(k, k)

    def ssim_grid(self, w, h, mode, ext=0, wnd=8, stride=4):
        out = (C.c_int32 * 4)()
        self.L.or_ssim_grid(w, h, mode, ext, wnd, stride, out)
        return tuple(out[i] for i in range(4))  # first x, count x, first y, count y

    def ssim_ex(self, p0, p1, bd, mode, ext=0, wnd=8, stride=4, ms=False):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 3)()
        fn = self.L.or_msssim_ex if ms else self.L.or_ssim_ex
        fn(pa, pb, w, h, bd, mode, ext, wnd, stride, out)
        return [out[i] for i in range(3)]

    def ssim_plane_ex(self, t, r, bd, mode, ext=0, wnd=8, stride=4, calc_l=True, erp_w=None):
        t = np.ascontiguousarray(t, np.uint16)
        r = np.ascontiguousarray(r, np.uint16)
        h, w = t.shape
        maxv = float((1 << bd) - 1)
        c1 = (0.01 * maxv) * (0.01 * maxv)
        c2 = (0.03 * maxv) * (0.03 * maxv)
        ew = np.ascontiguousarray(erp_w, np.float64) if erp_w is not None else None
        return self.L.or_ssim_plane_ex(t.ctypes.data_as(_u16p), r.ctypes.data_as(_u16p), w, h, mode, ext, wnd,
                                       stride, C.c_double(c1), C.c_double(c2), int(calc_l),
                                       ew.ctypes.data_as(_f64p) if ew is not None else None)

    def ivssim_ex(self, p0, p1, scp0, scp1, bd, mode, ext=0, wnd=8, stride=4, cwa=(4, 1, 1, 0), erp=False, lat=180,
                  ms=False):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        s0, ps0 = _planes(scp0)
        s1, ps1 = _planes(scp1)
        h, w = a[0].shape
        if ms:
            return self.L.or_ivmsssim_ex(pa, pb, ps0, ps1, w, h, bd, mode, ext, wnd, stride, _i32x4(cwa)), None, None
        ew = self.ws_weights(h, lat)[0] if erp else None
        r2t, t2r = C.c_double(0), C.c_double(0)
        v = self.L.or_ivssim_ex(pa, pb, ps0, ps1, w, h, bd, mode, ext, wnd, stride, _i32x4(cwa),
                                ew.ctypes.data_as(_f64p) if ew is not None else None, C.byref(r2t), C.byref(t2r))
        return v, r2t.value, t2r.value

    def num_blocks(self, n, wnd=8, stride=4):
        return self.L.or_ssim_num_blocks(n, wnd, stride)

    def ms_num_scales(self, w, h):
        return self.L.or_ms_num_scales(w, h)


def ref_library_path(level: str | None = None) -> str:
    if level is None:
        level = "v4" if host_has_avx512() else "v3"
    return os.path.join(REF_DIR, f"libqmiv_ref_{level}.so")


def reference_available(level: str | None = None) -> bool:
    return os.path.exists(ref_library_path(level))


class Reference:
    """The reference libraries themselves (built by oracle/Makefile into oracle/_ref)."""

    def __init__(self, level: str | None = None):
        self.level = level or ("v4" if host_has_avx512() else "v3")
        L = C.CDLL(ref_library_path(self.level))
        self.L = L
        L.ref_sd.restype = C.c_int64
        L.ref_ssd.restype = C.c_uint64
        L.ref_bench_pairs.restype = C.c_double
        L.ref_bench_pairs.argtypes = [C.c_int32, C.POINTER(_u16p), C.POINTER(_u16p), C.c_int32, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _f64p, _f64p, _f64p]
        L.ref_hardware_concurrency.restype = C.c_uint

    def upsample(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint16)
        h, w = src.shape
        dst = np.empty((2 * h, 2 * w), np.uint16)
        self.L.ref_upsample_hv(dst.ctypes.data_as(_u16p), src.ctypes.data_as(_u16p), 2 * w, 2 * h)
        return dst

    def upsample_h(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint16)
        h, w = src.shape
        dst = np.empty((h, 2 * w), np.uint16)
        self.L.ref_upsample_h(dst.ctypes.data_as(_u16p), src.ctypes.data_as(_u16p), 2 * w, h)
        return dst

    def downsample(self, src):
        src = np.ascontiguousarray(src, dtype=np.uint16)
        h, w = src.shape
        dst = np.empty((h >> 1, w >> 1), np.uint16)
        self.L.ref_downsample_hv(dst.ctypes.data_as(_u16p), src.ctypes.data_as(_u16p), w, h)
        return dst

    def sd(self, a, b, bd):
        a = np.ascontiguousarray(a, np.uint16)
        b = np.ascontiguousarray(b, np.uint16)
        h, w = a.shape
        return self.L.ref_sd(a.ctypes.data_as(_u16p), b.ctypes.data_as(_u16p), w, h, bd)

    def ssd(self, a, b, bd):
        a = np.ascontiguousarray(a, np.uint16)
        b = np.ascontiguousarray(b, np.uint16)
        h, w = a.shape
        return self.L.ref_ssd(a.ctypes.data_as(_u16p), b.ctypes.data_as(_u16p), w, h, bd)

    def gcd(self, p0, p1, bd, coef=(0.01, 0.01, 0.01, 0.0), R=2):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_int32 * 4)()
        self.L.ref_gcd(pa, pb, w, h, bd, _f32x4(coef), R, out)
        return [out[i] for i in range(4)]

    def cps_rows(self, tst, ref, bd, gcs, R=2, cw=(4, 1, 1, 0)):
        a, pa = _planes(tst)
        b, pb = _planes(ref)
        h, w = a[0].shape
        rows = np.zeros((h, 4), np.uint64)
        self.L.ref_cps_rows(pa, pb, w, h, bd, _i32x4(gcs), R, _i32x4(cw), rows.ctypes.data_as(_u64p))
        return rows

    def scp(self, p0, p1, bd, gcd, R=2, cw=(4, 1, 1, 0)):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        s0 = [np.zeros((h, w), np.uint16) for _ in range(3)]
        s1 = [np.zeros((h, w), np.uint16) for _ in range(3)]
        self.L.ref_scp(pa, pb, w, h, bd, _i32x4(gcd), R, _i32x4(cw),
                       (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s0]),
                       (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s1]))
        return s0, s1

    def psnr(self, p0, p1, bd):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 4)()
        self.L.ref_psnr(pa, pb, w, h, bd, out)
        return [out[i] for i in range(3)]

    def wspsnr(self, p0, p1, bd, erp=True, lat=180):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 4)()
        self.L.ref_wspsnr(pa, pb, w, h, bd, int(erp), lat, out)
        return [out[i] for i in range(3)]

    def ivpsnr(self, p0, p1, bd, gcd, R=2, cws=(4, 1, 1, 0), cwa=(4, 1, 1, 0), erp=False, lat=180):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        v, r2t, t2r = C.c_double(0), C.c_double(0), C.c_double(0)
        self.L.ref_ivpsnr(pa, pb, w, h, bd, _i32x4(gcd), R, _i32x4(cws), _i32x4(cwa), int(erp), lat,
                          C.byref(v), C.byref(r2t), C.byref(t2r))
        return v.value, r2t.value, t2r.value

    def ssim(self, p0, p1, bd, wnd=8, stride=4, ms=False):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 4)()
        self.L.ref_ssim(pa, pb, w, h, bd, wnd, stride, int(ms), out)
        return [out[i] for i in range(3)]

    def ivssim(self, p0, p1, scp0, scp1, bd, wnd=8, stride=4, cwa=(4, 1, 1, 0), erp=False, lat=180, ms=False):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        s0, ps0 = _planes(scp0)
        s1, ps1 = _planes(scp1)
        h, w = a[0].shape
        v = C.c_double(0)
        self.L.ref_ivssim(pa, pb, ps0, ps1, w, h, bd, wnd, stride, _i32x4(cwa), int(erp), lat, int(ms), C.byref(v))
        return v.value

    def ssim_ex(self, p0, p1, bd, mode, ext=0, wnd=8, stride=4, margin=8, ms=False):
        """Any mode / border extension on pictures with a margin of `margin` samples."""
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        h, w = a[0].shape
        out = (C.c_double * 4)()
        self.L.ref_ssim_ex(pa, pb, w, h, bd, mode, ext, wnd, stride, margin, int(ms), out)
        return [out[i] for i in range(3)]

    def ivssim_ex(self, p0, p1, scp0, scp1, bd, mode, ext=0, wnd=8, stride=4, margin=8, cwa=(4, 1, 1, 0),
                  erp=False, lat=180, ms=False):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        s0, ps0 = _planes(scp0)
        s1, ps1 = _planes(scp1)
        h, w = a[0].shape
        v, r2t, t2r = C.c_double(0), C.c_double(0), C.c_double(0)
        self.L.ref_ivssim_ex(pa, pb, ps0, ps1, w, h, bd, mode, ext, wnd, stride, margin, _i32x4(cwa), int(erp), lat,
                             int(ms), C.byref(v), C.byref(r2t), C.byref(t2r))
        return v.value, r2t.value, t2r.value

    def gcd_masked(self, p0, p1, msk, bd, bdm, R=2, coef=(0.01, 0.01, 0.01, 0.0)):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        out = (C.c_int32 * 4)()
        self.L.ref_gcd_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, bdm, R, _f32x4(coef), out)
        return [out[i] for i in range(4)]

    def psnr_family_masked(self, p0, p1, msk, bd, bdm, erp=False, lat=180):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        mse, psnr, ws = (C.c_double * 4)(), (C.c_double * 4)(), (C.c_double * 4)()
        self.L.ref_psnr_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, bdm, int(erp), lat, mse, psnr, ws)
        return list(mse)[:3], list(psnr)[:3], list(ws)[:3]

    def scp_masked(self, p0, p1, msk, bd, gcd, R=2, cws=(4, 1, 1, 0), fill=0):
        """(SCP0, SCP1) with a mask; pixels whose mask sample is 0 keep `fill`."""
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        s0 = [np.full((h, w), fill, np.uint16) for _ in range(3)]
        s1 = [np.full((h, w), fill, np.uint16) for _ in range(3)]
        self.L.ref_scp_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, _i32x4(gcd), R, _i32x4(cws), int(fill),
                              (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s0]),
                              (_u16p * 3)(*[s.ctypes.data_as(_u16p) for s in s1]))
        return s0, s1

    def ivpsnr_masked(self, p0, p1, msk, bd, bdm, gcd, R=2, cws=(4, 1, 1, 0), cwa=(4, 1, 1, 0), erp=False, lat=180):
        a, pa = _planes(p0)
        b, pb = _planes(p1)
        m = np.ascontiguousarray(msk, np.uint16)
        h, w = a[0].shape
        v, r2t, t2r = C.c_double(0), C.c_double(0), C.c_double(0)
        self.L.ref_ivpsnr_masked(pa, pb, m.ctypes.data_as(_u16p), w, h, bd, bdm, _i32x4(gcd), R, _i32x4(cws),
                                 _i32x4(cwa), int(erp), lat, C.byref(v), C.byref(r2t), C.byref(t2r))
        return v.value, r2t.value, t2r.value

    def convert_colour(self, planes, bd, direction, space):
        out = [np.ascontiguousarray(p, np.uint16).copy() for p in planes]
        h, w = out[0].shape
        self.L.ref_convert_colour(*[o.ctypes.data_as(_u16p) for o in out], w, h, bd, direction, space)
        return out

    def hardware_concurrency(self):
        return int(self.L.ref_hardware_concurrency())

    def bench_pairs(self, pairs, w, h, bd, chroma420, erp, metrics, threads, stages=None):
        """Times the app-order frame loop over in-memory pairs; returns (seconds, ivpsnr[], ivssim[]).
        `stages`, a list, receives the seconds per stage (unpack, PREPROC, Rearrange, GCD, SCP, IV-PSNR,
        IV-SSIM) summed over the pairs."""
        n = len(pairs)
        keep = []
        p0 = (_u16p * (3 * n))()
        p1 = (_u16p * (3 * n))()
        for f, (a, b) in enumerate(pairs):
            for c in range(3):
                aa = np.ascontiguousarray(a[c], np.uint16)
                bb = np.ascontiguousarray(b[c], np.uint16)
                keep += [aa, bb]
                p0[3 * f + c] = aa.ctypes.data_as(_u16p)
                p1[3 * f + c] = bb.ctypes.data_as(_u16p)
        iv = np.zeros(n, np.float64)
        ss = np.zeros(n, np.float64)
        st = np.zeros(7, np.float64)
        t = self.L.ref_bench_pairs(n, p0, p1, w, h, bd, int(chroma420), int(erp), metrics, threads,
                                   iv.ctypes.data_as(_f64p), ss.ctypes.data_as(_f64p), st.ctypes.data_as(_f64p))
        if stages is not None:
            stages[:] = st.tolist()
        return t, iv, ss
