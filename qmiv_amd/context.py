"""Python handle over one C-ABI context (tests, bench and smoke use it).

Pictures are given as three numpy planes (host, uploaded) or three torch CUDA
tensors (device-resident, bound without a copy).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import Config, FrameResult, Picture, check


def default_config(width, height, bit_depth=10, **kw) -> Config:
    L = _lib.load()
    cfg = Config()
    check(L.ivqm_config_init(C.byref(cfg)))
    cfg.width, cfg.height, cfg.bit_depth = width, height, bit_depth
    for k, v in kw.items():
        if k in ("cmp_weights_search", "cmp_weights_average"):
            getattr(cfg, k)[:] = list(v)
        elif k == "unnoticeable_coef":
            cfg.unnoticeable_coef[:] = [float(x) for x in v]
        else:
            setattr(cfg, k, int(v))
    return cfg


def _picture_from_numpy(planes, chroma_420: bool):
    arrs = [np.ascontiguousarray(p) for p in planes]
    bps = arrs[0].dtype.itemsize
    pic = Picture()
    for k in range(3):
        pic.plane[k] = arrs[k].ctypes.data
        pic.stride[k] = arrs[k].shape[1]
    pic.chroma_420 = int(chroma_420)
    pic.bytes_per_sample = bps
    return pic, arrs


def _picture_from_torch(planes, chroma_420: bool):
    pic = Picture()
    for k in range(3):
        t = planes[k]
        assert t.is_cuda and t.dim() == 2 and t.stride(1) == 1
        pic.plane[k] = t.data_ptr()
        pic.stride[k] = t.stride(0)
    pic.chroma_420 = int(chroma_420)
    pic.bytes_per_sample = planes[0].element_size()
    return pic


class IvqmContext:
    def __init__(self, width, height, bit_depth=10, **kw):
        self.L = _lib.load()
        self.cfg = default_config(width, height, bit_depth, **kw)
        h = C.c_void_p()
        check(self.L.ivqm_create(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self.width, self.height, self.bit_depth = width, height, bit_depth
        self._keep = [None, None]

    def close(self):
        if self.h:
            self.L.ivqm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- pictures ------------------------------------------------------------
    def upload(self, slot, planes, chroma_420=True):
        pic, arrs = _picture_from_numpy(planes, chroma_420)
        check(self.L.ivqm_upload(self.h, slot, C.byref(pic)))

    def bind(self, slot, planes, chroma_420=True):
        pic = _picture_from_torch(planes, chroma_420)
        self._keep[slot] = planes
        check(self.L.ivqm_bind(self.h, slot, C.byref(pic)))

    def upload_pair(self, p0, p1, chroma_420=True):
        self.upload(0, p0, chroma_420)
        self.upload(1, p1, chroma_420)

    # -- mask mode -------------------------------------------------------------
    def upload_mask(self, plane, bit_depth):
        """plane: 2-D numpy uint8 / uint16 weights at luma resolution."""
        a = np.ascontiguousarray(plane)
        m = _lib.Mask(a.ctypes.data, a.shape[1], a.dtype.itemsize, bit_depth, 0)
        check(self.L.ivqm_upload_mask(self.h, C.byref(m)))

    def bind_mask(self, plane, bit_depth=8):
        """plane: 2-D torch CUDA tensor (int16 / uint8 storage of the weights), or None to unbind."""
        if plane is None:
            self._mask = None
            check(self.L.ivqm_bind_mask(self.h, None))
            return
        m = _lib.Mask(plane.data_ptr(), plane.stride(0), plane.element_size(), bit_depth, 0)
        self._mask = plane
        check(self.L.ivqm_bind_mask(self.h, C.byref(m)))

    # -- per-metric calls ---------------------------------------------------------
    def gcd(self):
        g = (C.c_int32 * 4)()
        check(self.L.ivqm_gcd(self.h, g))
        return list(g)

    def ssd(self):
        s = (C.c_uint64 * 4)()
        check(self.L.ivqm_ssd(self.h, s))
        return list(s)

    def _v4(self, fn):
        v = (C.c_double * 4)()
        check(getattr(self.L, fn)(self.h, v))
        return list(v)

    def mse(self):
        return self._v4("ivqm_mse")

    def psnr(self):
        return self._v4("ivqm_psnr")

    def wspsnr(self):
        return self._v4("ivqm_wspsnr")

    def ssim(self):
        return self._v4("ivqm_ssim")

    def msssim(self):
        return self._v4("ivqm_msssim")

    def ivpsnr(self, gcd=None):
        v, a, b = C.c_double(), C.c_double(), C.c_double()
        g = (C.c_int32 * 4)(*gcd) if gcd is not None else None
        check(self.L.ivqm_ivpsnr(self.h, g, C.byref(v), C.byref(a), C.byref(b)))
        return v.value, a.value, b.value

    def gen_scp(self, gcd=None):
        g = (C.c_int32 * 4)(*gcd) if gcd is not None else None
        check(self.L.ivqm_gen_scp(self.h, g))

    def ivssim(self):
        v, a, b = C.c_double(), C.c_double(), C.c_double()
        check(self.L.ivqm_ivssim(self.h, C.byref(v), C.byref(a), C.byref(b)))
        return v.value, a.value, b.value

    def ivmsssim(self):
        v, a, b = C.c_double(), C.c_double(), C.c_double()
        check(self.L.ivqm_ivmsssim(self.h, C.byref(v), C.byref(a), C.byref(b)))
        return v.value, a.value, b.value

    def frame(self, metrics=_lib.IVQM_M_ALL) -> FrameResult:
        r = FrameResult()
        check(self.L.ivqm_frame(self.h, metrics, C.byref(r)))
        return r

    def enqueue(self, metrics, ring):
        check(self.L.ivqm_frame_enqueue(self.h, metrics, ring))

    def collect(self, ring) -> FrameResult:
        r = FrameResult()
        check(self.L.ivqm_frame_collect(self.h, ring, C.byref(r)))
        return r

    def profile(self, kernel, iters=20):
        """Mean duration (ms) of one kernel family over `iters` launches (HIP events, context stream)."""
        ms = C.c_double()
        check(self.L.ivqm_profile(self.h, kernel, iters, C.byref(ms)))
        return ms.value

    # -- debug ------------------------------------------------------------------
    def rows(self, direction):
        out = np.zeros((self.height, 4), np.uint64)
        check(self.L.ivqm_get_rows(self.h, direction, out.ctypes.data_as(C.POINTER(C.c_uint64))))
        return out

    def scp(self, idx):
        planes = [np.zeros((self.height, self.width), np.uint16) for _ in range(3)]
        ptrs = (C.POINTER(C.c_uint16) * 3)(*[p.ctypes.data_as(C.POINTER(C.c_uint16)) for p in planes])
        check(self.L.ivqm_get_scp(self.h, idx, ptrs))
        return planes

    def enable_best_map(self, on=True):
        check(self.L.ivqm_enable_best_map(self.h, int(on)))

    def best_map(self, direction):
        out = np.zeros((self.height, self.width), np.uint8)
        check(self.L.ivqm_get_best_map(self.h, direction, out.ctypes.data_as(C.POINTER(C.c_uint8))))
        return out


def _torch_picture(planes, chroma_420: bool) -> Picture:
    return _picture_from_torch(planes, chroma_420)


class IvqmBatch:
    """Handle over one batch object (include/ivqm_batch.h); pairs are torch CUDA planes."""

    def __init__(self, bit_depth_unused=None, **kw):
        self.L = _lib.load()
        self.cfg = default_config(64, 64, 10, **kw)
        h = C.c_void_p()
        check(self.L.ivqm_batch_create(C.byref(self.cfg), C.byref(h)))
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            self.L.ivqm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def make_pairs(items):
        """items: list of (p0_planes, p1_planes, width, height, bit_depth, chroma_420) with torch CUDA planes."""
        arr = (_lib.Pair * len(items))()
        for i, (p0, p1, w, h, bd, c420) in enumerate(items):
            arr[i].p0 = _picture_from_torch(p0, c420)
            arr[i].p1 = _picture_from_torch(p1, c420)
            arr[i].width, arr[i].height, arr[i].bit_depth = w, h, bd
        return arr

    @staticmethod
    def make_masks(items):
        """items: list of (torch CUDA 2-D plane or None, bit_depth)."""
        arr = (_lib.Mask * len(items))()
        for i, (t, bd) in enumerate(items):
            if t is not None:
                arr[i] = _lib.Mask(t.data_ptr(), t.stride(0), t.element_size(), bd, 0)
        return arr

    def run_masked(self, pairs, masks, metrics):
        res = (FrameResult * len(pairs))()
        check(self.L.ivqm_batch_enqueue_masked(self.h, pairs, masks, len(pairs), metrics, 0))
        check(self.L.ivqm_batch_collect(self.h, 0, res))
        return list(res)

    def set_tuning(self, search_tile_rows=0, ssim_band_rows=0, use_graph=False):
        check(self.L.ivqm_batch_set_tuning(self.h, search_tile_rows, ssim_band_rows, int(use_graph)))

    def tuning(self):
        a, b = C.c_int32(), C.c_int32()
        check(self.L.ivqm_batch_get_tuning(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, pairs, metrics):
        a, b = C.c_int32(), C.c_int32()
        check(self.L.ivqm_batch_autotune(self.h, pairs, len(pairs), metrics, C.byref(a), C.byref(b)))
        return a.value, b.value

    def run(self, pairs, metrics):
        res = (FrameResult * len(pairs))()
        check(self.L.ivqm_batch_run(self.h, pairs, len(pairs), metrics, res))
        return list(res)

    def enqueue(self, pairs, metrics, ring):
        self._keep[ring] = pairs
        check(self.L.ivqm_batch_enqueue(self.h, pairs, len(pairs), metrics, ring))

    def collect(self, ring, n):
        res = (FrameResult * n)()
        check(self.L.ivqm_batch_collect(self.h, ring, res))
        return list(res)

    def profile(self, kernel, iters=10):
        ms = C.c_double()
        check(self.L.ivqm_batch_profile(self.h, kernel, iters, C.byref(ms)))
        return ms.value


def _plane_ptrs(frames):
    """frames: list of 3-plane pictures (numpy arrays, or torch tensors of one device) ->
    (void* array, int32 stride array, on_device flag, objects to keep alive)."""
    n = len(frames)
    ptrs = (C.c_void_p * (3 * n))()
    strides = (C.c_int32 * (3 * n))()
    keep = []
    on_device = False
    for f, planes in enumerate(frames):
        for k in range(3):
            p = planes[k]
            if hasattr(p, "data_ptr"):
                on_device = bool(p.is_cuda)
                ptrs[3 * f + k] = p.data_ptr()
                strides[3 * f + k] = p.stride(0)
            else:
                a = np.ascontiguousarray(p)
                ptrs[3 * f + k] = a.ctypes.data
                strides[3 * f + k] = a.shape[1]
                p = a
            keep.append(p)
    return ptrs, strides, on_device, keep


class IvqmServer:
    """Handle over one serving loop (include/ivqm_serve.h)."""

    def __init__(self, memory_budget=0, max_batch_pairs=0, max_inflight=0, **kw):
        self.L = _lib.load()
        cfg = _lib.ServerConfig()
        check(self.L.ivqm_server_config_init(C.byref(cfg)))
        base = default_config(64, 64, 10, **kw)
        cfg.params = base
        if memory_budget:
            cfg.memory_budget = int(memory_budget)
        if max_batch_pairs:
            cfg.max_batch_pairs = int(max_batch_pairs)
        if max_inflight:
            cfg.max_inflight = int(max_inflight)
        self.cfg = cfg
        h = C.c_void_p()
        check(self.L.ivqm_server_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self._jobs = {}

    def close(self):
        if self.h:
            check(self.L.ivqm_server_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def add_sequence(self, frames, width, height, bit_depth, chroma_420=True):
        """frames: list of 3-plane pictures (host numpy or device torch); copied to the device once."""
        ptrs, strides, _, keep = _plane_ptrs(frames)
        p0 = keep[0]
        bps = p0.element_size() if hasattr(p0, "element_size") else p0.dtype.itemsize
        d = _lib.SequenceDesc(width, height, bit_depth, int(chroma_420), bps, len(frames), ptrs, strides)
        sid = C.c_int32()
        check(self.L.ivqm_server_add_sequence(self.h, C.byref(d), C.byref(sid)))
        return sid.value

    def remove_sequence(self, seq_id):
        check(self.L.ivqm_server_remove_sequence(self.h, seq_id))

    def submit(self, tag, seq_id, first_frame, frames, metrics, masks=None, mask_bit_depth=8):
        """frames: test pictures (host numpy or device torch); kept referenced until polled.
        masks: optional weight planes, one per frame, on the same side as the frames (mask mode)."""
        ptrs, strides, on_dev, keep = _plane_ptrs(frames)
        j = _lib.Job(tag, seq_id, first_frame, len(frames), metrics, int(on_dev), 0, ptrs, strides)
        if masks is not None:
            mp, ms, _, mkeep = _plane_ptrs([[m, m, m] for m in masks])
            n = len(masks)
            mptrs = (C.c_void_p * n)(*[mp[3 * i] for i in range(n)])
            mstr = (C.c_int32 * n)(*[ms[3 * i] for i in range(n)])
            j.mask_planes, j.mask_strides = mptrs, mstr
            j.mask_bit_depth = mask_bit_depth
            m0 = mkeep[0]
            j.mask_bytes_per_sample = m0.element_size() if hasattr(m0, "element_size") else m0.dtype.itemsize
            keep = keep + [mptrs, mstr, mkeep]
        check(self.L.ivqm_server_submit(self.h, C.byref(j)))
        self._jobs[tag] = (ptrs, strides, keep)

    def poll(self, max_results=64, wait_us=0):
        out = (_lib.JobResult * max_results)()
        n = C.c_int32()
        check(self.L.ivqm_server_poll(self.h, out, max_results, C.byref(n), wait_us))
        res = [out[i] for i in range(n.value)]
        for r in res:
            self._jobs.pop(r.tag, None)
        return res

    def drain(self):
        check(self.L.ivqm_server_drain(self.h))

    def stats(self):
        st = _lib.ServerStats()
        check(self.L.ivqm_server_get_stats(self.h, C.byref(st)))
        return st.as_dict()
