"""Seeded synthetic frame pairs for the QMIV metric workloads.

Thin ctypes wrapper over ``libivqm_synth.so`` (qmiv_amd/csrc/synth.cpp, where
the generator specification is written out). Returns numpy planes:
``(p0, p1)`` with ``p0`` the test picture (QMIV InputFile0) and ``p1`` the
reference picture (InputFile1); each is a list of three 2-D arrays at their
native resolution (chroma halved in both axes for 4:2:0).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SYNTH_SO = os.path.join(HERE, "libivqm_synth.so")


class _Spec(C.Structure):
    _fields_ = [
        ("cfg", C.c_int32), ("width", C.c_int32), ("height", C.c_int32), ("bit_depth", C.c_int32),
        ("chroma_420", C.c_int32), ("sigma_r", C.c_double), ("sigma_t", C.c_double),
        ("block_shift", C.c_int32), ("global_shift", C.c_int32), ("offsets", C.c_int32 * 3),
        ("offset_rand", C.c_int32),
    ]


@dataclasses.dataclass(frozen=True)
class Workload:
    """One synthetic frame-pair stream."""
    cfg: int
    width: int
    height: int
    bit_depth: int = 10
    chroma_420: bool = True
    sigma_r: float = 4.0
    sigma_t: float = 3.0
    block_shift: int = 1
    global_shift: int = 0
    offsets: tuple = (2, -1, 1)
    offset_rand: int = 0


# The five workloads: 1080p CPU case, 4096x2048 ERP SSIM family, the batched
# mixed-size set (sizes and depths listed in BATCH_SIZES), the served stream
# (drawn from BATCH_SIZES) and the 8192x4096 16-bit stress case.
CPU_1080P = Workload(cfg=0, width=1920, height=1080)
ERP_4K2K = Workload(cfg=1, width=4096, height=2048, global_shift=2)
STRESS_8K_444 = Workload(cfg=4, width=8192, height=4096, bit_depth=16, chroma_420=False, sigma_r=12.0,
                         sigma_t=12.0, block_shift=3, offset_rand=8)
STRESS_8K_420 = dataclasses.replace(STRESS_8K_444, chroma_420=True)
BATCH_SIZES = [(1280, 720), (1920, 1080), (2048, 2048), (4096, 2048)]


def batch_workload(index: int) -> Workload:
    """Frame pair `index` (0..63) of the batched set: 16 of each size, 8/10-bit alternating."""
    w, h = BATCH_SIZES[index // 16]
    return Workload(cfg=2, width=w, height=h, bit_depth=8 if index % 2 == 0 else 10)


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(SYNTH_SO):
            raise RuntimeError(f"{SYNTH_SO} is missing: run qmiv_amd.build.build() first")
        _lib = C.CDLL(SYNTH_SO)
        _lib.ivqm_synth_pair.restype = C.c_int
    return _lib


def plane_shapes(wl: Workload):
    cs = 1 if wl.chroma_420 else 0
    return [(wl.height, wl.width), (wl.height >> cs, wl.width >> cs), (wl.height >> cs, wl.width >> cs)]


def make_pair(wl: Workload, frame: int, as_u8: bool = False):
    """Generates frame pair `frame` of `wl`; returns (p0, p1) lists of planes."""
    L = _load()
    sp = _Spec(wl.cfg, wl.width, wl.height, wl.bit_depth, int(wl.chroma_420), wl.sigma_r, wl.sigma_t,
               wl.block_shift, wl.global_shift, (C.c_int32 * 3)(*wl.offsets), wl.offset_rand)
    dt = np.uint8 if as_u8 else np.uint16
    p0 = [np.empty(s, dt) for s in plane_shapes(wl)]
    p1 = [np.empty(s, dt) for s in plane_shapes(wl)]
    if as_u8:
        ptr = C.POINTER(C.c_uint8)
        a = (ptr * 3)(*[x.ctypes.data_as(ptr) for x in p0])
        b = (ptr * 3)(*[x.ctypes.data_as(ptr) for x in p1])
        rc = L.ivqm_synth_pair(C.byref(sp), frame, None, None, a, b)
    else:
        ptr = C.POINTER(C.c_uint16)
        a = (ptr * 3)(*[x.ctypes.data_as(ptr) for x in p0])
        b = (ptr * 3)(*[x.ctypes.data_as(ptr) for x in p1])
        rc = L.ivqm_synth_pair(C.byref(sp), frame, a, b, None, None)
    if rc != 0:
        raise ValueError(f"ivqm_synth_pair failed ({rc}) for {wl}")
    return p0, p1


def make_mask(width: int, height: int, bit_depth_m: int, seed: int = 0, zero_per_mille: int = 350) -> np.ndarray:
    """Deterministic weight mask (one uint16 plane, 0 .. 2^bit_depth_m - 1) for mask-mode tests and
    benches: hashed weights, about zero_per_mille / 1000 zeros, a fully masked band of rows and a
    band of maximum weights."""
    y, x = np.mgrid[0:height, 0:width].astype(np.uint64)
    h = x * np.uint64(2654435761) + y * np.uint64(40503) + np.uint64(seed * 97531 + 1)
    h ^= h >> np.uint64(13)
    h *= np.uint64(0x5BD1E995)
    h ^= h >> np.uint64(15)
    maxm = (1 << bit_depth_m) - 1
    m = (h % np.uint64(maxm + 1)).astype(np.uint16)
    m[((h >> np.uint64(24)) % np.uint64(1000)).astype(np.int64) < zero_per_mille] = 0
    if height > 24:
        m[height // 3:height // 3 + 6, :] = 0
    if width > 24:
        m[:, width // 5:width // 5 + 4] = maxm
    return m
