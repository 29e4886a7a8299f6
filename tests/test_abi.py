"""CPU checks of the C-ABI boundary: the library loads, exports every symbol the
headers declare, and reports errors instead of crashing when misused."""
import ctypes as C
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ivqm_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_the_abi():
    names = declared_symbols()
    assert "ivqm_create" in names and "ivqm_frame" in names
    from qmiv_amd._lib import EXPORTED
    missing = [n for n in names if n not in EXPORTED]
    assert not missing, f"bindings do not list: {missing}"


def test_library_exports_every_declared_symbol():
    from qmiv_amd._lib import LIB_PATH, load
    if not os.path.exists(LIB_PATH):
        pytest.skip("libivqm_hip.so not built (run python -m qmiv_amd.build)")
    L = load()
    for n in declared_symbols():
        assert hasattr(L, n), n


def test_config_defaults_match_reference():
    """Defaults of the reference's parameter classes (xCorrespPixelShiftPrms.h:20-21,
    xGlobClrDiff.h:20, xSSIM.h:35-37, xWeightedSpherically.h:21)."""
    from qmiv_amd._lib import LIB_PATH
    if not os.path.exists(LIB_PATH):
        pytest.skip("libivqm_hip.so not built")
    from qmiv_amd.context import default_config
    cfg = default_config(64, 64)
    assert cfg.search_range == 2
    assert list(cfg.cmp_weights_search) == [4, 1, 1, 0]
    assert list(cfg.cmp_weights_average) == [4, 1, 1, 0]
    assert [round(x, 6) for x in cfg.unnoticeable_coef] == [0.01, 0.01, 0.01, 0.0]
    assert (cfg.ssim_window, cfg.ssim_stride) == (8, 4)
    assert (cfg.ssim_mode, cfg.ssim_border) == (4, 0)  # BlockAveraged, no extension (xSSIM.h:35, xAppQMIV.cpp:327)
    assert (cfg.lon_range_deg, cfg.lat_range_deg) == (360, 180)


@pytest.mark.parametrize("field,value,ok", [
    ("ssim_mode", 5, False), ("ssim_mode", -1, False), ("ssim_border", 6, False),
    ("ssim_stride", 0, False), ("ssim_stride", 9, False),
    (("ssim_mode", "ssim_window"), (3, 32), False), (("ssim_mode", "ssim_window"), (3, 16), True),
    (("ssim_mode", "ssim_stride"), (0, 11), True), (("ssim_mode", "ssim_stride"), (1, 12), False),
    (("ssim_mode", "ssim_window", "ssim_stride"), (2, 8, 11), True),  # regular modes ignore ssim_window
    (("ssim_window", "ssim_stride"), (65, 4), False), (("ssim_window", "ssim_stride"), (32, 32), True),
])
def test_ssim_config_validation(field, value, ok):
    """SSIM mode / window / stride / border rules (xAppQMIV.cpp:325-332, xStructSimSTD.cpp:353-419):
    rejected settings return IVQM_EINVAL before any device work; accepted ones get past
    validation (to the HIP error of a GPU-less host, or a context on a GPU host)."""
    from qmiv_amd._lib import LIB_PATH, Config, load
    if not os.path.exists(LIB_PATH):
        pytest.skip("libivqm_hip.so not built")
    L = load()
    cfg = Config()
    L.ivqm_config_init(C.byref(cfg))
    cfg.width, cfg.height = 64, 64
    fields = field if isinstance(field, tuple) else (field,)
    values = value if isinstance(value, tuple) else (value,)
    for f, v in zip(fields, values):
        setattr(cfg, f, v)
    h = C.c_void_p()
    rc = L.ivqm_create(C.byref(cfg), C.byref(h))
    if ok:
        assert rc != -22, L.ivqm_last_error()
        if rc == 0:
            L.ivqm_destroy(h)
    else:
        assert rc == -22


def test_invalid_arguments_return_status_not_crash():
    from qmiv_amd._lib import LIB_PATH, Config, load
    if not os.path.exists(LIB_PATH):
        pytest.skip("libivqm_hip.so not built")
    L = load()
    assert L.ivqm_config_init(None) == -22
    assert L.ivqm_create(None, None) == -22
    cfg = Config()
    L.ivqm_config_init(C.byref(cfg))
    cfg.width, cfg.height, cfg.bit_depth = 64, 64, 20
    h = C.c_void_p()
    assert L.ivqm_create(C.byref(cfg), C.byref(h)) == -22
    assert L.ivqm_strerror(-22) == b"invalid argument"
    assert L.ivqm_frame(None, 0, None) == -22
    # no GPU in this container: creating a valid context reports a HIP error, it does not abort
    cfg.bit_depth = 10
    rc = L.ivqm_create(C.byref(cfg), C.byref(h))
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        assert rc == 0
        L.ivqm_destroy(h)
    else:
        assert rc <= -1000
        assert len(L.ivqm_last_error()) > 0


def test_noalloc_rejects_requests_beyond_its_limits():
    """ivqm_batch_run_noalloc checks its limits before any HIP call: too many pairs, heights adding up
    past IVQM_NOALLOC_MAX_ROWS, metrics outside the PSNR family and empty batches return IVQM_EINVAL.
    (The plane addresses are placeholders that nothing reads.)"""
    from qmiv_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libivqm_hip.so not built")
    from qmiv_amd.context import default_config
    L = _lib.load()
    cfg = default_config(64, 64)
    fam = _lib.IVQM_M_PSNR | _lib.IVQM_M_MSE | _lib.IVQM_M_WSPSNR | _lib.IVQM_M_IVPSNR

    def pairs(n, w, h):
        arr = (_lib.Pair * n)()
        for p in arr:
            for pic in (p.p0, p.p1):
                for k in range(3):
                    pic.plane[k] = 0x10000 * (k + 1)
                    pic.stride[k] = w if k == 0 else w // 2
                pic.chroma_420, pic.bytes_per_sample = 1, 2
            p.width, p.height, p.bit_depth = w, h, 10
        return arr

    res = (_lib.FrameResult * 32)()
    n = _lib.NOALLOC_MAX_PAIRS + 1
    assert L.ivqm_batch_run_noalloc(C.byref(cfg), pairs(n, 64, 64), n, fam, res, None) == -22
    tall = _lib.NOALLOC_MAX_ROWS // 4 + 2  # four such heights pass the row limit
    assert L.ivqm_batch_run_noalloc(C.byref(cfg), pairs(4, 64, tall), 4, fam, res, None) == -22
    assert b"IVQM_NOALLOC_MAX_ROWS" in L.ivqm_last_error()
    for m in (_lib.IVQM_M_IVSSIM, _lib.IVQM_M_SSIM | _lib.IVQM_M_PSNR, 0):
        assert L.ivqm_batch_run_noalloc(C.byref(cfg), pairs(1, 64, 64), 1, m, res, None) == -22
    assert L.ivqm_batch_run_noalloc(C.byref(cfg), pairs(1, 64, 64), 0, fam, res, None) == -22
