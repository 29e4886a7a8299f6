"""The front end's PNG / BMP readers (qmiv_amd/csrc/ivqm_images.cpp, host only,
no GPU): files written by tests/imagefiles.py from the PNG specification decode
to the RGB planes they were written from, for every 8-bit colour type, every
filter type, plain and Adam7-interlaced, at sizes smaller than one Adam7 block
and ragged ones. The reference decodes any 8-bit PNG to RGB8 through spng
(libs/pmbb-slst/src/xSeqLST.cpp:186-238) and reads BI_RGB BMPs (:337-415)."""
import os
import subprocess

import numpy as np
import pytest

from tests.imagefiles import bmp_bytes, png_bytes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "qmiv_amd", "bin", "ivqm_image_dump")


@pytest.fixture(scope="module")
def exe():
    if not os.path.exists(EXE):
        from qmiv_amd import build
        build.build_image_dump()
    return EXE


def _decode(exe, path, fmt, w, h):
    out = subprocess.run([exe, fmt, str(path), str(w), str(h)], capture_output=True, timeout=60)
    return out.returncode, out.stdout, out.stderr.decode()


def _expect_rgb(samples, ctype, palette):
    if ctype in (0, 4):
        g = samples[:, :, 0]
        return np.stack([g, g, g])
    if ctype == 3:
        pal = np.asarray(palette, np.uint8)
        return np.moveaxis(pal[samples[:, :, 0]], -1, 0)
    return np.moveaxis(samples[:, :, :3], -1, 0)


@pytest.mark.parametrize("interlace", [False, True])
@pytest.mark.parametrize("ctype", [0, 2, 3, 4, 6])
@pytest.mark.parametrize("size", [(1, 1), (3, 2), (7, 5), (37, 23), (64, 16)])
def test_png_decodes_to_its_planes(tmp_path, exe, interlace, ctype, size):
    w, h = size
    rng = np.random.default_rng(1000 * w + 10 * h + ctype + (500 if interlace else 0))
    palette = None
    if ctype == 3:
        palette = [tuple(int(v) for v in rng.integers(0, 256, 3)) for _ in range(37)]
        samples = rng.integers(0, len(palette), (h, w, 1), dtype=np.uint8)
    else:
        samples = rng.integers(0, 256, (h, w, {0: 1, 2: 3, 4: 2, 6: 4}[ctype]), dtype=np.uint8)
    path = tmp_path / "f.png"
    path.write_bytes(png_bytes(samples, ctype, filters=(0, 1, 2, 3, 4, 4, 3), interlace=interlace, palette=palette,
                               idat_split=7))
    rc, out, err = _decode(exe, path, "PNG", w, h)
    assert rc == 0, err
    got = np.frombuffer(out, np.uint8).reshape(3, h, w)
    np.testing.assert_array_equal(got, _expect_rgb(samples, ctype, palette))


def test_png_adam7_equals_plain(tmp_path, exe):
    """An Adam7 file and a plain file of the same image decode to the same planes."""
    rng = np.random.default_rng(7)
    img = rng.integers(0, 256, (45, 71, 3), dtype=np.uint8)
    (tmp_path / "a.png").write_bytes(png_bytes(img, 2, interlace=False))
    (tmp_path / "b.png").write_bytes(png_bytes(img, 2, interlace=True))
    ra, oa, _ = _decode(exe, tmp_path / "a.png", "PNG", 71, 45)
    rb, ob, _ = _decode(exe, tmp_path / "b.png", "PNG", 71, 45)
    assert ra == rb == 0 and oa == ob


def test_png_rejections(tmp_path, exe):
    img = np.zeros((8, 8, 3), np.uint8)
    p = tmp_path / "f.png"
    p.write_bytes(png_bytes(img, 2))
    assert _decode(exe, p, "PNG", 9, 8)[0] == 1
    data = bytearray(png_bytes(img, 2, interlace=True))
    data[8 + 8 + 12] = 2  # interlace method 2 (CRC is not checked, the method is)
    p.write_bytes(bytes(data))
    rc, _, err = _decode(exe, p, "PNG", 8, 8)
    assert rc == 1 and "interlace" in err
    p.write_bytes(png_bytes(img, 2)[:-30])
    assert _decode(exe, p, "PNG", 8, 8)[0] == 1
    p.write_bytes(b"not a png at all")
    rc, _, err = _decode(exe, p, "PNG", 8, 8)
    assert rc == 1 and "not a PNG" in err


@pytest.mark.parametrize("top_down", [False, True])
def test_bmp_decodes_to_its_planes(tmp_path, exe, top_down):
    rng = np.random.default_rng(3 + top_down)
    img = rng.integers(0, 256, (13, 11, 3), dtype=np.uint8)
    p = tmp_path / "f.bmp"
    p.write_bytes(bmp_bytes(img, top_down))
    rc, out, err = _decode(exe, p, "BMP", 11, 13)
    assert rc == 0, err
    np.testing.assert_array_equal(np.frombuffer(out, np.uint8).reshape(3, 13, 11), np.moveaxis(img, -1, 0))
