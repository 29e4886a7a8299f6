"""Colour-space conversion on the GPU (ivqm_convert_colour, k_colour) against the oracle, bit for bit:
fixed-point YCbCr <-> RGB of the reference (xColorSpace.h dispatch of the _I32 kernels,
libs/pmbb-core/src/xColorSpaceSTD.cpp) for the four colour spaces, 8/10/16-bit, 4:2:0 and 4:4:4
sources; and the command-line front end with -csi / -csm (xAppQMIV.cpp:853-892)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import qmiv_amd as Q
from qmiv_amd import _lib, synth
from tests.imagefiles import bmp_bytes, png_bytes

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "qmiv_amd", "bin", "ivqm_qmiv")


@pytest.mark.parametrize("bd", [8, 10, 16])
@pytest.mark.parametrize("c420", [True, False])
def test_convert_colour_matches_oracle(oracle, bd, c420):
    torch = pytest.importorskip("torch")
    dev = torch.device("cuda:0")
    wl = synth.Workload(cfg=200 + bd, width=130, height=66, bit_depth=bd, chroma_420=c420,
                        sigma_r=(1 << bd) / 6, sigma_t=(1 << bd) / 10)
    p0n, _ = synth.make_pair(wl, 0)
    W, H = wl.width, wl.height
    L = _lib.load()
    src = [torch.from_numpy(p.astype(np.uint16).view(np.int16)).to(dev) for p in p0n]
    pic = Q.context._picture_from_torch(src, c420)
    full = oracle.to444(p0n, c420)
    for direction in (0, 1):
        if direction == 1 and c420:
            continue  # RGB sources are 4:4:4 in practice (the kernel accepts 4:2:0 as upsampled planes)
        for space in range(4):
            out = torch.zeros((3, H, W), dtype=torch.int16, device=dev)
            _lib.check(L.ivqm_convert_colour(C.byref(pic), W, H, bd, direction, space, out.data_ptr(), W, None))
            torch.cuda.synchronize()
            got = out.cpu().numpy().view(np.uint16)
            want = oracle.convert_colour(full, bd, direction, space)
            for k in range(3):
                assert np.array_equal(got[k], want[k]), (bd, c420, direction, space, k)


def _lines_of(out):
    return out.stdout.splitlines()


@pytest.mark.parametrize("csi,csm,c420", [("YCbCr_BT709", "RGB", True), ("BGR", "RGB", False),
                                          ("RGB", "YCbCr_BT2020", False)])
def test_front_end_colour_spaces(tmp_path, oracle, csi, csm, c420):
    """PSNR / IV-PSNR / SSIM lines computed on the converted (or reordered) pictures, with the R:G:B
    suffixes when the metric colour space is RGB (xMetricQMIV.cpp:61-82)."""
    if not os.path.exists(EXE):
        pytest.fail("front end not built (python -m qmiv_amd.build)")
    wl = synth.Workload(cfg=413, width=128, height=96, chroma_420=c420, global_shift=1)
    bd, n = 10, 2
    pairs = [synth.make_pair(wl, i) for i in range(n)]
    f0, f1 = tmp_path / "a.yuv", tmp_path / "b.yuv"
    raw = lambda pic: b"".join(np.ascontiguousarray(p, np.uint16).tobytes() for p in pic)
    f0.write_bytes(b"".join(raw(p[0]) for p in pairs))
    f1.write_bytes(b"".join(raw(p[1]) for p in pairs))
    out = subprocess.run([EXE, "-i0", str(f0), "-i1", str(f1), "-ps", f"{wl.width}x{wl.height}", "-bd", str(bd),
                          "-cf", "420" if c420 else "444", "-ml", "PSNR, IVPSNR, SSIM", "-csi", csi, "-csm", csm,
                          "-v", "2"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = _lines_of(out)
    rgb = csm == "RGB"
    sfx_cmp, sfx_pic = (" R:G:B    ", "-RGB      ") if rgb else (" Y:Cb:Cr  ", "-YCbCr    ")
    spaces = {"YCbCr_BT709": 1, "YCbCr_BT2020": 3}
    for f, (p0n, p1n) in enumerate(pairs):
        pics = []
        for pn in (p0n, p1n):
            p = oracle.to444(pn, c420)
            if csi.startswith("YCbCr"):
                p = oracle.convert_colour(p, bd, 0, spaces[csi])
            elif csm.startswith("YCbCr"):
                p = oracle.convert_colour(p, bd, 1, spaces[csm])
            elif csi == "BGR":
                p = [p[2], p[1], p[0]]
            pics.append(p)
        p0, p1 = pics
        exact = [bool(np.array_equal(p0[c], p1[c])) for c in range(3)]
        psnr = oracle.psnr(p0, p1, bd)
        fake = oracle.psnr_from_ssd(1, wl.width * wl.height, bd)
        psnr = [fake if exact[c] else v for c, v in enumerate(psnr)]
        line = "Frame %08d " % f + "%8s%s " % ("PSNR", sfx_cmp) + "".join("%8.4f " % v for v in psnr)
        line += "".join(" Exact" + "YUV"[c] for c in range(3) if exact[c])
        assert line in lines, (line, [l for l in lines if "PSNR" in l][:4])
        g = oracle.gcd(p0, p1, bd)
        iv = oracle.ivpsnr(p0, p1, bd, g, R=2)[0]
        sfx_iv = "-RGB      " if rgb else " " * 10
        assert ("Frame %08d " % f + "%8s%s " % ("IVPSNR", sfx_iv) + "%8.4f " % iv + " " * 18) in lines
        ss = oracle.ssim(p0, p1, bd)
        sl = next(l for l in lines if l.startswith("Frame %08d " % f + "%8s%s " % ("SSIM", sfx_cmp)))
        got = [float(x) for x in sl.split()[-3:]]
        assert np.max(np.abs(np.array(got) - np.array(ss))) <= 1.5e-6
        assert any(l.startswith("Frame %08d " % f + "%8s%s " % ("SSIM", sfx_pic)) for l in lines)


@pytest.mark.parametrize("fmt", ["PNG", "PNG-Adam7", "BMP"])
def test_front_end_image_lists(tmp_path, oracle, fmt):
    """-ff PNG / BMP: frames decoded from image files named by a pattern (first index 1; PNG rows use
    every filter type and the data spans two IDAT chunks, plain and Adam7-interlaced; BMP top-down and
    bottom-up) give the PSNR / IV-PSNR of the RGB planes, with R:G:B suffixes (xSeqLST.cpp,
    xAppQMIV.cpp:297)."""
    if not os.path.exists(EXE):
        pytest.fail("front end not built (python -m qmiv_amd.build)")
    wl = synth.Workload(cfg=414, width=90, height=60, bit_depth=8, chroma_420=False, global_shift=1)
    n = 2
    pairs = [synth.make_pair(wl, i) for i in range(n)]
    adam7 = fmt == "PNG-Adam7"
    fmt = fmt.split("-")[0]
    ext = fmt.lower()
    for f, (p0n, p1n) in enumerate(pairs):
        for tag, pic in (("a", p0n), ("b", p1n)):
            rgb = np.stack([np.asarray(p, np.uint8) for p in pic], axis=-1)
            data = (png_bytes(rgb, 2, [0, 1, 2, 3, 4], interlace=adam7) if fmt == "PNG"
                    else bmp_bytes(rgb, top_down=(f % 2 == 1)))
            (tmp_path / f"{tag}_{f + 1:03d}.{ext}").write_bytes(data)
    out = subprocess.run([EXE, "-i0", str(tmp_path / f"a_{{:03d}}.{ext}"), "-i1", str(tmp_path / f"b_{{:03d}}.{ext}"),
                          "-ps", f"{wl.width}x{wl.height}", "-ff", fmt, "-ml", "PSNR, IVPSNR", "-v", "2"],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    for f, (p0n, p1n) in enumerate(pairs):
        p0 = [np.asarray(p, np.uint16) for p in p0n]
        p1 = [np.asarray(p, np.uint16) for p in p1n]
        psnr = oracle.psnr(p0, p1, 8)
        line = "Frame %08d " % f + "%8s%s " % ("PSNR", " R:G:B    ") + "".join("%8.4f " % v for v in psnr)
        assert line in lines, (line, [l for l in lines if "PSNR" in l][:3])
        g = oracle.gcd(p0, p1, 8)
        iv = oracle.ivpsnr(p0, p1, 8, g, R=2)[0]
        assert ("Frame %08d " % f + "%8s%s " % ("IVPSNR", "-RGB      ") + "%8.4f " % iv + " " * 18) in lines
