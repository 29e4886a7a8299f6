"""Option handling of the QMIV command-line front end (qmiv_amd/bin/ivqm_qmiv):
rejected configurations exit with status 1 and the reference's style of
message before any GPU call (apps/QMIV/src/xAppQMIV.cpp:217-370)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "qmiv_amd", "bin", "ivqm_qmiv")


def _run(*args):
    if not os.path.exists(EXE):
        pytest.skip("front end not built (python -m qmiv_amd.build)")
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


def test_help():
    r = _run("-h")
    assert r.returncode == 0 and "-i0" in r.stdout


@pytest.mark.parametrize("args,msg", [
    (["-i0", "a", "-ps", "64x64"], "InputFile1 is empty"),
    (["-i0", "a", "-i1", "b"], "PictureSize"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-cf", "411"], "ChromaFormat"),
    (["-i0", "a", "-i1", "b", "-ps", "65x64", "-cf", "422"], "4:2:2 needs an even picture width"),
    (["-i0", "a", "-i1", "b", "-ps", "65x64", "-cf", "420"], "even picture sizes"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-bd", "17"], "BitDepth"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ml", "PSNR, FOO"], 'Token "FOO"'),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ssm", "GaussianFlt"], "StructSimMode value is not valid"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ssm", "RegularGaussianFlt", "-ssw", "8"],
     "only StructSimWindow==11"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ssb", "Wrap"], "StructSimBrdExt value is not valid"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-sss", "9"], "StructSimStride must be in range"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ssm", "RegularAveraged", "-sss", "12"], "StructSimStride"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ssm", "BlockGaussianInt", "-ssw", "32"], "BlockGaussianInt"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-sr", "5"], "SearchRange"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-im", "m.yuv"], "cannot be combined with Mask mode"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-im", "m.yuv", "-ml", "PSNR", "-bdm", "16"], "BitDepthM"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-im", "m.yuv", "-ml", "PSNR", "-cfm", "411"], "ChromaFormatM"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ff", "TIFF"], "FileFormat is invalid"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ff", "PNG", "-bd", "10"], "contains 8-bit per pixel data"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-ff", "BMP", "-csi", "YCbCr_BT709"], "contains data in RGB color space"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-csi", "XYZ"], "ColorSpaceInput value is not valid"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-csi", "YCbCr_BT709", "-csm", "YCbCr_BT601"],
     "YCbCr to YCbCr conversion is not supported"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-csi", "YCbCr", "-csm", "RGB"], "Generic YCbCr cannot be used"),
    (["-i0", "a", "-i1", "b", "-ps", "64x64", "-cws", "4:1:1"], "component weights"),
    (["-i0", "a", "-i1", "b", "-pf", "rgb24", "-ps", "64x64"], "PictureFormat"),
])
def test_rejected_options(args, msg):
    r = _run(*args)
    assert r.returncode == 1 and msg in r.stderr, r.stderr


@pytest.mark.parametrize("args", [
    ["-ssm", "regulargaussianint", "-ssb", "Mirror", "-sss", "1"],
    ["-ssm", "BlockGaussianInt", "-ssw", "16", "-sss", "16"],
    ["-ssm", "BlockAveraged", "-ssw", "32", "-sss", "8"],
    ["-ssm", "RegularGaussianFlt", "-ssb", " Constant "],
])
def test_accepted_ssim_options(tmp_path, args):
    # accepted SSIM settings get past option checking (an empty file then has no frames)
    a = tmp_path / "a.yuv"
    a.write_bytes(b"")
    r = _run("-i0", str(a), "-i1", str(a), "-ps", "64x64", *args)
    assert r.returncode == 1 and "no frames to process" in r.stderr, r.stderr


def test_mask_file_checks(tmp_path):
    """The mask sequence must exist and hold at least the frames to process (xAppQMIV.cpp:627)."""
    a = tmp_path / "a.yuv"
    a.write_bytes(bytes(64 * 64 * 3 // 2) * 3)  # three 8-bit 4:2:0 frames
    base = ["-i0", str(a), "-i1", str(a), "-ps", "64x64", "-ml", "PSNR, IVPSNR"]
    r = _run(*base, "-im", str(tmp_path / "none.yuv"))
    assert r.returncode == 1 and "InputFileM does not exist" in r.stderr
    m = tmp_path / "m.yuv"
    m.write_bytes(bytes(64 * 64) * 2)  # two luma-only frames
    r = _run(*base, "-im", str(m), "-cfm", "400")
    assert r.returncode == 1 and "FramesToProcess > NumOfFramesM" in r.stderr


def test_image_list_not_found(tmp_path):
    r = _run("-i0", str(tmp_path / "a_{:03d}.png"), "-i1", str(tmp_path / "b_{}.png"), "-ps", "64x64", "-ff", "PNG",
             "-ml", "PSNR")
    assert r.returncode == 1 and "First Idx=(0 or 1) not found" in r.stderr


def test_missing_input_file(tmp_path):
    r = _run("-i0", str(tmp_path / "none0.yuv"), "-i1", str(tmp_path / "none1.yuv"), "-ps", "64x64")
    assert r.returncode == 1 and "does not exist" in r.stderr


def test_pix_fmt_sets_depth_and_chroma(tmp_path):
    # yuv444p12le: 12-bit 4:4:4; an empty file has no frames, reported after the options are accepted
    a = tmp_path / "a.yuv"
    a.write_bytes(b"")
    r = _run("-i0", str(a), "-i1", str(a), "-pf", "yuv444p12le", "-ps", "64x64")
    assert r.returncode == 1 and "no frames to process" in r.stderr
