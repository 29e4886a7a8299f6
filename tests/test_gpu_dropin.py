"""The C++ drop-in classes (include/ivqm_dropin.hpp), driven like the QMIV frame loop,
against the oracle: GCD and IV-PSNR / PSNR / WS-PSNR bit-exact, SSIM family within 1e-6."""
import json
import os
import subprocess

import numpy as np
import pytest

from qmiv_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "qmiv_amd", "bin", "ivqm_dropin_check")


@pytest.mark.parametrize("erp", [0, 1])
def test_dropin_classes_match_oracle(tmp_path, oracle, erp):
    if not os.path.exists(EXE):
        pytest.fail("drop-in check program not built (python -m qmiv_amd.build)")
    wl = synth.Workload(cfg=400, width=544, height=528, global_shift=1)
    p0n, p1n = synth.make_pair(wl, 0)
    p0, p1 = oracle.to444(p0n, True), oracle.to444(p1n, True)
    f0, f1 = tmp_path / "p0.raw", tmp_path / "p1.raw"
    f0.write_bytes(b"".join(np.ascontiguousarray(p).tobytes() for p in p0))
    f1.write_bytes(b"".join(np.ascontiguousarray(p).tobytes() for p in p1))
    out = subprocess.run([EXE, str(wl.width), str(wl.height), "10", str(f0), str(f1), str(erp)],
                         capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    r = json.loads(out.stdout)
    g = oracle.gcd(p0, p1, 10)
    assert r["gcd"] == g
    assert r["psnr"] == oracle.psnr(p0, p1, 10)
    assert r["wspsnr"] == oracle.wspsnr(p0, p1, 10, erp=bool(erp))
    assert [r["ivpsnr"], r["r2t"], r["t2r"]] == list(oracle.ivpsnr(p0, p1, 10, g, erp=bool(erp)))
    s0, s1 = oracle.scp(p0, p1, 10, g)
    assert r["scp_sum"] == sum(int(s0[c].sum(dtype=np.uint64)) * 3 + int(s1[c].sum(dtype=np.uint64)) for c in range(3))
    assert np.max(np.abs(np.array(r["ssim"]) - oracle.ssim(p0, p1, 10))) <= 1e-6
    assert np.max(np.abs(np.array(r["msssim"]) - oracle.msssim(p0, p1, 10))) <= 1e-6
    assert abs(r["ivssim"] - oracle.ivssim(p0, p1, s0, s1, 10, erp=bool(erp))[0]) <= 1e-6
    assert abs(r["ivmsssim"] - oracle.ivmsssim(p0, p1, s0, s1, 10)) <= 1e-6
