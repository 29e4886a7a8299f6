"""Host-side bounds rule of the batch engine (CPU, no GPU needed).

Every picture a batch's tables name inside the engine's SCP or pyramid arena
is turned into device pointers by ivqm::resolve_picture
(qmiv_amd/csrc/ivqm_engine.h) before any launch reads the tables. A picture
that would reach past its arena is rejected with IVQM_EINVAL. The program
tests/plan_check/arena_check.cpp (built by qmiv_amd/build.py) drives that
rule to each of its rejections and checks that accepted pictures lie inside
their arena."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "qmiv_amd", "bin", "ivqm_arena_check")


def test_arena_bounds_rule_rejects_pictures_outside_their_arena():
    if not os.path.exists(EXE):
        from qmiv_amd import build
        build.build_arena_check()
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    print(out.stdout)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.splitlines()
    assert lines[-1] == "OK"
    assert sum("rc=-22" in ln for ln in lines) == 9
