"""Builds the native libraries of qmiv_amd in-tree.

  qmiv_amd/libivqm_hip.so    gfx950 kernels + C ABI (hipcc)
  qmiv_amd/bin/ivqm_qmiv     QMIV command-line front end over the batch ABI
  qmiv_amd/libivqm_synth.so  seeded synthetic frame generator (g++)

Run ``python -m qmiv_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["ivqm_kernels.hip", "ivqm_ssim_modes.hip", "ivqm_masked.hip", "ivqm_engine.hip", "ivqm_runtime.hip",
               "ivqm_serve.hip", "ivqm_noalloc.hip", "ivqm_search420.hip"]
OBJ_DIR = os.path.join(HERE, "_build")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 VALU forms stay off (hipcc -O3 would otherwise emit v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _run(cmd, **kw):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, **kw)


def build_hip(verbose_isa: bool = False) -> str:
    """Compiles every HIP source to an object (in parallel), then links the shared library."""
    out = os.path.join(HERE, "libivqm_hip.so")
    os.makedirs(OBJ_DIR, exist_ok=True)
    procs, objs = [], []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
        cmd = [HIPCC, *HIP_FLAGS, "-I" + os.path.join(ROOT, "include"), "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, "hipcc " + " ".join(failed))
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-soname,libivqm_hip.so", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_synth() -> str:
    out = os.path.join(HERE, "libivqm_synth.so")
    _run(["g++", "-std=c++17", "-O2", "-fopenmp", "-fPIC", "-shared", os.path.join(CSRC, "synth.cpp"), "-o", out])
    return out


def build_dropin_check(lib: str) -> str:
    """C++ program that drives include/ivqm_dropin.hpp (run by tests/test_gpu_dropin.py)."""
    out_dir = os.path.join(HERE, "bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "ivqm_dropin_check")
    _run(["g++", "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "dropin", "dropin_check.cpp"), "-L" + os.path.dirname(lib), "-livqm_hip",
          "-Wl,-rpath,$ORIGIN/..", "-o", out])
    return out


def build_arena_check() -> str:
    """Host-only program over the batch engine's arena bounds rule (run by tests/test_plan_check.py)."""
    out_dir = os.path.join(HERE, "bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "ivqm_arena_check")
    _run([HIPCC, "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
          os.path.join(ROOT, "tests", "plan_check", "arena_check.cpp"), "-o", out])
    return out


def build_image_dump() -> str:
    """Host-only program over the front end's PNG / BMP readers (run by tests/test_images.py)."""
    out_dir = os.path.join(HERE, "bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "ivqm_image_dump")
    _run(["g++", "-std=c++17", "-O2", os.path.join(ROOT, "tests", "images_check", "image_dump.cpp"),
          os.path.join(CSRC, "ivqm_images.cpp"), "-lz", "-o", out])
    return out


def build_app(lib: str) -> str:
    """ivqm_qmiv: the QMIV command-line front end (raw YUV files -> per-frame and average metrics)."""
    out_dir = os.path.join(HERE, "bin")
    os.makedirs(out_dir, exist_ok=True)
    out = os.path.join(out_dir, "ivqm_qmiv")
    _run([HIPCC, "-std=c++17", "-O2", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, "ivqm_app.cpp"), os.path.join(CSRC, "ivqm_images.cpp"),
          "-L" + os.path.dirname(lib), "-livqm_hip", "-lz", "-lpthread", "-Wl,-rpath,$ORIGIN/..", "-o", out])
    return out


def build_oracle() -> None:
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j8"])


def build(oracle: bool = True) -> None:
    build_synth()
    lib = build_hip()
    build_dropin_check(lib)
    build_arena_check()
    build_image_dump()
    build_app(lib)
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build(oracle="--no-oracle" not in sys.argv)
