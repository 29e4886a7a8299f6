#!/usr/bin/env python3
"""Throughput bench of the QMIV metric hot path on one MI355X.

Workload: 17 frame pairs of 4096x2048 equirectangular, 10-bit 4:2:0 (the
seeded synthetic generator of qmiv_amd/csrc/synth.cpp), resident in HBM before
timing. One step = one pass over all 17 pairs computing IV-PSNR and IV-SSIM
(4:2:0 reads with on-the-fly replication, GCD, both search directions with the
shift-compensated pictures, both IV-SSIM directions, FP64 totals, result
download and the host FP64 epilogue). value = 17 * 4096 * 2048 px-pairs / s / 1e6.

Also reported: the roofline of the dominant kernel (the search: VALU-bound) and
of the streaming SSIM kernel (HBM-bound), both timed with HIP events on the
context's stream; and the reference library's own CPU path (oracle/_ref,
compiled from the reference sources) timed on a bounded sample on the host.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FRAMES = 17
W, H, BD = 4096, 2048, 10
PEAK_HBM_GBS = 8000.0
PEAK_INT32_VALU_TOPS = 39.3  # 256 CU x 64 lanes x 2.4 GHz, full-rate 32-bit integer VALU ops
SEARCH_OPS_PER_PX = 2 * 25 * 12  # two directions x (2R+1)^2 candidates x 12 ops (R = 2)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--metrics", default="ivpsnr+ivssim", choices=["ivpsnr+ivssim", "all"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=2)
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(frame pairs shard across GPUs by pair; see DESIGN.md)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import qmiv_amd as Q
    from qmiv_amd import synth

    dev = torch.device("cuda:0")
    wl = synth.ERP_4K2K
    t_gen = time.time()
    pairs_np = [synth.make_pair(wl, f) for f in range(FRAMES)]
    t_gen = time.time() - t_gen
    # device-resident planes
    pairs = []
    for p0, p1 in pairs_np:
        pairs.append(([torch.from_numpy(p.view(np.int16)).to(dev) for p in p0],
                      [torch.from_numpy(p.view(np.int16)).to(dev) for p in p1]))
    torch.cuda.synchronize()
    metrics = (Q.IVQM_M_IVPSNR | Q.IVQM_M_IVSSIM) if args.metrics == "ivpsnr+ivssim" else Q.IVQM_M_ALL
    ctx = Q.IvqmContext(W, H, BD, erp=1)

    def step():
        res = []
        for f, (a, b) in enumerate(pairs):
            ring = f % 8
            if f >= 8:
                res.append(ctx.collect(ring))
            ctx.bind(0, a, True)
            ctx.bind(1, b, True)
            ctx.enqueue(metrics, ring)
        for f in range(max(0, FRAMES - 8), FRAMES):
            res.append(ctx.collect(f % 8))
        return res

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        last = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    px = FRAMES * W * H
    value = px / (ms_per_step / 1e3) / 1e6

    # kernel rooflines (HIP events on the context stream), pair 0 bound
    ctx.bind(0, pairs[0][0], True)
    ctx.bind(1, pairs[0][1], True)
    t_search = ctx.profile(1, 20)
    t_ssim = ctx.profile(2, 20)
    t_stats = ctx.profile(0, 20)
    search_tops = W * H * SEARCH_OPS_PER_PX / (t_search * 1e-3) / 1e12
    # IV-SSIM jobs read P0 and P1 (4:2:0, 3 B each per px) and both SCP pictures (6 B each)
    ssim_bytes = W * H * (3 + 3 + 6 + 6)
    ssim_gbs = ssim_bytes / (t_ssim * 1e-3) / 1e9
    stats_gbs = W * H * 6 / (t_stats * 1e-3) / 1e9

    out = {
        "metric": "IV-PSNR+IV-SSIM Mpixel-pairs/s at 4096x2048 10b",
        "value": round(value, 1),
        "unit": "Mpixel-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32 (search, SSIM moments) + fp32 (SSIM ratios) + fp64 (totals)",
        "data": "synthetic (seeded generator, qmiv_amd/csrc/synth.cpp)",
        "config": {"workload": "ERP 4096x2048 10-bit 4:2:0, 17 frame pairs, IV-PSNR + IV-SSIM per pair "
                               "(R=2, weights 4:1:1, SSIM 8x8/4), inputs resident in HBM",
                   "frames": FRAMES, "width": W, "height": H, "bit_depth": BD, "metrics": args.metrics,
                   "parallelism": "1 GPU"},
        "roofline": {"kernel": "k_search (both directions + SCP)", "bound": "valu",
                     "achieved": round(search_tops, 2), "peak": PEAK_INT32_VALU_TOPS, "unit": "TOP/s",
                     "frac": round(search_tops / PEAK_INT32_VALU_TOPS, 4), "traffic": None,
                     "ms_per_launch": round(t_search, 4),
                     "algorithmic_ops_per_px_pair": SEARCH_OPS_PER_PX},
        "roofline_streaming": {"kernel": "k_ssim (IV-SSIM, both directions)", "bound": "hbm",
                               "achieved": round(ssim_gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                               "frac": round(ssim_gbs / PEAK_HBM_GBS, 4), "ms_per_launch": round(t_ssim, 4),
                               "stats_kernel_gbs": round(stats_gbs, 1)},
        "results_frame0": {"ivpsnr": last[0].ivpsnr, "ivssim": last[0].ivssim},
        "gen_seconds": round(t_gen, 1),
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pairs_np, args.cpu_frames)
    ctx.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(pairs_np, nframes):
    """Reference library (oracle/_ref) on a bounded sample of the same workload."""
    try:
        from oracle.oracle import Reference, reference_available
    except Exception as e:  # noqa: BLE001
        return {"value": None, "error": f"reference build unavailable: {e}"}
    if not reference_available():
        return {"value": None, "error": "oracle/_ref not built"}
    ref = Reference()
    threads = min(16, os.cpu_count() or 1)
    sample = pairs_np[:nframes]
    t, iv, ss = ref.bench_pairs(sample, W, H, BD, chroma420=True, erp=True, metrics=3, threads=threads)
    return {"value": round(len(sample) * W * H / t / 1e6, 2), "unit": "Mpixel-pairs/s", "cores": threads,
            "kind": "reference",
            "sample": f"{len(sample)} of the 17 frame pairs, IV-PSNR + IV-SSIM in the app's stage order "
                      f"(4:2:0 unpack, margins, rearrange, GCD, SCP, IV-PSNR, IV-SSIM), "
                      f"reference pmbb libs x86-64-{ref.level}, {threads} pool threads",
            "seconds": round(t, 2)}


if __name__ == "__main__":
    main()
