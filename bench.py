"""Benchmark of the per-frame non-rigid step on one MI355X.

Workload: the full-scale capsule-body sequence (256^3 TSDF at 4 mm,
640x480 depth, ~2.4k nodes, 5 GN x 30 PCG iterations, FP32 with FP64 PCG
dots).  One step = GN solve (assembly, block-Jacobi PCG, validated twist
update) + DQ-warped integration + marching cubes on one frame whose depth
and association inputs are already resident in HBM.

Prints one JSON line (value = frames/s over the timed steps).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "non-rigid frames/s at 256³,640×480; PCG ms to 1e-6 rel resid: learned vs Jacobi"
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(h, frames: int, threads: int):
    """The oracle (C++/OpenMP restatement, FP32) on `frames` frames of the
    same workload: GN solve + warped integration + marching cubes."""
    import numpy as np

    from oracle import pyoracle as po
    po.build()
    po.threads(threads)
    vox = h.download_voxels().copy()
    knn, knn_w = h.download_knn()
    T = np.zeros(h.N, h.download_transforms().dtype)
    T["real"][:, 3] = 1
    pose = h.pose
    times = []
    for f in range(1, frames + 1):
        corrs, can = h.frame_corrs_host(f)
        depth = h.frames_host[f].depth.reshape(-1)
        t0 = time.perf_counter()
        T, _, _, _, _ = po.solve_gn_f32(corrs, can, h.nodes, T, h.gn, h.sp)
        po.integrate_f32(h.vp, vox, depth, h.cam, pose, h.nodes, T, knn, knn_w)
        po.mc_f32(h.vp, vox, normals=False)
        times.append(time.perf_counter() - t0)
        log(f"[cpu] frame {f}: {times[-1]:.2f} s")
    return times


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--cpu-frames", type=int, default=2)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the step on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    import torch

    from cudafusion_amd.harness import Harness

    t0 = time.perf_counter()
    h = Harness("body", frames=args.frames, use_graph=args.graph)
    h.setup()
    log(f"[bench] setup {time.perf_counter() - t0:.1f} s, {h.N} nodes, "
        f"{int(h.frames_host[1].valid.sum())} valid pixels in frame 1")
    nf = args.frames
    seq = [1 + (i % nf) for i in range(args.warmup + args.steps)]
    for f in seq[: args.warmup]:
        h.step(f)
    torch.cuda.synchronize()
    stage = [0.0, 0.0, 0.0]
    t0 = time.perf_counter()
    for f in seq[args.warmup:]:
        h.step(f)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms = h.stage_ms()
    fps = args.steps / elapsed
    log(f"[bench] {args.steps} steps in {elapsed * 1e3:.1f} ms -> {fps:.1f} frames/s; last step stages "
        f"solve {ms[0]:.3f} ms, integrate {ms[1]:.3f} ms, mc {ms[2]:.3f} ms; tris {h.triangle_count()}")
    # integration roofline: knn (32 B/voxel) + voxel read/write of updated voxels + depth + node table
    V = 256 ** 3
    integ_bytes = V * 32 + 4 * 640 * 480 + 32 * h.N
    achieved = integ_bytes / (ms[1] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        times = cpu_baseline(h, args.cpu_frames, threads)
        cpu = {"value": round(len(times) / sum(times), 4), "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"{len(times)} frames of the same workload (frames 1..{len(times)}), oracle FP32 "
                         f"solve_gn + warped integrate + marching cubes, OpenMP {threads} threads"}
    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "capsule-body 256^3 @4mm, 640x480, %d nodes, 5 GN x 30 PCG" % h.N,
                   "graph": bool(args.graph)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "integrate"},
        "stages_ms": {"solve": ms[0], "integrate": ms[1], "mc": ms[2]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
