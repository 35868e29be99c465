"""Benchmark of the per-frame non-rigid step on one MI355X.

Workload: the full-scale capsule-body sequence (256^3 TSDF at 4 mm,
640x480 depth, ~2.4k nodes, 5 GN x 30 PCG iterations, FP32 with FP64 PCG
dots).  One step = GN solve (assembly, block-Jacobi PCG, validated twist
update) + DQ-warped integration + marching cubes on one frame whose depth
and association inputs are already resident in HBM.

Solver policy.  The solver runs with the ARAP rotation Jacobian in the
sign convention of the data term (solver option arap_consistent_sign) and
applies every GN step through the reference caller's per-node clamp
(include/fusion_app.h:1353-1375) without the whole-step rejection of
fusion_app.h:409-491.  Every frame then runs its full 5 GN x 30 PCG and the
point-to-plane RMS after each frame's solve sits at the depth-noise floor
(~1.7 mm) over the sequence.  With the reference's negated ARAP rotation rows
(kernels/solver_kernels.cu:145-153) the rotations drift each GN iteration,
and with its rejection rule one poorly constrained node (raw |omega| > 1 rad)
rejects a whole frame's step, after which tracking is lost and most GN
iterations of later frames do no work.  Both policies run the same kernels;
--reference-policy selects the reference's.

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
PCG_FORMS = {1: "overlapped (A z + beta A p)", 2: "single reduction (Chronopoulos-Gear)",
             3: "pipelined (Ghysels-Vanroose)"}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def measured_traffic():
    """HBM bytes per dispatch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (profiles/hbm_traffic.json, calibrated on a copy of known size), keyed by
    kernel group; None when the file is missing or was measured on other
    kernel sources than the ones built here."""
    import hashlib
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        meas = json.load(f)
    csrc = os.path.join(ROOT, "cudafusion_amd", "csrc")
    h = hashlib.sha256()
    for fn in sorted(os.listdir(csrc)):
        with open(os.path.join(csrc, fn), "rb") as f:
            h.update(f.read())
    if h.hexdigest() != meas.get("source_sha256"):
        return None
    per = meas["bytes_per_dispatch"]
    out = dict(per)
    if "mc_classify" in per and "mc_emit" in per:
        out["mc"] = per["mc_classify"] + per["mc_emit"]
    return out


def observed_voxels(h) -> int:
    """Voxels whose kNN snapshot holds a node (the others integrate unwarped
    and their kNN weights are not read)."""
    import numpy as np
    knn, _ = h.download_knn()
    return int((knn.reshape(-1, 4).max(axis=1) >= 0).sum())


def knn_id_bytes(h) -> int:
    """kNN id bytes the warped integration reads: the 16-byte ids of every
    64-voxel run along x that holds an observed voxel, plus the run map's
    4-byte word per run (runs without one integrate unwarped unread)."""
    import numpy as np
    X, Y, Z = (int(d) for d in h.w.dims)
    knn, _ = h.download_knn()
    obs = (knn.reshape(Z, Y, X, 4).max(axis=3) >= 0)
    nsx = (X + 63) // 64
    pad = np.zeros((Z, Y, nsx * 64), bool)
    pad[:, :, :X] = obs
    occ = pad.reshape(Z, Y, nsx, 64).any(axis=3)
    run_len = np.minimum(64, X - 64 * np.arange(nsx))
    return int((occ * run_len[None, None, :]).sum()) * 16 + Z * Y * nsx * 4


def _cpu_frames(h, po, vox, T, frames, threads):
    """Per-frame seconds of the oracle's FP32 step (solve, fusion, mc) on
    frames 1..frames, the first one a warm-up."""
    knn, knn_w = h.download_knn()
    po.threads(threads)
    rec = []
    for f in range(1, frames + 1):
        corrs, can = h.frame_corrs_host(f)
        depth = h.frames_host[f].depth.reshape(-1)
        t0 = time.perf_counter()
        T, _, _, _, _ = po.solve_gn_f32(corrs, can, h.nodes, T, h.gn, h.sp)
        t1 = time.perf_counter()
        po.integrate_f32(h.vp, vox, depth, h.cam, h.pose, h.nodes, T, knn, knn_w)
        t2 = time.perf_counter()
        po.mc_f32(h.vp, vox, normals=False)
        t3 = time.perf_counter()
        rec.append((t3 - t0, t1 - t0, t2 - t1, t3 - t2))
        log(f"[cpu {threads} threads] frame {f}{' (warm-up)' if f == 1 else ''}: {t3 - t0:.2f} s "
            f"(solve {t1 - t0:.2f}, fusion {t2 - t1:.2f}, mc {t3 - t2:.2f})")
    return T, rec[1:]


def _cpu_host():
    """The host CPU as lscpu names it (model, sockets, cores per socket)."""
    import subprocess
    info = {"hardware_concurrency": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def cpu_baseline(h, frames: int, threads: int, frames_1t: int):
    """The oracle (the C++/OpenMP FP32 restatement of the step) on the same
    workload, on this host's cores: `threads` OpenMP threads over frames
    1..frames and one thread over frames 1..frames_1t, each run's first frame
    a warm-up; median and p99 of the per-frame times and per-stage medians
    under the reference profiler's names (include/fusion_app.h:700-739:
    tracking.solve, fusion, mc)."""
    import numpy as np

    from oracle import pyoracle as po
    po.build()
    vox0 = h.download_voxels().copy()
    T0 = np.zeros(h.N, h.download_transforms().dtype)
    T0["real"][:, 3] = 1
    omp_max = po.threads(0)

    def summary(rec, nthreads):
        t = np.array([r[0] for r in rec])
        return {"threads": nthreads, "frames": len(rec), "median_s": round(float(np.median(t)), 4),
                "p99_s": round(float(np.percentile(t, 99)), 4),
                "frames_per_s": round(1.0 / float(np.median(t)), 4),
                "stages_median_s": {"solve": round(float(np.median([r[1] for r in rec])), 4),
                                    "fusion": round(float(np.median([r[2] for r in rec])), 4),
                                    "mc": round(float(np.median([r[3] for r in rec])), 4)}}

    _, rec = _cpu_frames(h, po, vox0.copy(), T0.copy(), frames, threads)
    many = summary(rec, threads)
    _, rec1 = _cpu_frames(h, po, vox0.copy(), T0.copy(), frames_1t, 1)
    one = summary(rec1, 1)
    po.threads(omp_max)
    return many, one, dict(_cpu_host(), omp_get_max_threads=omp_max)


def equal_residual_leg(host, frames: int):
    """The full step with every PCG run judged by the FP64 true residual
    (below 1e-6 ||b||, correction rounds after the PCG, at most 2000
    iterations) in three forms, each on a fresh pipeline of the same
    workload: the reference's own GPU structure of the solve (solver option
    plain_form: float-atomic assembly, one kernel per operation, a host loop
    reading every PCG dot product back; FP32 with FP64 dot sums), this
    library's FP32 block-Jacobi form (fixed-order MFMA assembly, persistent
    single-reduction PCG), and the precision form (split-BF16 JtJ, FP16 TSDF
    + u8 weight, FP64 Cholesky inverses).  frames/s over `frames` frames
    after two warm-up frames, PCG work per frame and the runs that reached
    the residual."""
    import torch

    from cudafusion_amd.harness import Harness
    forms = {"reference_structure_fp32": dict(jtj_precision=0, storage16=False, precond_fp64=False, plain_form=True),
             "optimised_fp32": dict(jtj_precision=0, storage16=False, precond_fp64=False),
             "precision_bf16": dict(jtj_precision=3, storage16=True, precond_fp64=True)}
    out = {}
    for name, kw in forms.items():
        h = Harness("body", frames=host.frames, arap_consistent_sign=True, validate=False, pcg_pipelined=2,
                    relative_tolerance=True, pcg_tolerance=1e-6, true_residual_rounds=2, **kw)
        h.sp.pcg_iterations = 2000
        h.setup(host=host)
        for f in (1, 2):
            h.step(f)
        torch.cuda.synchronize()
        t0s = h.solver_totals()
        t0 = time.perf_counter()
        for f in range(3, 3 + frames):
            h.step(f)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        t1s = h.solver_totals()
        runs = t1s.pcg_runs - t0s.pcg_runs
        out[name] = {"frames_per_s": round(frames / el, 2), "ms_per_frame": round(el * 1e3 / frames, 3),
                     "pcg_iterations_per_frame": round((t1s.pcg_iterations - t0s.pcg_iterations) / frames, 1),
                     "pcg_runs": int(runs), "reached_1e-6": int(t1s.pcg_converged - t0s.pcg_converged)}
        log(f"[equal residual] {name}: {out[name]}")
        h.close()
    base = out["reference_structure_fp32"]["frames_per_s"]
    out["speedup"] = round(out["precision_bf16"]["frames_per_s"] / base, 3)
    out["speedup_optimised_fp32"] = round(out["optimised_fp32"]["frames_per_s"] / base, 3)
    out["speedup_precision_vs_optimised_fp32"] = round(out["precision_bf16"]["frames_per_s"] /
                                                       out["optimised_fp32"]["frames_per_s"], 3)
    out["target"] = ">= 5x (full step, equal 1e-6 relative PCG residual, against the plain FP32 block-Jacobi " \
                    "form in the reference's GPU structure)"
    return out


def tracking_leg(host, frames: int, line_search: bool, reference_policy: bool):
    """Frames 1..frames tracked with the device association
    (nrf_pipeline_track) on a fresh pipeline of the same workload: frames/s
    over the frames after the first, the first frame's stage times, and the
    loop's record per frame."""
    import numpy as np
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import Harness
    h = Harness("body", frames=host.frames, arap_consistent_sign=not reference_policy, validate=reference_policy,
                pcg_pipelined=3)
    h.setup(host=host)
    L.check(h.lib.nrf_solver_set_pcg_layout(h.lib.nrf_pipeline_solver(h.pipe), 4, 16), "pcg_layout")
    tp = L.track_params(min_valid_corrs=2000, line_search=line_search)
    h.track(1, tp)
    torch.cuda.synchronize()
    st = h.track_stats()
    first = h.stage_ms()
    recs = []
    t0 = time.perf_counter()
    for f in range(2, frames + 1):
        h.track(f, tp)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st_last = h.track_stats()
    n = frames - 1
    out = {"frames_per_s": round(n / el, 2), "ms_per_frame": round(el * 1e3 / n, 3), "frames": n,
           "line_search": bool(line_search),
           "policy": "reference (ARAP sign, step validation)" if reference_policy else
                     "consistent ARAP sign, per-node clamp",
           "first_frame_ms": {"tracking": round(first[0], 3), "fusion": round(first[1], 3), "mc": round(first[2], 3)},
           "first_frame": {"gn_run": st.gn_run, "gn_applied": st.gn_applied, "stop_reason": st.stop_reason,
                           "num_valid": list(st.num_valid[: st.gn_run]),
                           "best_scale": [round(v, 4) for v in st.best_scale[: st.gn_run]],
                           "rmse_m": [round(v, 6) for v in st.base_rmse[: st.gn_run]]},
           "last_frame": {"gn_run": st_last.gn_run, "gn_applied": st_last.gn_applied,
                          "stop_reason": st_last.stop_reason, "num_valid": list(st_last.num_valid[: st_last.gn_run]),
                          "valid_pixels": int(h.frames_host[frames].valid.sum())}}
    log(f"[track ls={int(line_search)} ref={int(reference_policy)}] {out}")
    h.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--frames", type=int, default=40)
    ap.add_argument("--cpu-frames", type=int, default=21,
                    help="frames of the CPU baseline on all threads (the first a warm-up)")
    ap.add_argument("--cpu-frames-1t", type=int, default=11,
                    help="frames of the one-thread CPU baseline (the first a warm-up)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--track-frames", type=int, default=12,
                    help="frames of the tracking legs (device association; 0 skips them)")
    ap.add_argument("--equal-residual-frames", type=int, default=10,
                    help="frames of the equal-residual comparison (full step, PCG to a 1e-6 true residual, "
                         "plain FP32 against the precision form; 0 skips it)")
    ap.add_argument("--pcg-sample-every", type=int, default=3, metavar="K",
                    help="time the PCG launch (HIP events around one launch, GN iterations in turn) on every K-th "
                         "timed step")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--graph", action="store_true")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run each step's integration and marching cubes on the caller's stream instead of "
                         "beside the next step's first PCG launches")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="before the timed steps, run one elementwise copy of MIB MiB (x * 1 -> y): a kernel of "
                         "known HBM traffic that calibrates hardware byte counters")
    ap.add_argument("--textbook-pcg", action="store_true",
                    help="A p recomputed each iteration instead of A z + beta A p overlapping the all-reduce")
    ap.add_argument("--pcg-form", type=int, choices=[1, 2, 3], default=None,
                    help="1: A p = A z + beta A p with the product overlapping the (r.r, r.z) all-reduce; 2: also "
                         "p.Ap from z.Az, one all-reduce per iteration (Chronopoulos-Gear); 3: pipelined "
                         "(Ghysels-Vanroose), the all-reduce overlapping the halo exchange and the product.  "
                         "Default 3 for the fixed-iteration step, 2 for the batched form (PCG to a tolerance)")
    ap.add_argument("--pcg-layout", default="", metavar="LANES,NODES",
                    help="PCG lanes per scalar row (1, 2, 4) and target nodes per workgroup (default 4,16 for the "
                         "pipelined form, the solver's 4,24 otherwise)")
    ap.add_argument("--precision", choices=["bf16", "fp16", "bf16x1", "fp16x1"], default=None,
                    help="precision form: FP16 TSDF + u8 weight, compact kNN, JtJ blocks from split BF16 / FP16 "
                         "operands on the matrix cores (bf16x1 / fp16x1: one unsplit product), FP64 Cholesky "
                         "inverses, one FP64-residual refinement step per solve; reports errors against an FP64 "
                         "run of one frame")
    ap.add_argument("--refine", type=int, default=0,
                    help="PCG iterations of the precision form's refinement step (0: the correction solve is the "
                         "block-Jacobi inverse, x += M^-1 (b - A x), one kernel)")
    ap.add_argument("--reference-policy", action="store_true",
                    help="reference ARAP rotation sign and whole-step rejection (frames lose GN work)")
    ap.add_argument("--batched", type=int, default=0, metavar="COUNT",
                    help="batched form: COUNT uneven capsule-body instances (node counts log-uniform in "
                         "500-8000, 128^3-320^3 volumes at 4-8 mm, 640x480 or 320x240 depth) stepped together, "
                         "PCG to 1e-6 relative residual; reports aggregate frames/s against the same instances "
                         "stepped one after the other")
    ap.add_argument("--batch-frames", type=int, default=12,
                    help="distinct frames per batched instance (the steps cycle over them)")
    args = ap.parse_args()
    if args.pcg_form is None:
        args.pcg_form = 2 if args.batched else 3
    # the pipelined form overlaps its waits best with more, smaller workgroups
    # (sweep on the capsule body: 16 nodes 0.62 ms of PCG per frame, 20: 0.65,
    # 24: 0.67, 14: 0.67, 32: 0.72); the other forms keep the solver's 24
    if not args.pcg_layout and args.pcg_form == 3 and not args.textbook_pcg and not args.batched:
        args.pcg_layout = "4,16"
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the step on one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    if args.batched:
        return run_batched(args)

    import ctypes as C

    import numpy as np
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import Harness

    t0 = time.perf_counter()
    ref = args.reference_policy
    prec = {"bf16": 3, "fp16": 4, "bf16x1": 1, "fp16x1": 2}.get(args.precision, 0)
    h = Harness("body", frames=args.frames, use_graph=args.graph, arap_consistent_sign=not ref, validate=ref,
                pcg_pipelined=0 if args.textbook_pcg else args.pcg_form, jtj_precision=prec,
                refine_iterations=args.refine if prec else 0, refine_jacobi=bool(prec) and args.refine == 0,
                precond_fp64=bool(prec), storage16=bool(prec))
    h.setup()
    if args.pcg_layout:
        lanes, wg_nodes = (int(v) for v in args.pcg_layout.split(","))
        L.check(h.lib.nrf_solver_set_pcg_layout(h.lib.nrf_pipeline_solver(h.pipe), lanes, wg_nodes), "pcg_layout")
    overlap = not args.no_overlap and not args.graph
    if overlap:  # each frame's fusion runs beside the next frame's first PCG launches
        L.check(h.lib.nrf_pipeline_set_overlap(h.pipe, 1), "overlap")
    log(f"[bench] setup {time.perf_counter() - t0:.1f} s, {h.N} nodes, "
        f"{int(h.frames_host[1].valid.sum())} valid pixels in frame 1")
    nf = args.frames
    seq = [1 + (i % nf) for i in range(args.warmup + args.steps)]
    lib = h.lib
    # live timing of the dominant kernel: two events around one PCG launch on
    # every K-th timed step (GN iterations in turn), which cost the step far
    # less than events around every kernel group (a few us of device time each)
    stream_ptr = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for f in seq[: args.warmup]:
        h.step(f)
    if overlap:  # the last step's fusion is enqueued by the next step or by a join
        L.check(lib.nrf_pipeline_join(h.pipe, stream_ptr), "join")
    torch.cuda.synchronize()
    if args.copy_probe > 0:
        n = args.copy_probe * (1 << 20) // 4
        xs = torch.ones(n, dtype=torch.float32, device="cuda")
        ys = torch.empty_like(xs)
        torch.cuda.synchronize()
        torch.mul(xs, 1.0, out=ys)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {n * 4} bytes read, {n * 4} bytes written")
        del xs, ys
    if not args.graph:  # the sampled steps' PCG events of the timed steps
        L.check(lib.nrf_pipeline_profile_pcg(h.pipe, max(64, args.steps)), "profile")
        L.check(lib.nrf_pipeline_profile_every(h.pipe, args.pcg_sample_every), "profile_every")
    upd0 = lib.nrf_pipeline_updated_voxels(h.pipe)
    tot0 = h.solver_totals()
    t0 = time.perf_counter()
    for f in seq[args.warmup:]:
        h.step(f)
    if overlap:
        L.check(lib.nrf_pipeline_join(h.pipe, stream_ptr), "join")
    t_enq = time.perf_counter() - t0  # host time to enqueue the steps
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    upd = (lib.nrf_pipeline_updated_voxels(h.pipe) - upd0) / args.steps
    tot1 = h.solver_totals()
    gn_per_step = (tot1.gn_done - tot0.gn_done) / args.steps
    it_per_step = (tot1.pcg_iterations - tot0.pcg_iterations) / args.steps
    rejected = tot1.rejected - tot0.rejected
    ms = h.stage_ms()
    fps = args.steps / elapsed
    log(f"[bench] {args.steps} steps in {elapsed * 1e3:.1f} ms -> {fps:.1f} frames/s; last step stages "
        f"solve {ms[0]:.3f} ms, integrate {ms[1]:.3f} ms, mc {ms[2]:.3f} ms; tris {h.triangle_count()}; "
        f"GN steps applied per frame {gn_per_step:.2f}, PCG iterations per frame {it_per_step:.1f}, "
        f"frames ending on a rejected step {rejected}")
    kern = None
    if not args.graph:
        pcg_launch = []
        for k in range((args.steps + args.pcg_sample_every - 1) // args.pcg_sample_every):
            out = (C.c_float * 6)()
            L.check(lib.nrf_pipeline_profile_read(h.pipe, k, out), "profile_read")
            pcg_launch.append(out[2])
        # [pcg per launch (mean over the timed steps), integrate, mc (last step's stages)]
        kern = [float(np.mean(pcg_launch)), ms[1], ms[2]]
        log("[bench] PCG launch %.4f ms (mean of %d, GN iterations sampled in turn); last step integrate %.4f, "
            "mc %.4f ms" % (kern[0], len(pcg_launch), kern[1], kern[2]))
    w = h.w
    V = int(np.prod(w.dims))
    G = w.gn_iterations
    nb = h.num_blocks()
    # algorithmic bytes per launch (counted once, minimum traffic of the algorithm)
    n_obs = observed_voxels(h)
    algo = {
        # kNN ids of the runs holding an observed voxel (16 B each) and the run map, kNN weights of the
        # voxels with a snapshot neighbour (16 B), read + write of updated voxels, depth, centred node table
        "integrate": knn_id_bytes(h) + n_obs * 16 + upd * 24 + 4 * w.width * w.height + 32 * h.N,
        # per iteration: BSR values + column ids, row pointers, p gather + Ap, block-Jacobi
        # blocks, r/z/x/p vector traffic of the updates and the three dot products
        "pcg": w.pcg_iterations * (nb * (36 * 4 + 4) + 4 * (h.N + 1) + 2 * 24 * h.N + h.N * 36 * 4
                                   + 2 * 24 * h.N + 3 * 2 * 24 * h.N + 3 * 3 * 24 * h.N),
        # one read of the 12-byte voxels + 36 B of vertices and 12 B of indices per triangle
        "mc": V * 12 + h.triangle_count() * 48,
    }
    # PCG: algorithmic bytes per iteration times the iterations one launch runs
    # on average (launches after a rejected step run none)
    pcg_launch_iters = it_per_step / G
    algo["pcg"] = algo["pcg"] / w.pcg_iterations * pcg_launch_iters
    if kern is not None:
        per_launch = {"integrate": kern[1], "pcg": kern[0], "mc": kern[2]}
        per_step = {"integrate": kern[1], "pcg": kern[0] * G, "mc": kern[2]}
    else:
        per_launch = {"integrate": ms[1], "pcg": float("nan"), "mc": ms[2]}
        per_step = dict(per_launch)
    if kern is None:  # graph replays: no per-kernel split; the eager line carries the roofline
        per_step = {k: (v if v == v and v > 0 else 0.0) for k, v in per_step.items()}
    dom = max(per_step, key=lambda k: per_step[k])
    achieved = algo[dom] / (per_launch[dom] * 1e-3) / 1e9 if per_launch[dom] > 0 else float("nan")
    traffic = measured_traffic()
    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle as po
        po.build()
        threads = args.cpu_threads or po.threads(0)
        many, one, host_cpu = cpu_baseline(h, args.cpu_frames, threads, args.cpu_frames_1t)
        cpu = {"value": many["frames_per_s"], "unit": "frames/s", "cores": threads, "kind": "port",
               "sample": f"frames 2..{args.cpu_frames} of the same workload after frame 1 as a warm-up, oracle "
                         f"FP32 solve_gn + warped integrate + marching cubes, OpenMP {threads} threads; value = "
                         f"1 / median frame time",
               "all_threads": many, "one_thread": one, "host": host_cpu}
    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1e3 / args.steps, 4),
        "host_enqueue_ms_per_step": round(t_enq * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "capsule-body 256^3 @4mm, 640x480, %d nodes, 5 GN x 30 PCG" % h.N,
                   "graph": bool(args.graph),
                   "overlapped_fusion": overlap,
                   "graph_captures": int(lib.nrf_pipeline_graph_captures(h.pipe)) if args.graph else None,
                   "solver_policy": ("reference (ARAP sign, whole-step rejection)" if ref else
                                     "consistent ARAP sign, per-node clamp, no whole-step rejection"),
                   "pcg_layout": args.pcg_layout or "4,24",
                   "pcg": "textbook" if args.textbook_pcg else
                          PCG_FORMS[args.pcg_form]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None or dom not in traffic else traffic[dom], "kernel": dom,
                     "launch_ms": round(per_launch[dom], 5), "algorithmic_bytes": int(algo[dom])},
        "kernel_ms_per_step": None if kern is None else {"pcg": round(kern[0] * G, 5), "integrate": round(kern[1], 5),
                                                           "mc": round(kern[2], 5)},
        "roofline_other": {k: round(algo[k] / (per_launch[k] * 1e-3) / 1e9, 1) for k in algo
                           if per_launch[k] == per_launch[k]},
        "work": {"gn_applied_per_step": round(gn_per_step, 3), "pcg_iterations_per_step": round(it_per_step, 2),
                 "rejected_steps": int(rejected), "updated_voxels_per_step": int(upd), "observed_voxels": n_obs,
                 "nodes": h.N,
                 "bsr_blocks": nb},
        "cpu_baseline": cpu,
    }
    out["pcg_to_1e-6"] = pcg_to_tolerance(h, 1e-6)
    if args.equal_residual_frames > 0 and not prec:
        out["equal_residual"] = equal_residual_leg(h, args.equal_residual_frames)
    if args.track_frames > 1 and not prec:
        # the device association (live prediction + projective correspondences every GN iteration):
        # the reference's policy with its line search, and this step's policy without one
        out["tracking"] = {"reference_line_search": tracking_leg(h, args.track_frames, True, True),
                           "step_policy": tracking_leg(h, args.track_frames, False, False)}
    if prec:
        out["config"]["workload"] += ", precision form (%s JtJ, FP16 TSDF + u8 weight, FP64 Cholesky, one " \
                                     "FP64-residual refinement step: %s)" % (
                                         {"bf16": "split BF16", "fp16": "split FP16"}.get(args.precision,
                                                                                          args.precision),
                                         "x += M^-1 (b - A x)" if args.refine == 0 else
                                         "%d PCG iterations on b - A x" % args.refine)
        out["dtype"] = args.precision + "/f16/f32"
        out["errors_vs_fp64"] = precision_errors(h, seq[-1] % nf + 1)
    print(json.dumps(out))
    h.close()
    return 0


def pcg_to_tolerance(h, rel_tol, reps=20):
    """Second half of the metric: the PCG alone on the last frame's last
    assembled system, run to ||r|| < rel_tol ||b|| (the recurrence's residual,
    at most 5000 iterations) from x0 = 0; ms per solve from HIP events over
    `reps` solves, and the true relative residual ||b - A x|| / ||b|| of the
    result recomputed in FP64.  Preconditioners: block-Jacobi, and the
    learned block-diagonal one when its trained weights are present (the
    network's forward time is reported apart from the PCG's)."""
    import ctypes as C

    import torch

    from cudafusion_amd import _lib as L
    lib = h.lib
    solver = lib.nrf_pipeline_solver(h.pipe)
    sp = L.solver_params(**{f: getattr(h.sp, f) for f, _ in L.SolverParams._fields_})
    sp.pcg_iterations, sp.pcg_tolerance, sp.relative_tolerance, sp.refine_iterations = 5000, rel_tol, 1, 0
    sp.refine_jacobi = 0
    # convergence judged by the FP64 true residual: after the PCG, correction
    # rounds on b - A x while ||b - A x|| >= tol ||b||
    sp.true_residual_rounds = 2
    # the pipelined form's recurrence residual drifts from the true one near
    # 1e-6 (6.8e-5 true at 1e-6 on this system), which would cost correction
    # rounds; this leg runs the single-reduction form
    if sp.pcg_pipelined == 3:
        sp.pcg_pipelined = 2
    L.check(lib.nrf_solver_set_params(solver, C.byref(sp)), "set_params")
    x = torch.zeros(6 * h.N, dtype=torch.float32, device="cuda")
    res = L.PcgResult()
    stream = torch.cuda.current_stream()
    L.check(lib.nrf_solver_resolve(solver, x.data_ptr(), C.byref(res), stream.cuda_stream), "resolve")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        L.check(lib.nrf_solver_resolve(solver, x.data_ptr(), None, stream.cuda_stream), "resolve")
    e1.record(stream)
    torch.cuda.synchronize()
    out = (C.c_double * 2)()
    L.check(lib.nrf_solver_true_residual(solver, x.data_ptr(), out, stream.cuda_stream), "true_residual")
    result = {"preconditioner": "block-Jacobi", "pcg": PCG_FORMS.get(int(sp.pcg_pipelined), "textbook"),
              "stop": "FP64 true residual < 1e-6 ||b|| (correction rounds after the PCG)",
              "ms": round(e0.elapsed_time(e1) / reps, 4), "converged": bool(res.converged),
              "iterations": int(res.iterations), "true_rel_residual": float(out[0] / out[1]), "learned": None}
    # learned block-diagonal preconditioner (trained weights of cudafusion_amd.train_lp): network
    # forward + fold into the block-Jacobi inverses once, then the same PCG
    from cudafusion_amd import learned as LP
    path = os.path.join(ROOT, "cudafusion_amd", "lp_weights_v0.npz")
    if os.path.exists(path):
        lp = LP.LearnedPreconditioner(0, max_nodes=h.N, max_probes=1)
        lp.load(path)
        lp.set_graph(h.nodes)
        d_nodes = lib.nrf_pipeline_nodes(h.pipe)
        f0, f1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        lp.forward(solver, d_nodes)  # warm-up (first launches load the code objects); the fold stays single
        torch.cuda.synchronize()
        f0.record(stream)
        lp.forward(solver, d_nodes)
        lp.fold(solver)
        f1.record(stream)
        fallbacks = lp.fallbacks()
        L.check(lib.nrf_solver_resolve(solver, x.data_ptr(), C.byref(res), stream.cuda_stream), "resolve")
        e0.record(stream)
        for _ in range(reps):
            L.check(lib.nrf_solver_resolve(solver, x.data_ptr(), None, stream.cuda_stream), "resolve")
        e1.record(stream)
        torch.cuda.synchronize()
        L.check(lib.nrf_solver_true_residual(solver, x.data_ptr(), out, stream.cuda_stream), "true_residual")
        result["learned"] = {"preconditioner": "BJ^-1 + L L^T, L block diagonal (3-layer MPNN, BF16 weights)",
                             "ms": round(e0.elapsed_time(e1) / reps, 4), "network_ms": round(f0.elapsed_time(f1), 4),
                             "converged": bool(res.converged), "fallbacks": int(fallbacks),
                             "iterations": int(res.iterations), "true_rel_residual": float(out[0] / out[1])}
        lp.close()
    L.check(lib.nrf_solver_set_params(solver, C.byref(h.sp)), "set_params")
    return result


def precision_errors(h, f):
    """One more (untimed) frame f, and the same frame from the same state
    run by the FP64 oracle with the same algorithm (PCG and refinement
    iterations): node transforms (max rotation / translation difference),
    TSDF (max difference on voxels away from a rounding boundary of the FP64
    projection, in truncation units; the FP64 integration from the device's
    decoded state with the device's transforms and its compact kNN decoded,
    so that it measures the integration alone) and vertices (largest
    distance from a device vertex to the nearest FP64 vertex, and back, in
    voxels)."""
    import numpy as np
    import torch
    from scipy.spatial import cKDTree

    from cudafusion_amd.harness import transform_errors
    from oracle import pyoracle as po
    po.build()
    import ctypes as C

    from cudafusion_amd import _lib as L
    T0 = h.download_transforms().copy()
    vox0 = h.download_voxels().copy()
    V = int(np.prod(h.w.dims))
    t16, w8, k16, kw16 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    L.check(h.lib.nrf_pipeline_volume16(h.pipe, C.byref(t16), C.byref(w8), C.byref(k16), C.byref(kw16)), "volume16")
    knn = h._d2h(k16.value, 4 * V, np.uint16).astype(np.int32)
    knn[knn == 0xFFFF] = -1
    knn_w = h._d2h(kw16.value, 4 * V, np.float16).astype(np.float32)
    h.step(f)
    torch.cuda.synchronize()
    T1 = h.download_transforms().copy()
    vox1 = h.download_voxels()
    mesh = h.mesh().copy()
    corrs, can = h.frame_corrs_host(f)
    T64, _, _, _, _ = po.solve_gn_f64(corrs, can, h.nodes, T0.view(np.float32).reshape(-1, 8).astype(np.float64),
                                      h.gn, h.sp)
    er, et = transform_errors(T1, T64)
    t64, w64 = vox0["tsdf"].astype(np.float64), vox0["weight"].astype(np.float64)
    depth = h.frames_host[f].depth.reshape(-1)
    Td = T1.view(np.float32).reshape(-1, 8).astype(np.float64)
    po.integrate_f64(h.vp, t64, w64, depth, h.cam, h.pose, h.nodes, Td, knn, knn_w)
    margins = po.integrate_margin_f64(h.vp, depth, h.cam, h.pose, h.nodes, Td, knn, knn_w)
    tsdf_err, amb, wmis = po.tsdf_fp64_errors(vox1, t64, w64, margins)
    _, v64, _ = po.mc_f64(h.vp, t64, w64, normals=False)
    s = h.w.voxel_size
    d1, _ = cKDTree(v64).query(mesh.reshape(-1, 3).astype(np.float64))
    d2, _ = cKDTree(mesh.reshape(-1, 3).astype(np.float64)).query(v64)
    return {"frame": int(f), "transform_rot_rad": float(er), "transform_trans_m": float(et),
            "tsdf_max": float(tsdf_err), "tsdf_weight_mismatches": int(wmis), "tsdf_ambiguous_frac": float(amb),
            "vertices_max_vox": float(max(d1.max(), d2.max()) / s), "vertices_p99_vox": float(np.percentile(d1, 99) / s),
            "triangles": [int(len(mesh) // 3), int(len(v64) // 3)]}



def run_batched(args) -> int:
    """Batched form: args.batched instances through nrf_batch_step (one launch
    per stage over all instances, tuned PCG layout and integration z chunk)
    against the same instances stepped one after the other through their own
    pipelines (default layout), both with the PCG run to 1e-6 relative
    residual (at most 300 iterations) in every GN iteration."""
    import numpy as np
    import torch

    from cudafusion_amd import scene as S
    from cudafusion_amd.harness import BatchHarness, Harness

    count = args.batched
    nf = max(1, args.batch_frames)
    t0 = time.perf_counter()
    ws = S.batch_workloads(count, frames=nf, seed=0, pcg_iterations=300)
    # convergence judged by the FP64 true residual (correction rounds after a PCG
    # that stops short of it), and data-term pairs outside the pattern kept on
    # the diagonal: with the reference's dropped pairs some instances' systems
    # are indefinite and no PCG reaches the tolerance on them
    kw = dict(frames=nf, use_graph=args.graph, pcg_tolerance=1e-6, relative_tolerance=True,
              arap_consistent_sign=True, validate=False, pcg_pipelined=args.pcg_form, true_residual_rounds=2,
              psd_dropped_pairs=True)
    solo, batch = [], []
    for w in ws:
        a = Harness(workload=w, **kw)
        a.setup()
        b = Harness(workload=w, **kw)
        b.setup(host=a)
        solo.append(a)
        batch.append(b)
        log(f"[batched] instance {len(solo) - 1}: {w.name}, {a.N} nodes ({time.perf_counter() - t0:.0f} s)")
    nodes = [h.N for h in solo]
    seq = [1 + (i % nf) for i in range(args.warmup + args.steps)]

    def pcg_outcomes(t0s, t1s):
        runs = sum(b.pcg_runs - a.pcg_runs for a, b in zip(t0s, t1s))
        conv = sum(b.pcg_converged - a.pcg_converged for a, b in zip(t0s, t1s))
        brk = sum(b.pcg_breakdowns - a.pcg_breakdowns for a, b in zip(t0s, t1s))
        return {"pcg_runs": int(runs), "reached_1e-6": int(conv), "stopped_on_breakdown": int(brk),
                "not_reached": int(runs - conv - brk)}

    # batched: tune on frame 1 (state left as it was), then warm up and time
    bh = BatchHarness(batch, use_graph=args.graph)
    tt = time.perf_counter()
    tune = bh.autotune(1, use_cache=True)
    tune_s = time.perf_counter() - tt
    log(f"[batched] tuning: {tune.lanes_per_row} lanes/row, {tune.wg_nodes} nodes/workgroup, z chunk "
        f"{tune.z_chunk}, {tune.launch_groups} PCG launch group(s); PCG {tune.pcg_ms:.3f} ms, integration "
        f"{tune.integrate_ms:.3f} ms ({tune_s:.1f} s)")
    for f in seq[: args.warmup]:
        bh.step(f)
    torch.cuda.synchronize()
    tot0 = [h.solver_totals() for h in batch]
    upd0 = [h.lib.nrf_pipeline_updated_voxels(h.pipe) for h in batch]
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t1 = time.perf_counter()
    for k, f in enumerate(seq[args.warmup:]):
        ev[k][0].record()
        bh.step(f)
        ev[k][1].record()
    torch.cuda.synchronize()
    el_b = time.perf_counter() - t1
    lat_b = np.array([a.elapsed_time(b) for a, b in ev])
    stages = bh.stage_ms()
    tot1 = [h.solver_totals() for h in batch]
    upd = [(h.lib.nrf_pipeline_updated_voxels(h.pipe) - u) / args.steps for h, u in zip(batch, upd0)]
    it_b = sum(t1_.pcg_iterations - t0_.pcg_iterations for t0_, t1_ in zip(tot0, tot1)) / (count * args.steps)
    conv_b = pcg_outcomes(tot0, tot1)
    # unbatched: every instance's own pipeline, one after the other
    for f in seq[: args.warmup]:
        for h in solo:
            h.step(f)
    torch.cuda.synchronize()
    tot0 = [h.solver_totals() for h in solo]
    evs = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in solo]
           for _ in range(args.steps)]
    t1 = time.perf_counter()
    for k, f in enumerate(seq[args.warmup:]):
        for i, h in enumerate(solo):
            evs[k][i][0].record()
            h.step(f)
            evs[k][i][1].record()
    torch.cuda.synchronize()
    el_s = time.perf_counter() - t1
    lat_s = np.array([[a.elapsed_time(b) for a, b in row] for row in evs])
    tot1 = [h.solver_totals() for h in solo]
    it_s = sum(t1_.pcg_iterations - t0_.pcg_iterations for t0_, t1_ in zip(tot0, tot1)) / (count * args.steps)
    conv_s = pcg_outcomes(tot0, tot1)
    fps_b = count * args.steps / el_b
    fps_s = count * args.steps / el_s
    log(f"[batched] batched {fps_b:.1f} frames/s (step p50 {np.percentile(lat_b, 50):.2f} ms, p99 "
        f"{np.percentile(lat_b, 99):.2f} ms; solve {stages[0]:.2f} integrate {stages[1]:.2f} mc {stages[2]:.2f} ms)"
        f" vs one after the other {fps_s:.1f} frames/s; PCG iterations per step {it_b:.1f} vs {it_s:.1f}; "
        f"outcomes {conv_b} vs {conv_s}")
    # integration roofline of the batched launch (algorithmic bytes as in the single-instance line)
    algo = 0
    for h, u in zip(batch, upd):
        algo += knn_id_bytes(h) + observed_voxels(h) * 16 + u * 24 + 4 * h.w.width * h.w.height + 32 * h.N
    int_ms = stages[1] if stages[1] > 0 else tune.integrate_ms  # graph replays have no stage split
    ach = algo / (int_ms * 1e-3) / 1e9
    out = {
        "metric": "batched non-rigid frames/s, %d uneven capsule-body instances, PCG to 1e-6 rel resid" % count,
        "value": round(fps_b, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el_b * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "%d capsule-body instances, %d-%d nodes (%d total), 128^3-320^3 at 4-8 mm, "
                               "640x480 / 320x240, 5 GN x PCG to a 1e-6 relative true residual (max 300 "
                               "iterations, up to 2 correction rounds), dropped data-term pairs on the diagonal"
                               % (count, min(nodes), max(nodes), sum(nodes)),
                   "instances": [w.name for w in ws], "graph": bool(args.graph),
                   "solver_policy": "consistent ARAP sign, per-node clamp, no whole-step rejection",
                   "pcg": PCG_FORMS[args.pcg_form]},
        "tuning": {"lanes_per_row": tune.lanes_per_row, "wg_nodes": tune.wg_nodes, "z_chunk": tune.z_chunk,
                   "pcg_launch_groups": tune.launch_groups, "tune_s": round(tune_s, 2)},
        "batched": {"frames_per_s": round(fps_b, 2), "step_ms_p50": round(float(np.percentile(lat_b, 50)), 3),
                    "step_ms_p99": round(float(np.percentile(lat_b, 99)), 3),
                    "stage_ms": {"solve": round(stages[0], 3), "integrate": round(stages[1], 3),
                                 "mc": round(stages[2], 3)},
                    "pcg_iterations_per_run": round(it_b / 5, 2), "pcg_outcomes": conv_b},
        "one_after_another": {"frames_per_s": round(fps_s, 2),
                              "instance_step_ms_p50": round(float(np.percentile(lat_s, 50)), 3),
                              "instance_step_ms_p99": round(float(np.percentile(lat_s, 99)), 3),
                              "pcg_iterations_per_run": round(it_s / 5, 2), "pcg_outcomes": conv_s},
        "speedup": round(fps_b / fps_s, 3),
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "integrate (batched)",
                     "launch_ms": round(int_ms, 5), "algorithmic_bytes": int(algo)},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    bh.close()
    for h in solo + batch:
        h.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
