"""Batched form: several independent pipelines of uneven size stepped by
common launches (nrf_batch_*).  A batched step of an instance must give the
same bits as that pipeline's own step with the same PCG layout: node
transforms, solver record, TSDF volume and mesh, over several frames, with
eager launches and with the step captured in a hipGraph; the autotuner must
leave the instances' state untouched."""
import ctypes as C
from dataclasses import replace

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _workloads():
    from cudafusion_amd import scene as S
    base = S.sphere_workload(frames=3)
    return [
        (replace(base, seed=0), dict(validate=True)),
        (replace(base, dims=(96, 96, 96), voxel_size=0.008, truncation=0.024, node_spacing=0.020, seed=1),
         dict(validate=False, arap_consistent_sign=True)),
        (replace(base, dims=(112, 112, 112), voxel_size=0.007, truncation=0.021, origin=(-0.392, -0.392, 0.608),
                 node_spacing=0.032, seed=2), dict(validate=False, arap_consistent_sign=True)),
    ]


def _sets(pcg_form=1):
    """Two harness sets over the same workloads (the second reuses the
    first's rendered frames and nodes); pcg_form 1 overlapped, 2 single
    reduction, 3 pipelined."""
    from cudafusion_amd.harness import Harness
    a, b = [], []
    for w, kw in _workloads():
        ha = Harness(workload=w, frames=3, pcg_pipelined=pcg_form, pcg_tolerance=1e-6, relative_tolerance=True,
                     **kw)
        ha.setup()
        hb = Harness(workload=w, frames=3, pcg_pipelined=pcg_form, pcg_tolerance=1e-6, relative_tolerance=True,
                     **kw)
        hb.setup(host=ha)
        a.append(ha)
        b.append(hb)
    return a, b


def _state(h):
    r = h.solver_result()
    return (h.download_transforms().copy(), h.download_voxels().copy(), h.mesh().copy(),
            (r.iterations, r.total_iterations, r.flags, r.gn_done))


def _compare(a, b, tag):
    for k, (ha, hb) in enumerate(zip(a, b)):
        Ta, va, ma, ra = _state(ha)
        Tb, vb, mb, rb = _state(hb)
        print(f"[batch {tag}] instance {k}: N={ha.N} solver {ra} vs {rb}, triangles {len(ma) // 3} vs "
              f"{len(mb) // 3}", flush=True)
        assert ra == rb
        assert Ta.tobytes() == Tb.tobytes()
        assert va.tobytes() == vb.tobytes()
        assert ma.tobytes() == mb.tobytes() and len(ma) > 300


def _set_layout(hs, t):
    from cudafusion_amd import _lib as L
    for h in hs:
        L.check(h.lib.nrf_solver_set_pcg_layout(h.lib.nrf_pipeline_solver(h.pipe), t.lanes_per_row, t.wg_nodes),
                "set_pcg_layout")


@pytest.mark.parametrize("pcg_form", [1, 3], ids=["overlapped", "pipelined"])
def test_batch_matches_unbatched_eager(gpu, pcg_form):
    import torch

    from cudafusion_amd.harness import BatchHarness
    a, b = _sets(pcg_form)
    bh = BatchHarness(b)
    t = bh.tuning()
    assert (t.lanes_per_row, t.wg_nodes, t.z_chunk, t.launch_groups) == (4, 32, 4, 1)
    _set_layout(a, t)
    for f in (1, 2, 3):
        for h in a:
            h.step(f)
        bh.step(f)
        torch.cuda.synchronize()
        _compare(a, b, f"eager f{f}")
    bh.close()
    for h in a + b:
        h.close()


def test_batch_autotune_and_graph(gpu):
    """Tuned layout, graph replays and the single-reduction PCG form."""
    import torch

    from cudafusion_amd.harness import BatchHarness
    a, b = _sets(pcg_form=2)
    bh = BatchHarness(b, use_graph=True)
    before = [_state(h) for h in b]
    tot_before = [(t.solves, t.gn_done, t.pcg_iterations) for t in (h.solver_totals() for h in b)]
    t = bh.autotune(1, use_cache=False)
    print(f"[batch tune] lanes {t.lanes_per_row} nodes/wg {t.wg_nodes} z_chunk {t.z_chunk} groups "
          f"{t.launch_groups} pcg {t.pcg_ms:.3f} ms integrate {t.integrate_ms:.3f} ms", flush=True)
    assert t.lanes_per_row in (1, 2, 4) and t.z_chunk in (1, 2, 4, 8, 16) and t.pcg_ms > 0 and t.integrate_ms > 0
    after = [_state(h) for h in b]
    for x, y in zip(before, after):  # the tuner leaves transforms, volume, mesh and record as they were
        for u, v in zip(x, y):
            assert (u.tobytes() == v.tobytes()) if isinstance(u, np.ndarray) else u == v
    assert tot_before == [(t2.solves, t2.gn_done, t2.pcg_iterations) for t2 in (h.solver_totals() for h in b)]
    t2 = bh.autotune(1, use_cache=True)  # same shapes: the cached choice
    assert (t2.lanes_per_row, t2.wg_nodes, t2.z_chunk) == (t.lanes_per_row, t.wg_nodes, t.z_chunk)
    _set_layout(a, t)
    for i, f in enumerate((1, 1, 2, 3)):  # the second step replays the first one's graph
        for h in a:
            h.step(f)
        bh.step(f)
        torch.cuda.synchronize()
        _compare(a, b, f"graph step {i} f{f}")
    bh.close()
    for h in a + b:
        h.close()


def test_batch_rejects_mismatched_instances(gpu):
    """The batch needs one kernel form for all: different GN iteration
    counts (or the textbook PCG) are refused at creation."""
    from cudafusion_amd import _lib as L
    from cudafusion_amd import scene as S
    from cudafusion_amd.harness import Harness
    w = S.sphere_workload(frames=1)
    h1 = Harness(workload=w, frames=1, pcg_pipelined=True)
    h1.setup()
    h2 = Harness(workload=replace(w, gn_iterations=2), frames=1, pcg_pipelined=True)
    h2.setup(host=h1)
    h3 = Harness(workload=w, frames=1, pcg_pipelined=False)
    h3.setup(host=h1)
    lib = L.load()
    for pair in ((h1, h2), (h1, h3), (h1, h1)):
        arr = (C.c_void_p * 2)(*[x.pipe for x in pair])
        assert not lib.nrf_batch_create(arr, 2, None, 0)
    for h in (h1, h2, h3):
        h.close()


def test_pipeline_graph_replay_matches_eager(gpu):
    """One pipeline with use_graph: captured once on its own stream and
    replayed on the caller's for every frame, whose inputs (other depth,
    correspondence and canonical buffers, a camera pose that moves each step)
    are copied into the captured step's buffers ahead of the replay.  Every
    replay equals the eager step bit for bit."""
    import torch

    from cudafusion_amd import scene as S
    from cudafusion_amd.harness import Harness
    w = S.sphere_workload(frames=3)
    a = Harness(workload=w, frames=3, pcg_pipelined=True)
    a.setup()
    b = Harness(workload=w, frames=3, pcg_pipelined=True, use_graph=True)
    b.setup(host=a)
    for i, f in enumerate((1, 1, 2, 2, 3, 1)):
        pose = np.eye(4, dtype=np.float32)
        pose[0, 3] = 0.0005 * i  # the camera moves 0.5 mm per step
        a.pose = b.pose = pose.reshape(16)
        a.step(f)
        b.step(f)
        torch.cuda.synchronize()
        _compare([a], [b], f"pipeline graph step {i} f{f}")
        ms = b.stage_ms()
        assert ms[0] > 0 and ms[1] == -1
        assert b.lib.nrf_pipeline_graph_captures(b.pipe) == 1
    a.close()
    b.close()


def test_pipeline_pcg_profile_sampling(gpu):
    """nrf_pipeline_profile_pcg: two events per step around the PCG launch of
    GN iteration (step mod G).  Each sampled launch takes positive time, the
    other entries of the record are zero, and the step's results do not
    depend on the profiling mode.  The sampled mean and the per-kernel
    ring's PCG time per launch are printed side by side (wall-clock values
    of a shared device: not asserted)."""
    import ctypes as C

    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd import scene as S
    from cudafusion_amd.harness import Harness
    w = S.sphere_workload(frames=3)
    a = Harness(workload=w, frames=3, pcg_pipelined=3)
    a.setup()
    b = Harness(workload=w, frames=3, pcg_pipelined=3)
    b.setup(host=a)
    G = w.gn_iterations
    L.check(a.lib.nrf_pipeline_profile_pcg(a.pipe, 16), "profile_pcg")
    L.check(b.lib.nrf_pipeline_profile(b.pipe, 16), "profile")
    frames = [1, 2, 3] * G
    sampled, full = [], []
    for f in frames:
        a.step(f)
        b.step(f)
        torch.cuda.synchronize()
        _compare([a], [b], f"profiled step f{f}")
        out = (C.c_float * 6)()
        L.check(a.lib.nrf_pipeline_profile_read(a.pipe, 0, out), "profile_read")
        assert 0 < out[2] and out[0] == 0 and out[4] == 0
        sampled.append(out[2])
        L.check(b.lib.nrf_pipeline_profile_read(b.pipe, 0, out), "profile_read")
        full.append(out[2] / G)
    ms_s, ms_f = float(np.mean(sampled)), float(np.mean(full))
    print(f"[profile] PCG launch {ms_s:.4f} ms sampled, {ms_f:.4f} ms from the full ring", flush=True)
    assert ms_s > 0 and ms_f > 0
    # every third step only: 7 steps give the sampled steps 0, 3, 6, whose
    # records are the ring's last three (a fourth does not exist)
    L.check(a.lib.nrf_pipeline_profile_pcg(a.pipe, 16), "profile_pcg")
    L.check(a.lib.nrf_pipeline_profile_every(a.pipe, 3), "profile_every")
    assert a.lib.nrf_pipeline_profile_every(a.pipe, 0) != 0
    for f in [1, 2, 3, 1, 2, 3, 1]:
        a.step(f)
    torch.cuda.synchronize()
    out = (C.c_float * 6)()
    for k in range(3):
        L.check(a.lib.nrf_pipeline_profile_read(a.pipe, k, out), "profile_read")
        assert 0 < out[2] and out[0] == 0
    assert a.lib.nrf_pipeline_profile_read(a.pipe, 3, out) != 0
    a.close()
    b.close()


def _chunks_equal(pa, pb, nbytes, chunk=1 << 26):
    """Device buffers pa and pb hold the same bytes (compared in chunks)."""
    from cudafusion_amd.fusion import copy_device_to_numpy
    for off in range(0, nbytes, chunk):
        n = min(chunk, nbytes - off)
        if copy_device_to_numpy(pa + off, n, np.uint8).tobytes() != copy_device_to_numpy(pb + off, n, np.uint8).tobytes():
            return False
    return True


def test_batch_config4_workload(gpu, orc):
    """The batched form at its workload: 16 uneven capsule-body instances
    (node counts log-uniform in 500-8000, 128^3-320^3 volumes at 4-8 mm,
    640x480 or 320x240 depth), 5 GN iterations each with the single-reduction
    PCG run to a 1e-6 relative residual judged by the FP64 true residual (at
    most 300 iterations, up to two correction rounds), the tuner's layout,
    and data-term pairs outside the pattern kept on the diagonal
    (psd_dropped_pairs): with the reference's dropped pairs the system of
    some instances is indefinite (instance 2: two negative eigenvalues in
    FP64, where CG breaks down at a 0.29 relative residual), so that no PCG
    reaches 1e-6 on it.
    One frame:
      * every instance gives the bits of its own pipeline's step at the same
        PCG layout (transforms, solver record, TSDF volume, mesh);
      * every PCG run of every instance reached 1e-6 by its true residual;
      * the smallest and the largest instance match the FP64 oracle's solve
        from the same state (PCG to 1e-6 relative) within 1e-4 rad /
        1e-5 m."""
    import time

    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd import scene as S
    from cudafusion_amd.harness import BatchHarness, Harness, transform_errors
    t0 = time.perf_counter()
    ws = S.batch_workloads(16, frames=1, seed=0, pcg_iterations=300)
    kw = dict(frames=1, pcg_tolerance=1e-6, relative_tolerance=True, arap_consistent_sign=True, validate=False,
              pcg_pipelined=2, true_residual_rounds=2, psd_dropped_pairs=True)
    solo, batch = [], []
    for w in ws:
        a = Harness(workload=w, **kw)
        a.setup()
        b = Harness(workload=w, **kw)
        b.setup(host=a)
        solo.append(a)
        batch.append(b)
    print(f"[config4] {len(ws)} instances, nodes {[h.N for h in solo]}, setup {time.perf_counter() - t0:.1f} s",
          flush=True)
    bh = BatchHarness(batch)
    tune = bh.autotune(1, use_cache=False)
    print(f"[config4] tuning {tune.lanes_per_row} lanes/row, {tune.wg_nodes} nodes/wg, z chunk {tune.z_chunk}, "
          f"{tune.launch_groups} launch group(s)", flush=True)
    _set_layout(solo, tune)
    T0 = [h.download_transforms().copy() for h in solo]
    tot0 = [h.solver_totals() for h in batch]
    bh.step(1)
    for h in solo:
        h.step(1)
    torch.cuda.synchronize()
    for k, (a, b) in enumerate(zip(solo, batch)):
        ra, rb = a.solver_result(), b.solver_result()
        ta = b.solver_totals()
        runs, conv = ta.pcg_runs - tot0[k].pcg_runs, ta.pcg_converged - tot0[k].pcg_converged
        print(f"[config4] instance {k}: N={a.N} gn_done {rb.gn_done}, PCG runs {runs} reached 1e-6 {conv}, "
              f"iterations {rb.total_iterations}, true residual {rb.final_residual / max(rb.rhs_norm, 1e-300):.2e}",
              flush=True)
        assert (ra.iterations, ra.total_iterations, ra.flags, ra.gn_done) == \
            (rb.iterations, rb.total_iterations, rb.flags, rb.gn_done)
        assert rb.gn_done == 5 and runs == 5 and conv == runs
        assert a.download_transforms().tobytes() == b.download_transforms().tobytes()
        V = int(np.prod(a.w.dims))
        assert _chunks_equal(a.lib.nrf_pipeline_voxels(a.pipe), b.lib.nrf_pipeline_voxels(b.pipe), 12 * V)
        assert a.mesh().tobytes() == b.mesh().tobytes()
    order = np.argsort([h.N for h in solo])
    for k in (int(order[0]), int(order[-1])):
        h = solo[k]
        corrs, can = h.frame_corrs_host(1)
        T64, _, _, done, _ = orc.solve_gn_f64(corrs, can, h.nodes, T0[k].view(np.float32).reshape(-1, 8)
                                              .astype(np.float64), h.gn, h.sp)
        er, et = transform_errors(h.download_transforms(), T64)
        print(f"[config4] instance {k} (N={h.N}) vs FP64: rot {er:.2e} rad, trans {et:.2e} m", flush=True)
        assert done == 5 and er < 1e-4 and et < 1e-5
    bh.close()
    for h in solo + batch:
        h.close()
