"""Overlapped fusion (nrf_pipeline_set_overlap): a step's warped integration
and marching cubes are deferred to the next step's first PCG launch and run
on the pipeline's own stream alongside that PCG, from a copy of the step's
depth and a snapshot of its node transforms; the next solve's first twist
update waits only for the snapshot.  Every frame must compute
exactly what it computes unoverlapped: the same transforms, volume, mesh and
solver record, bit for bit, with the steps enqueued back to back (no host
synchronisation between them)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _pointers(h):
    """The device buffers of the volume and mesh (the getters wait for a
    deferred fusion when called; the buffers stay where they are)."""
    import ctypes as C
    t16, w8, k16, kw16 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    L_check = h.lib.nrf_pipeline_volume16(h.pipe, C.byref(t16), C.byref(w8), C.byref(k16), C.byref(kw16))
    assert L_check == 0
    cnt = C.c_void_p()
    verts = h.lib.nrf_pipeline_mesh(h.pipe, C.byref(cnt))
    return dict(vox=h.lib.nrf_pipeline_voxels(h.pipe), t16=t16.value, w8=w8.value, verts=verts, counts=cnt.value)


def _state(h, ptr=None):
    """Transforms, volume (the FP32 voxels, and the compact planes of the
    precision form), mesh and triangle count.  With `ptr` (from _pointers)
    the buffers are read by plain device copies, which wait for nothing."""
    from cudafusion_amd.fusion import copy_device_to_numpy
    if ptr is None:
        ptr = _pointers(h)
    V = int(np.prod(h.w.dims))
    vox = copy_device_to_numpy(ptr["vox"], 12 * V, np.uint8).tobytes()
    if ptr["t16"]:
        vox += copy_device_to_numpy(ptr["t16"], 2 * V, np.uint8).tobytes()
        vox += copy_device_to_numpy(ptr["w8"], V, np.uint8).tobytes()
    ntri = int(copy_device_to_numpy(ptr["counts"], 2, np.int32)[1])
    mesh = copy_device_to_numpy(ptr["verts"], 9 * min(ntri, h.cfg.max_triangles), np.float32).tobytes()
    return h.download_transforms().tobytes(), vox, mesh, ntri


def _check_same(sa, sb):
    """Equal states, compared part by part (a failure names the parts, not
    a diff of the bytes)."""
    names = ("transforms", "volume", "mesh", "triangles")
    diff = [n for n, x, y in zip(names, sa, sb) if x != y]
    assert not diff, f"overlapped steps differ in {diff} (triangles {sa[3]} / {sb[3]})"


@pytest.mark.parametrize("storage16", [False, True], ids=["fp32", "precision"])
def test_overlapped_steps_equal_sequential(gpu, storage16):
    import ctypes as C

    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import Harness
    kw = dict(frames=8, arap_consistent_sign=True, validate=False, pcg_pipelined=3)
    if storage16:
        kw.update(jtj_precision=3, storage16=True, precond_fp64=True, refine_jacobi=True)
    a = Harness("body", **kw)
    a.setup()
    b = Harness("body", **kw)
    b.setup(host=a)
    for h in (a, b):
        L.check(h.lib.nrf_solver_set_pcg_layout(h.lib.nrf_pipeline_solver(h.pipe), 4, 16), "layout")
    L.check(b.lib.nrf_pipeline_set_overlap(b.pipe, 1), "overlap")
    stream = torch.cuda.current_stream()
    pb = _pointers(b)
    # frames 1..3 back to back, then the stream-ordered join (the reads below
    # wait for nothing but the stream)
    for f in (1, 2, 3):
        a.step(f)
        torch.cuda.synchronize()
    for f in (1, 2, 3):
        b.step(f)
    L.check(b.lib.nrf_pipeline_join(b.pipe, C.c_void_p(stream.cuda_stream)), "join")
    stream.synchronize()
    sa, sb = _state(a), _state(b, pb)
    print(f"[overlap {storage16}] frame 3: {sa[3]} / {sb[3]} triangles", flush=True)
    _check_same(sa, sb)
    # frames 4.. back to back right after the join (the first waits for the
    # joined fusion), then a device synchronisation; the getters in _state
    # enqueue the last deferred fusion and wait for it
    for f in range(4, kw["frames"] + 1):
        b.step(f)
    for f in range(4, kw["frames"] + 1):
        a.step(f)
    torch.cuda.synchronize()
    _check_same(_state(a), _state(b))
    ta, tb = a.solver_totals(), b.solver_totals()
    assert (ta.pcg_iterations, ta.gn_done, ta.pcg_runs) == (tb.pcg_iterations, tb.gn_done, tb.pcg_runs)
    # a pipeline call after overlapped steps sees their results (it drains the side stream)
    assert a.lib.nrf_pipeline_updated_voxels(a.pipe) == b.lib.nrf_pipeline_updated_voxels(b.pipe)
    a.close()
    b.close()


@pytest.mark.parametrize("storage16", [False, True], ids=["fp32", "precision"])
def test_overlapped_steps_with_one_depth_buffer(gpu, storage16):
    """A caller that uploads every frame into one depth buffer, right after
    the previous step call returns and on the same stream, still gets each
    frame fused from its own depth (the deferred fusion reads the step's
    copy)."""
    import ctypes as C

    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import Harness
    kw = dict(frames=6, arap_consistent_sign=True, validate=False, pcg_pipelined=3)
    if storage16:
        kw.update(jtj_precision=3, storage16=True, precond_fp64=True, refine_jacobi=True)
    a = Harness("body", **kw)
    a.setup()
    b = Harness("body", **kw)
    b.setup(host=a)
    for h in (a, b):
        L.check(h.lib.nrf_solver_set_pcg_layout(h.lib.nrf_pipeline_solver(h.pipe), 4, 16), "layout")
    L.check(b.lib.nrf_pipeline_set_overlap(b.pipe, 1), "overlap")
    for f in range(1, 7):
        a.step(f)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    buf = torch.empty_like(b.depth[1])
    buf.copy_(b.depth[1])
    for f in range(1, 7):
        L.check(b.lib.nrf_pipeline_step(b.pipe, buf.data_ptr(), b.pose.ctypes.data, b.corrs[f].data_ptr(),
                                        b.canonical[f].data_ptr(), b.w.width * b.w.height, stream.cuda_stream),
                "pipeline_step")
        if f < 6:
            buf.copy_(b.depth[f + 1])  # the next frame into the same buffer, stream-ordered
    L.check(b.lib.nrf_pipeline_join(b.pipe, C.c_void_p(stream.cuda_stream)), "join")
    stream.synchronize()
    _check_same(_state(a), _state(b))
    a.close()
    b.close()


def test_overlap_refused_in_a_batch(gpu):
    """Batch members step unoverlapped: a pipeline with overlap on is refused."""
    import ctypes as C

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import Harness
    h = Harness("sphere", frames=2, arap_consistent_sign=True, validate=False, pcg_pipelined=2)
    h.setup()
    L.check(h.lib.nrf_pipeline_set_overlap(h.pipe, 1), "overlap")
    arr = (C.c_void_p * 1)(h.pipe)
    assert not h.lib.nrf_batch_create(arr, 1, None, 0)
    assert h.lib.nrf_pipeline_set_overlap(h.pipe, 2) != 0
    h.close()
