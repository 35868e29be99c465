"""Overlapped fusion (nrf_pipeline_set_overlap): a step's warped integration
and marching cubes run on the pipeline's own stream while the next step's
memberships and first assembly run on the caller's, and the next solve waits
for them before its first PCG and twist update.  Every frame must compute
exactly what it computes unoverlapped: the same transforms, volume, mesh and
solver record, bit for bit, with the steps enqueued back to back (no host
synchronisation between them)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _state(h):
    """Transforms, volume (the FP32 voxels, and the compact planes of the
    precision form), mesh and triangle count."""
    import ctypes as C

    from cudafusion_amd.fusion import copy_device_to_numpy
    V = int(np.prod(h.w.dims))
    vox = copy_device_to_numpy(h.lib.nrf_pipeline_voxels(h.pipe), 12 * V, np.uint8).tobytes()
    t16, w8, k16, kw16 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    h.lib.nrf_pipeline_volume16(h.pipe, C.byref(t16), C.byref(w8), C.byref(k16), C.byref(kw16))
    if t16.value:
        vox += copy_device_to_numpy(t16.value, 2 * V, np.uint8).tobytes()
        vox += copy_device_to_numpy(w8.value, V, np.uint8).tobytes()
    return h.download_transforms().tobytes(), vox, h.mesh().tobytes(), h.triangle_count()


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
    # frames 1..3 back to back, then the stream-ordered join
    for f in (1, 2, 3):
        a.step(f)
        torch.cuda.synchronize()
    for f in (1, 2, 3):
        b.step(f)
    L.check(b.lib.nrf_pipeline_join(b.pipe, C.c_void_p(stream.cuda_stream)), "join")
    stream.synchronize()
    sa, sb = _state(a), _state(b)
    print(f"[overlap {storage16}] frame 3: {sa[3]} / {sb[3]} triangles", flush=True)
    assert sa == sb
    # frames 4..7 back to back, then a device synchronisation
    for f in (4, 5, 6, 7):
        a.step(f)
        b.step(f)
    torch.cuda.synchronize()
    assert _state(a) == _state(b)
    ta, tb = a.solver_totals(), b.solver_totals()
    assert (ta.pcg_iterations, ta.gn_done, ta.pcg_runs) == (tb.pcg_iterations, tb.gn_done, tb.pcg_runs)
    # a pipeline call after overlapped steps sees their results (it drains the side stream)
    assert a.lib.nrf_pipeline_updated_voxels(a.pipe) == b.lib.nrf_pipeline_updated_voxels(b.pipe)
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
