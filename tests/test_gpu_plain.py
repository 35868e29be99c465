"""The plain form of the solve (solver option plain_form): the reference's
own GPU structure (kernels/solver_kernels.cu:543-741), with float atomics
into the BSR blocks, one kernel per operation and a host loop around the PCG
that reads each dot product back.  It is the baseline of the equal-residual
comparison, so it is held to the same bars as the optimised forms: node
transforms within 1e-4 rad / 1e-5 m of the FP64 oracle's solve from the same
state, every PCG run at a 1e-6 relative true residual."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", ["sphere", "body"])
def test_plain_form_matches_fp64(gpu, orc, kind):
    import torch

    from cudafusion_amd.harness import Harness, transform_errors
    h = Harness(kind, frames=3, arap_consistent_sign=True, validate=False, pcg_pipelined=2, relative_tolerance=True,
                pcg_tolerance=1e-6, true_residual_rounds=2, plain_form=True)
    h.sp.pcg_iterations = 2000
    h.setup()
    for f in (1, 2):
        T0 = h.download_transforms().copy()
        t0 = h.solver_totals()
        h.step(f)
        torch.cuda.synchronize()
        T1 = h.download_transforms().copy()
        res = h.solver_result()
        t1 = h.solver_totals()
        corrs, can = h.frame_corrs_host(f)
        T64, _, tot, done, _ = orc.solve_gn_f64(corrs, can, h.nodes, T0.view(np.float32).reshape(-1, 8)
                                                .astype(np.float64), h.gn, h.sp)
        er, et = transform_errors(T1, T64)
        runs, conv = t1.pcg_runs - t0.pcg_runs, t1.pcg_converged - t0.pcg_converged
        print(f"[plain {kind}] frame {f}: N={h.N} gn_done {res.gn_done}/{done}, PCG runs {runs} at 1e-6 {conv}, "
              f"iterations {res.total_iterations} (oracle {tot}); vs FP64 rot {er:.2e} rad trans {et:.2e} m",
              flush=True)
        assert res.gn_done == done == h.w.gn_iterations
        assert runs == h.w.gn_iterations and conv == runs
        assert er < 1e-4 and et < 1e-5
    h.close()


def test_plain_form_refused_under_capture(gpu):
    """The host reads inside the plain form's PCG cannot be captured: a
    captured step reports an invalid argument instead of recording a graph
    that would skip them."""
    import ctypes as C

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import Harness
    h = Harness("sphere", frames=2, arap_consistent_sign=True, validate=False, use_graph=True, relative_tolerance=True,
                pcg_tolerance=1e-6, plain_form=True)
    h.setup()
    with pytest.raises(L.NrfError):
        h.step(1)
    h.close()
