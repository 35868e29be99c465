"""GPU parity of every kernel of the step against the CPU oracle, on the
128^3 / 320x240 sphere workload (all calls go through the C ABI).

Tolerances: TSDF within 1e-3 (in units of the truncation, i.e.
of the stored normalised tsdf) of the FP64 oracle; marching-cubes case
indices and triangle counts bit-exact given an identical TSDF; node
transforms within 1e-4 rad / 1e-5 m of the FP64 solver on identical inputs.
Against the FP32 oracle (same operation order) the kernels are expected to
agree to a few ulp."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sphere(gpu, orc):
    import torch

    from cudafusion_amd import scene as S
    sc = S.make_scene("sphere", frames=3)
    w = sc.w
    frames = [sc.render(f) for f in range(4)]
    nodes = S.scene_nodes(sc)
    print(f"[sphere] {len(nodes)} nodes, valid px {[int(f.valid.sum()) for f in frames]}", flush=True)
    return dict(scene=sc, w=w, frames=frames, nodes=nodes, torch=torch)


def _vol(sph):
    from cudafusion_amd.fusion import TSDFVolume
    w = sph["w"]
    return TSDFVolume(TSDFVolume.Params(w.dims, w.voxel_size, w.truncation, w.origin, w.decay_alpha,
                                        w.max_weight))


def _cam(w):
    from cudafusion_amd import _lib as L
    return L.camera(w.fx, w.fy, w.cx, w.cy, w.width, w.height)


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _vox_host(vol):
    from cudafusion_amd import _lib as L
    return vol.device_data().cpu().numpy().view(L.VOXEL)


def _perturbed_T(orc, nodes, seed, rot=0.02, trans=0.005):
    """Random small node motions (FP32 DQ records and their FP64 twin)."""
    from cudafusion_amd import _lib as L
    rng = np.random.default_rng(seed)
    N = len(nodes)
    T = np.zeros(N, L.DUALQUAT)
    T["real"][:, 3] = 1
    for i in range(N):
        tw = np.concatenate([rng.normal(0, rot, 3), rng.normal(0, trans, 3)]).astype(np.float32)
        src = T[i:i + 1].copy()
        out = np.zeros(1, L.DUALQUAT)
        orc.load().orc_twist_update_f32(tw.ctypes.data, src.ctypes.data, 0.0, 0.0, 1.0, out.ctypes.data)
        T[i] = out[0]
    return T, T.view(np.float32).reshape(-1, 8).astype(np.float64)


def test_integrate_rigid(sphere, orc):
    w = sphere["w"]
    vol = _vol(sphere)
    cam = _cam(w)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    depth = sphere["frames"][0].depth.reshape(-1)
    vol.integrate(_dev(depth), None, cam, pose)
    gpu = _vox_host(vol)
    ref = np.zeros(len(gpu), gpu.dtype)
    ref["tsdf"] = 1
    n_upd = orc.integrate_f32(vol.vp, ref, depth, cam, pose)
    t64 = np.ones(len(gpu))
    w64 = np.zeros(len(gpu))
    orc.integrate_f64(vol.vp, t64, w64, depth, cam, pose)
    assert n_upd > 10000
    np.testing.assert_array_equal(gpu["weight"], ref["weight"])
    np.testing.assert_allclose(gpu["tsdf"], ref["tsdf"], atol=2e-7)
    # FP64: a voxel whose projection sits on a pixel-rounding boundary (or
    # whose sdf sits on the truncation cut) may legitimately sample another
    # pixel in FP32; such voxels are identified from the FP64 projection
    err, amb, wmis = orc.tsdf_fp64_errors(gpu, t64, w64, orc.integrate_margin_f64(vol.vp, depth, cam, pose))
    print(f"[integrate] FP64 max err {err:.2e}, ambiguous {amb:.2e}, weight mismatches {wmis}", flush=True)
    assert wmis == 0 and amb < 0.05 and err < 1e-3


def test_voxel_knn_and_warped_integration(sphere, orc):
    from cudafusion_amd import _lib as L
    from cudafusion_amd.fusion import WarpField
    w = sphere["w"]
    nodes = sphere["nodes"]
    vol = _vol(sphere)
    cam = _cam(w)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    d0 = sphere["frames"][0].depth.reshape(-1)
    vol.integrate(_dev(d0), None, cam, pose)
    wf = WarpField(w.node_spacing)
    wf.upload_nodes(nodes)
    knn, knn_w = wf.compute_voxel_knn(vol)
    vox = _vox_host(vol)
    rk, rw = orc.knn_f32(vol.vp, vox, nodes)
    gk = knn.cpu().numpy()
    gw = knn_w.cpu().numpy()
    np.testing.assert_array_equal(gk, rk)          # ids bit-exact (same tie rule)
    np.testing.assert_allclose(gw, rw, atol=2e-6)  # expf ulp differences only
    assert (gk[0::4] >= 0).sum() > 100000
    # warped integration of frame 1 under random node motions
    T, T64 = _perturbed_T(orc, nodes, 5)
    wf.upload_transforms(T)
    d1 = sphere["frames"][1].depth.reshape(-1)
    vol_b = _vol(sphere)
    vol_b.device_data().copy_(vol.device_data())
    cen = wf.centered_table()
    vol.integrate(_dev(d1), None, cam, pose, wf.device_nodes(), wf.device_transforms(), len(nodes), knn, knn_w,
                  centered=cen)
    vol_b.integrate(_dev(d1), None, cam, pose, wf.device_nodes(), wf.device_transforms(), len(nodes), knn, knn_w)
    g1 = _vox_host(vol)
    g1b = _vox_host(vol_b)
    np.testing.assert_array_equal(g1["tsdf"], g1b["tsdf"])  # table == per-voxel dq_centered
    ref = vox.copy()
    orc.integrate_f32(vol.vp, ref, d1, cam, pose, nodes, T, rk, rw)
    np.testing.assert_array_equal(g1["weight"], ref["weight"])
    np.testing.assert_allclose(g1["tsdf"], ref["tsdf"], atol=1e-5)
    t64 = vox["tsdf"].astype(np.float64)
    w64 = vox["weight"].astype(np.float64)
    orc.integrate_f64(vol.vp, t64, w64, d1, cam, pose, nodes, T64, rk, rw)
    changed = g1["weight"] != vox["weight"]
    assert changed.sum() > 10000
    err, amb, wmis = orc.tsdf_fp64_errors(g1, t64, w64,
                                          orc.integrate_margin_f64(vol.vp, d1, cam, pose, nodes, T64, rk, rw))
    print(f"[warped] FP64 max err {err:.2e}, ambiguous {amb:.2e}, weight mismatches {wmis}", flush=True)
    assert wmis == 0 and amb < 0.05 and err < 1e-3


def test_marching_cubes(sphere, orc):
    w = sphere["w"]
    vol = _vol(sphere)
    cam = _cam(w)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    for f in range(2):
        vol.integrate(_dev(sphere["frames"][f].depth.reshape(-1)), None, cam, pose)
    v, n, t = vol.extract_surface()
    cases = vol.mc_cases()
    vox = _vox_host(vol)
    rc, rv, rn = orc.mc_f32(vol.vp, vox)
    np.testing.assert_array_equal(cases, rc)
    assert len(v) == len(rv) and len(v) > 3000
    np.testing.assert_array_equal(t.reshape(-1), np.arange(len(v)))
    np.testing.assert_allclose(v, rv, atol=1e-6)
    np.testing.assert_allclose(n, rn, atol=1e-5)
    # device form: same vertices, no normals
    vd, td = vol.extract_surface_device()
    np.testing.assert_array_equal(vd.cpu().numpy(), v)


def test_marching_cubes_reference_sphere(gpu, orc):
    """The reference's analytic sphere scene (tests/test_mesh_extraction.cu:180-219)."""
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.fusion import TSDFVolume
    p = TSDFVolume.Params((128, 128, 128), 0.01, 0.03, (-0.64, -0.64, 0.36), 1.0, 50.0)
    vol = TSDFVolume(p)
    X = np.arange(128)
    z, y, x = np.meshgrid(X, X, X, indexing="ij")
    wx = np.float32(-0.64) + x * np.float32(0.01)
    wy = np.float32(-0.64) + y * np.float32(0.01)
    wz = np.float32(0.36) + z * np.float32(0.01)
    d = np.sqrt(wx ** 2 + wy ** 2 + (wz - 1.0) ** 2) - 0.3
    vox = np.zeros(128 ** 3, L.VOXEL)
    vox["tsdf"] = np.clip(d, -0.03, 0.03).reshape(-1)
    vox["weight"] = 1
    vol.device_data().copy_(torch.from_numpy(vox.view(np.uint8).copy()).cuda())
    v, n, t = vol.extract_surface()
    assert len(v) > 100 and len(t) > 50  # test_mesh_extraction.cu:556-558
    rc, rv, _ = orc.mc_f32(vol.vp, vox)
    np.testing.assert_array_equal(vol.mc_cases(), rc)
    tri = v.reshape(-1, 3, 3)
    area = 0.5 * np.linalg.norm(np.cross(tri[:, 1] - tri[:, 0], tri[:, 2] - tri[:, 0]), axis=1).sum()
    assert abs(area - 4 * np.pi * 0.09) < 0.01 * 4 * np.pi * 0.09


def test_twist_and_warp_points(sphere, orc):
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.fusion import WarpField
    nodes = sphere["nodes"]
    N = len(nodes)
    wf = WarpField(0.025)
    wf.upload_nodes(nodes)
    T, _ = _perturbed_T(orc, nodes, 7)
    wf.upload_transforms(T)
    rng = np.random.default_rng(8)
    dx = rng.normal(0, 0.05, 6 * N).astype(np.float32)
    wf.apply_twist_increment(torch.from_numpy(dx).cuda(), 0.05, 0.01, 0.7)
    got = wf.download_transforms()
    ref = np.zeros(N, L.DUALQUAT)
    for i in range(N):
        src = T[i:i + 1].copy()
        d = dx[6 * i:6 * i + 6].copy()
        out = np.zeros(1, L.DUALQUAT)
        orc.load().orc_twist_update_f32(d.ctypes.data, src.ctypes.data, 0.05, 0.01, 0.7, out.ctypes.data)
        ref[i] = out[0]
    gv = got.view(np.float32).reshape(-1, 8)
    rv = ref.view(np.float32).reshape(-1, 8)
    np.testing.assert_allclose(gv, rv, atol=2e-6)
    # point warp through random 4-NN
    pts = (nodes["pos"][rng.integers(0, N, 500)] + rng.normal(0, 0.01, (500, 3))).astype(np.float32)
    ids = rng.integers(0, N, (500, 4)).astype(np.int32)
    ws = rng.uniform(0, 1, (500, 4)).astype(np.float32)
    ws[:, 3] = 0
    ids[:7, :] = -1
    out = torch.empty(500 * 3, dtype=torch.float32, device="cuda")
    lib = L.load()
    pt, it, wt = _dev(pts.reshape(-1)), _dev(ids.reshape(-1)), _dev(ws.reshape(-1))
    L.check(lib.nrf_warp_points(pt.data_ptr(), it.data_ptr(), wt.data_ptr(), 500, wf.device_nodes().data_ptr(),
                                wf.device_transforms().data_ptr(), N, out.data_ptr(), None), "warp_points")
    torch.cuda.synchronize()
    refp = np.zeros(1500, np.float32)
    gT = np.ascontiguousarray(gv)  # keep the array alive across the call
    orc.load().orc_warp_points_f32(pts.ctypes.data, ids.ctypes.data, ws.ctypes.data, 500, nodes.ctypes.data,
                                   gT.ctypes.data, N, refp.ctypes.data)
    np.testing.assert_allclose(out.cpu().numpy(), refp, atol=2e-6)
    np.testing.assert_array_equal(out.cpu().numpy()[:21], pts[:7].reshape(-1))


def _frame_corrs(sphere, orc, vol_knn, f):
    """Packed correspondences of frame f (device + host copies)."""
    import torch

    from cudafusion_amd import _lib as L
    w = sphere["w"]
    fr = sphere["frames"][f]
    npx = w.width * w.height
    vp = L.volume_params(w.dims, w.voxel_size, w.truncation, w.origin, w.decay_alpha, w.max_weight)
    knn, knn_w = vol_knn
    out = torch.empty(npx * 76, dtype=torch.uint8, device="cuda")
    can = _dev(fr.canonical.reshape(-1))
    dst = _dev(fr.dst.reshape(-1))
    nrm = _dev(fr.normal.reshape(-1))
    val = _dev(fr.valid.reshape(-1))
    L.check(L.load().nrf_corr_pack(C.byref(vp), knn.data_ptr(), knn_w.data_ptr(), can.data_ptr(), dst.data_ptr(),
                                   nrm.data_ptr(), None, val.data_ptr(), npx, out.data_ptr(), None), "corr_pack")
    torch.cuda.synchronize()
    return out, can, out.cpu().numpy().view(L.CORR).copy(), fr.canonical.copy()


@pytest.fixture(scope="module")
def solver_inputs(sphere, orc):
    from cudafusion_amd.fusion import WarpField
    w = sphere["w"]
    vol = _vol(sphere)
    cam = _cam(w)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    vol.integrate(_dev(sphere["frames"][0].depth.reshape(-1)), None, cam, pose)
    wf = WarpField(w.node_spacing)
    wf.upload_nodes(sphere["nodes"])
    knn = wf.compute_voxel_knn(vol)
    dcorr, dcan, hcorr, hcan = _frame_corrs(sphere, orc, knn, 2)
    print(f"[solver] valid corrs {int(hcorr['valid'].sum())}", flush=True)
    return dict(wf=wf, dcorr=dcorr, dcan=dcan, hcorr=hcorr, hcan=hcan)


@pytest.mark.parametrize("pcg", ["textbook", "pipelined", "single_reduction", "ghysels_vanroose"])
@pytest.mark.parametrize("form", ["mfma", "scalar"])
@pytest.mark.parametrize("consistent", [0, 1])
def test_solver_system_and_pcg(sphere, orc, solver_inputs, consistent, form, pcg, monkeypatch):
    """solve(): one linearisation with the stored src (GaussNewtonSolver::solve),
    in the reference's ARAP sign convention and in the opt-in consistent one,
    with the data term on the matrix cores and in the scalar LDS form, and
    with the textbook, the overlapped, the single-reduction and the pipelined
    (Ghysels-Vanroose) PCG forms."""
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.fusion import GaussNewtonSolver, copy_device_to_numpy
    w = sphere["w"]
    nodes = sphere["nodes"]
    N = len(nodes)
    wf = solver_inputs["wf"]
    T, T64 = _perturbed_T(orc, nodes, 11, rot=0.01, trans=0.003)
    wf.upload_transforms(T)
    corrs = solver_inputs["hcorr"].copy()
    # stored src = the canonical point warped by the current field (FP32 oracle)
    src = np.zeros_like(solver_inputs["hcan"])
    ids = np.ascontiguousarray(corrs["node_ids"])
    ws = np.ascontiguousarray(corrs["node_ws"])
    can = np.ascontiguousarray(solver_inputs["hcan"])
    Tc = np.ascontiguousarray(T)
    orc.load().orc_warp_points_f32(can.ctypes.data, ids.ctypes.data, ws.ctypes.data, len(corrs), nodes.ctypes.data,
                                   Tc.ctypes.data, N, src.ctypes.data)
    corrs["src"] = src.reshape(-1, 3)
    dcorr = torch.from_numpy(corrs.view(np.uint8).copy()).cuda()
    p = GaussNewtonSolver.Params(pcg_iterations=20, pcg_tolerance=0.0, lambda_smooth=w.lambda_smooth,
                                 arap_consistent_sign=bool(consistent),
                                 pcg_pipelined={"textbook": 0, "pipelined": 1, "single_reduction": 2,
                                                "ghysels_vanroose": 3}[pcg])
    monkeypatch.setenv("NRF_ASM_MFMA", "1" if form == "mfma" else "0")  # read when the solver is created
    s = GaussNewtonSolver(p, 8192)
    dx = torch.zeros(6 * N, dtype=torch.float32, device="cuda")
    res = s.solve(dcorr, len(corrs), wf.device_nodes(), wf.device_transforms(), N, dx)
    assert res.iterations == 20 and not (res.flags & 8)
    rp_p, ci_p, va_p, nb, rhs_p, pc_p = s.system_matrix()
    rp, ci, A, b, M = orc.assemble_f32(corrs, nodes, T, w.lambda_smooth, w.lambda_damping,
                                       consistent_sign=consistent)
    assert nb == len(ci)
    np.testing.assert_array_equal(copy_device_to_numpy(rp_p, N + 1, np.int32), rp)
    np.testing.assert_array_equal(copy_device_to_numpy(ci_p, nb, np.int32), ci)
    gA = copy_device_to_numpy(va_p, nb * 36, np.float32).reshape(nb, 6, 6)
    gb = copy_device_to_numpy(rhs_p, 6 * N, np.float32)
    gM = copy_device_to_numpy(pc_p, 36 * N, np.float32).reshape(N, 6, 6)
    scale = np.abs(A).max()
    np.testing.assert_allclose(gA, A, atol=2e-6 * scale)
    np.testing.assert_allclose(gb, b, atol=2e-6 * np.abs(b).max())
    _, _, A64, b64, M64 = orc.assemble_f64(corrs, nodes, T64, w.lambda_smooth, w.lambda_damping,
                                           consistent_sign=consistent)
    np.testing.assert_allclose(gA, A64, atol=1e-5 * scale)
    np.testing.assert_allclose(gM, M64, rtol=1e-3, atol=1e-4 * np.abs(M64).max())
    # PCG: 20 iterations from x0 = 0 on the same system, FP64 oracle
    x64, it, fl, r64, _ = orc.pcg_f64(rp, ci, A64, b64, M64, 20, 0.0)
    gx = dx.cpu().numpy()
    assert np.linalg.norm(gx - x64) <= 1e-3 * np.linalg.norm(x64)
    assert abs(res.final_residual - r64) <= 1e-2 * r64 + 1e-9
    # stopping rule: relative tolerance 1e-3 ends at the oracle's iteration (+-1 at the threshold)
    p.pcg_iterations, p.pcg_tolerance, p.relative_tolerance = 200, 1e-3, True
    s.set_params(p)
    res2 = s.solve(dcorr, len(corrs), wf.device_nodes(), wf.device_transforms(), N, dx)
    _, it64, fl64, r64b, _ = orc.pcg_f64(rp, ci, A64, b64, M64, 200, 1e-3, rel=1)
    print(f"[pcg {pcg}] rel 1e-3: gpu {res2.iterations} it (flags {res2.flags}), oracle {it64} it", flush=True)
    assert (res2.flags & 2) and (fl64 & 2) and abs(res2.iterations - it64) <= 1


@pytest.mark.parametrize("consistent", [0, 1])
def test_solve_gn_transforms(sphere, orc, solver_inputs, consistent):
    """3 GN x 20 PCG with re-warped src and the reference caller's step
    policy (fusion_app.h:72-119 defaults): node transforms vs the FP64 oracle,
    in both ARAP sign conventions."""
    import torch

    from cudafusion_amd import _lib as L
    from cudafusion_amd.harness import transform_errors
    from cudafusion_amd.fusion import GaussNewtonSolver
    w = sphere["w"]
    nodes = sphere["nodes"]
    N = len(nodes)
    wf = solver_inputs["wf"]
    wf.reset_transforms()
    p = GaussNewtonSolver.Params(pcg_iterations=20, pcg_tolerance=0.0, lambda_smooth=w.lambda_smooth,
                                 arap_consistent_sign=bool(consistent))
    s = GaussNewtonSolver(p, 8192)
    s.set_pattern(nodes)
    gn = L.gn_params(3, 0.25, 0.05, 0.03, 1, 0.03)
    cen = wf.centered_table()
    dx = torch.zeros(6 * N, dtype=torch.float32, device="cuda")
    res = s.solve_gn(solver_inputs["dcorr"], solver_inputs["dcan"], len(solver_inputs["hcorr"]), wf.device_nodes(),
                     wf.device_transforms(), cen, N, gn, dx)
    Tg = wf.download_transforms()
    sp = L.solver_params(3, 20, 0.0, w.lambda_smooth, 1.0, w.lambda_damping, arap_consistent_sign=consistent)
    T64, _, tot, done, _ = orc.solve_gn_f64(solver_inputs["hcorr"], solver_inputs["hcan"], nodes,
                                            orc.identity_dq64(N), gn, sp)
    print(f"[gn] gpu gn_done={res.gn_done} iters={res.total_iterations} oracle gn_done={done} iters={tot}", flush=True)
    assert res.gn_done == done == 3
    er, et = transform_errors(Tg, T64)
    print(f"[gn] rotation err {er:.3e} rad, translation err {et:.3e} m", flush=True)
    assert er < 1e-4 and et < 1e-5
    # the solve moved the nodes (non-trivial test)
    er0, et0 = transform_errors(Tg, orc.identity_dq64(N))
    assert et0 > 1e-4


def test_pipeline_step_sphere(gpu, orc):
    import torch

    from cudafusion_amd.harness import Harness
    h = Harness("sphere", frames=2)
    h.setup()
    for f in (1, 2):
        h.step(f)
    torch.cuda.synchronize()
    assert h.triangle_count() > 1000
    ms = h.stage_ms()
    assert all(m > 0 for m in ms)
    h.close()


def test_warp_normals(sphere, orc):
    import torch

    from cudafusion_amd import _lib as L
    nodes = sphere["nodes"]
    N = len(nodes)
    T, _ = _perturbed_T(orc, nodes, 13, rot=0.2, trans=0.01)
    rng = np.random.default_rng(14)
    nrm = rng.normal(0, 1, (400, 3)).astype(np.float32)
    ids = rng.integers(0, N, (400, 4)).astype(np.int32)
    ws = rng.uniform(0, 1, (400, 4)).astype(np.float32)
    ids[:5] = -1
    out = torch.empty(1200, dtype=torch.float32, device="cuda")
    dn, di, dw, dT = _dev(nrm.reshape(-1)), _dev(ids.reshape(-1)), _dev(ws.reshape(-1)), _dev(T.view(np.uint8))
    L.check(L.load().nrf_warp_normals(dn.data_ptr(), di.data_ptr(), dw.data_ptr(), 400, dT.data_ptr(), N,
                                      out.data_ptr(), None), "warp_normals")
    torch.cuda.synchronize()
    ref = np.zeros(1200, np.float32)
    Tc = np.ascontiguousarray(T)
    orc.load().orc_warp_normals_f32(nrm.ctypes.data, ids.ctypes.data, ws.ctypes.data, 400, Tc.ctypes.data, N,
                                    ref.ctypes.data)
    np.testing.assert_allclose(out.cpu().numpy(), ref, atol=2e-6)


def _sphere_system(sphere, orc, solver_inputs, **params):
    """Assembles the sphere system of frame 2 at perturbed transforms with the
    given solver options; returns (solver, dcorr, corrs, T, T64, dx, res)."""
    import torch

    from cudafusion_amd.fusion import GaussNewtonSolver
    w = sphere["w"]
    nodes = sphere["nodes"]
    N = len(nodes)
    wf = solver_inputs["wf"]
    T, T64 = _perturbed_T(orc, nodes, 11, rot=0.01, trans=0.003)
    wf.upload_transforms(T)
    corrs = solver_inputs["hcorr"].copy()
    src = np.zeros_like(solver_inputs["hcan"])
    ids = np.ascontiguousarray(corrs["node_ids"])
    ws = np.ascontiguousarray(corrs["node_ws"])
    can = np.ascontiguousarray(solver_inputs["hcan"])
    Tc = np.ascontiguousarray(T)
    orc.load().orc_warp_points_f32(can.ctypes.data, ids.ctypes.data, ws.ctypes.data, len(corrs), nodes.ctypes.data,
                                   Tc.ctypes.data, N, src.ctypes.data)
    corrs["src"] = src.reshape(-1, 3)
    dcorr = torch.from_numpy(corrs.view(np.uint8).copy()).cuda()
    p = GaussNewtonSolver.Params(pcg_iterations=20, pcg_tolerance=0.0, lambda_smooth=w.lambda_smooth, **params)
    s = GaussNewtonSolver(p, 8192)
    dx = torch.zeros(6 * N, dtype=torch.float32, device="cuda")
    res = s.solve(dcorr, len(corrs), wf.device_nodes(), wf.device_transforms(), N, dx)
    return s, dcorr, corrs, T, T64, dx, res


@pytest.mark.parametrize("prec", [1, 2], ids=["bf16", "fp16"])
def test_reduced_precision_jtj(sphere, orc, solver_inputs, prec):
    """Precision form: the data term's JtJ blocks from BF16 / FP16 matrix-core
    products (FP32 accumulation) stay within the type's rounding of the FP64
    oracle; the RHS (FP32 column) keeps the FP32 bar."""
    from cudafusion_amd.fusion import copy_device_to_numpy
    w = sphere["w"]
    nodes = sphere["nodes"]
    N = len(nodes)
    s, dcorr, corrs, T, T64, dx, res = _sphere_system(sphere, orc, solver_inputs, jtj_precision=prec)
    assert not (res.flags & 8)
    rp_p, ci_p, va_p, nb, rhs_p, pc_p = s.system_matrix()
    rp, ci, A, b, M = orc.assemble_f32(corrs, nodes, T, w.lambda_smooth, w.lambda_damping)
    _, _, A64, b64, _ = orc.assemble_f64(corrs, nodes, T64, w.lambda_smooth, w.lambda_damping)
    gA = copy_device_to_numpy(va_p, nb * 36, np.float32).reshape(nb, 6, 6)
    gb = copy_device_to_numpy(rhs_p, 6 * N, np.float32)
    scale = np.abs(A64).max()
    err = np.abs(gA - A64).max() / scale
    tol = {1: 1e-2, 2: 2e-3}[prec]
    print(f"[jtj prec {prec}] blocks max err {err:.2e} of the largest entry (bar {tol:.0e})", flush=True)
    assert 0 < err < tol
    np.testing.assert_allclose(gb, b, atol=2e-6 * np.abs(b).max())


def test_true_residual_and_refinement(sphere, orc, solver_inputs):
    """nrf_solver_true_residual equals ||b - A x|| computed in FP64 from the
    assembled FP32 system; one refinement step (FP64 residual, 10 PCG
    iterations on the correction) lowers the true residual of 10 iterations
    to that of about 20."""
    from cudafusion_amd.fusion import copy_device_to_numpy
    N = len(sphere["nodes"])
    s, dcorr, corrs, T, T64, dx, res = _sphere_system(sphere, orc, solver_inputs)
    rp_p, ci_p, va_p, nb, rhs_p, pc_p = s.system_matrix()
    rp = copy_device_to_numpy(rp_p, N + 1, np.int32)
    ci = copy_device_to_numpy(ci_p, nb, np.int32)
    A = copy_device_to_numpy(va_p, nb * 36, np.float32).reshape(nb, 6, 6).astype(np.float64)
    b = copy_device_to_numpy(rhs_p, 6 * N, np.float32).astype(np.float64)
    x = dx.cpu().numpy().astype(np.float64)
    Ax = np.zeros(6 * N)
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            Ax[6 * i:6 * i + 6] += A[q] @ x[6 * ci[q]:6 * ci[q] + 6]
    r_np, b_np = np.linalg.norm(b - Ax), np.linalg.norm(b)
    r_gpu, b_gpu = s.true_residual(dx)
    assert abs(r_gpu - r_np) <= 1e-9 * b_np and abs(b_gpu - b_np) <= 1e-12 * b_np
    rel20 = r_gpu / b_gpu
    p = s.params
    p.pcg_iterations = 10
    s.set_params(p)
    wf = solver_inputs["wf"]
    s.solve(dcorr, len(corrs), wf.device_nodes(), wf.device_transforms(), N, dx)
    rel10 = s.true_residual(dx)[0] / b_gpu
    p.refine_iterations = 10
    s.set_params(p)
    s.solve(dcorr, len(corrs), wf.device_nodes(), wf.device_transforms(), N, dx)
    rel_ref = s.true_residual(dx)[0] / b_gpu
    print(f"[refine] true relative residual: 10 it {rel10:.2e}, 10 + refine 10 {rel_ref:.2e}, 20 it {rel20:.2e}",
          flush=True)
    assert rel_ref < 0.5 * rel10 and rel_ref < 3 * rel20


def test_pipeline_step_odd_volume(gpu, orc):
    """A pipeline step on a volume whose sizes are not multiples of the
    kernels' tiles: x = 100 (one full and one partial 64-voxel run of the
    integration's kNN run map), 71 x 81 rows of cubes (several blocks of the
    marching-cubes row scan, the last one partial).  The warped integration
    equals the FP32 oracle's (weights bit-exact, TSDF to 1e-5) on the same
    transforms and kNN, and the mesh equals the oracle's marching cubes of
    the resulting volume."""
    from dataclasses import replace

    import torch

    from cudafusion_amd import scene as S
    from cudafusion_amd.harness import Harness
    w = replace(S.sphere_workload(frames=2), dims=(100, 71, 81))
    h = Harness(workload=w, frames=1, arap_consistent_sign=True, validate=False, pcg_pipelined=2)
    h.setup()
    vox0 = h.download_voxels().copy()
    knn, knn_w = h.download_knn()
    assert (knn.reshape(-1, 4).max(axis=1) >= 0).sum() > 1000
    h.step(1)
    torch.cuda.synchronize()
    vox1 = h.download_voxels().copy()
    T = h.download_transforms().copy()
    ref = vox0.copy()
    depth = h.frames_host[1].depth.reshape(-1)
    orc.integrate_f32(h.vp, ref, depth, h.cam, h.pose, h.nodes, T, knn, knn_w)
    changed = vox1["weight"] != vox0["weight"]
    print(f"[odd] {int(changed.sum())} voxels updated, {h.triangle_count()} triangles", flush=True)
    assert changed.sum() > 1000
    np.testing.assert_array_equal(vox1["weight"], ref["weight"])
    np.testing.assert_allclose(vox1["tsdf"], ref["tsdf"], atol=1e-5)
    _, rv, _ = orc.mc_f32(h.vp, vox1, normals=False)
    v = h.mesh()
    assert h.triangle_count() == len(rv) // 3 and len(rv) > 300
    np.testing.assert_allclose(v, rv, atol=1e-6)
    h.close()


@pytest.mark.parametrize("pcg", [0, 2, 3], ids=["textbook", "single_reduction", "ghysels_vanroose"])
def test_true_residual_rounds_reach_tolerance(sphere, orc, solver_inputs, pcg):
    """true_residual_rounds: after the PCG, r = b - A x in FP64 and, while
    ||r|| >= 1e-6 ||b||, correction PCG rounds on A d = r.  The solve then
    reports convergence by the true residual: one run, converged, its
    final_residual equal to nrf_solver_true_residual of the returned x and
    below 1e-6 ||b||; and x matches the FP64 oracle's dense solve of the same
    FP32 system to the accuracy the residual allows."""
    from cudafusion_amd.fusion import copy_device_to_numpy
    N = len(sphere["nodes"])
    s, dcorr, corrs, T, T64, dx, res = _sphere_system(sphere, orc, solver_inputs)
    wf = solver_inputs["wf"]
    p = s.params
    p.pcg_iterations, p.pcg_tolerance, p.relative_tolerance, p.pcg_pipelined = 300, 1e-6, True, pcg
    p.true_residual_rounds = 2
    s.set_params(p)
    res = s.solve(dcorr, len(corrs), wf.device_nodes(), wf.device_transforms(), N, dx)
    r_gpu, b_gpu = s.true_residual(dx)
    print(f"[true residual] form {pcg}: runs {res.runs} converged {res.converged} breakdowns {res.breakdowns}, "
          f"iterations {res.total_iterations}, reported {res.final_residual / res.rhs_norm:.2e}, "
          f"true {r_gpu / b_gpu:.2e}", flush=True)
    assert res.runs == 1 and res.converged == 1 and res.breakdowns == 0 and (res.flags & 2)
    assert r_gpu < 1e-6 * b_gpu
    assert abs(res.final_residual - r_gpu) <= 1e-9 * b_gpu and abs(res.rhs_norm - b_gpu) <= 1e-12 * b_gpu
    rp_p, ci_p, va_p, nb, rhs_p, _ = s.system_matrix()
    rp = copy_device_to_numpy(rp_p, N + 1, np.int32)
    ci = copy_device_to_numpy(ci_p, nb, np.int32)
    A = copy_device_to_numpy(va_p, nb * 36, np.float32).reshape(nb, 6, 6).astype(np.float64)
    b = copy_device_to_numpy(rhs_p, 6 * N, np.float32).astype(np.float64)
    D = np.zeros((6 * N, 6 * N))
    for i in range(N):
        for q in range(rp[i], rp[i + 1]):
            D[6 * i:6 * i + 6, 6 * ci[q]:6 * ci[q] + 6] = A[q]
    x_star = np.linalg.solve(D, b)
    x = dx.cpu().numpy().astype(np.float64)
    # ||x - x*|| <= ||A^-1|| ||b - A x||, ||A^-1|| = 1 / lambda_min (A SPD)
    bound = r_gpu / np.linalg.eigvalsh(0.5 * (D + D.T)).min()
    assert np.linalg.norm(x - x_star) <= 1.01 * bound + 1e-7 * np.linalg.norm(x_star)
    p.true_residual_rounds = 0
    s.set_params(p)


@pytest.mark.parametrize("prec", [0, 3], ids=["fp32", "bf16x3"])
@pytest.mark.parametrize("form", ["mfma", "scalar"])
def test_psd_dropped_pairs_system(sphere, orc, solver_inputs, form, prec, monkeypatch):
    """psd_dropped_pairs: the assembled system (data-term pairs outside the
    pattern on the diagonal) equals the FP32 oracle's with the same option,
    in the matrix-core and the scalar form, and in the split-BF16 form at
    16 significant bits."""
    from cudafusion_amd.fusion import copy_device_to_numpy
    if form == "scalar" and prec:
        pytest.skip("the reduced-precision products exist in the matrix-core form only")
    monkeypatch.setenv("NRF_ASM_MFMA", "1" if form == "mfma" else "0")
    w = sphere["w"]
    nodes = sphere["nodes"]
    N = len(nodes)
    s, dcorr, corrs, T, T64, dx, res = _sphere_system(sphere, orc, solver_inputs, psd_dropped_pairs=True,
                                                      arap_consistent_sign=True, jtj_precision=prec)
    rp_p, ci_p, va_p, nb, rhs_p, _ = s.system_matrix()
    rp, ci, A, b, M = orc.assemble_f32(corrs, nodes, T, w.lambda_smooth, w.lambda_damping, consistent_sign=1,
                                       psd_pairs=1)
    _, _, Aq, _, _ = orc.assemble_f32(corrs, nodes, T, w.lambda_smooth, w.lambda_damping, consistent_sign=1)
    gA = copy_device_to_numpy(va_p, nb * 36, np.float32).reshape(nb, 6, 6)
    gb = copy_device_to_numpy(rhs_p, 6 * N, np.float32)
    scale = np.abs(A).max()
    diff = np.abs(A - Aq).max()
    print(f"[psd pairs {form} {prec}] largest compensation {diff / scale:.2e} of the largest entry, device vs "
          f"oracle {np.abs(gA - A).max() / scale:.2e}", flush=True)
    assert diff > 1e-4 * scale  # the option changes this system
    np.testing.assert_allclose(gA, A, atol=(2e-6 if prec == 0 else 3e-5) * scale)
    np.testing.assert_allclose(gb, b, atol=2e-6 * np.abs(b).max())
