"""Pins of the oracle's tracking frame (oracle/pyoracle.py track_frame_f32,
restating include/fusion_app.h:253-669 from the oracle's stages):

* the correspondence RMSE (fusion_app.h:1331-1351) by its closed form on
  slots whose residual is known, with invalid and zero-weight slots
  ignored and infinity without weight;
* the trial step (upload_scaled_delta_x then apply_twist_increment,
  fusion_app.h:1353-1375): the per-node clamp comes before the scale, so a
  clamped step at scale s rotates by s * max_rot (scaling first and
  clamping after would give max_rot), and a negative scale reverses it;
* the loop on the 128^3 sphere: every applied line-search step lowers the
  RMSE below the iteration's base (the search keeps a trial only then), the
  association after the last step has a lower RMSE than the first, and
  without the line search the steps are those of the plain GN update.
"""
import numpy as np

from cudafusion_amd import _lib as L


def _corrs(d, w, valid):
    c = np.zeros(len(d), L.CORR)
    c["src"][:, 2] = 1.0 + np.asarray(d, np.float32)
    c["dst"][:, 2] = 1.0
    c["normal"][:, 2] = 1.0
    c["weight"] = w
    c["valid"] = valid
    return c


def test_corr_rmse_closed_form(orc):
    d = np.array([0.01, -0.02, 0.03, 0.5, 0.7], np.float32)
    w = np.array([1.0, 0.5, 0.25, 0.0, 2.0], np.float32)
    v = np.array([1, 1, 1, 1, 0], np.uint8)
    got = orc.corr_rmse(_corrs(d, w, v))
    # the residual is the FP32 difference src - dst (fusion_app.h:1341-1342)
    dd = ((np.float32(1) + d) - np.float32(1)).astype(np.float64)[:3]
    ww = w.astype(np.float64)[:3]
    assert abs(got - np.sqrt((ww * dd * dd).sum() / ww.sum())) < 1e-9
    assert np.isinf(orc.corr_rmse(_corrs(d, np.zeros(5, np.float32), v)))


def _angle_and_t(T):
    from cudafusion_amd.harness import dq_to_rt
    R, t = dq_to_rt(np.ascontiguousarray(T).view(np.float32).reshape(-1, 8)[0])
    return float(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1))), t, R


def test_trial_step_clamps_then_scales(orc):
    T0 = np.zeros(1, L.DUALQUAT)
    T0["real"][0, 3] = 1
    gn = L.gn_params(1, 1.0, 0.05, 0.03, 1, 0.0)
    dx = np.array([0, 0, 0.2, 0.1, 0, 0], np.float32)  # |w| = 0.2 > 0.05, |v| = 0.1 > 0.03
    for s in (1.0, 0.5, 0.25, -0.5):
        ang, t, R = _angle_and_t(orc.apply_step_f32(dx, T0, gn, s))
        assert abs(ang - abs(s) * 0.05) < 2e-6, s
        assert np.sign(R[1, 0]) == np.sign(s)
        # the translation of exp(twist): |v| scaled by s (a rotation about z
        # keeps the x-direction translation in the xy-plane)
        assert abs(np.linalg.norm(t) - abs(s) * 0.03) < 3e-4 * abs(s) + 1e-6, s
    # an unclamped step is simply scaled
    small = np.array([0, 0, 0.01, 0.005, 0, 0], np.float32)
    ang, t, _ = _angle_and_t(orc.apply_step_f32(small, T0, gn, 0.5))
    assert abs(ang - 0.005) < 1e-6


def _sphere_state(orc):
    from cudafusion_amd import scene as S
    sc = S.make_scene("sphere", frames=2)
    w = sc.w
    vp = L.volume_params(w.dims, w.voxel_size, w.truncation, w.origin, w.decay_alpha, w.max_weight)
    cam = L.camera(w.fx, w.fy, w.cx, w.cy, w.width, w.height)
    pose = np.eye(4, dtype=np.float32).reshape(16)
    frames = [sc.render(f) for f in range(2)]
    vox = np.zeros(int(np.prod(w.dims)), L.VOXEL)
    vox["tsdf"] = 1
    orc.integrate_f32(vp, vox, frames[0].depth.reshape(-1).astype(np.float32), cam, pose)
    nodes = S.scene_nodes(sc)
    knn, knn_w = orc.knn_f32(vp, vox, nodes)
    _, verts, _ = orc.mc_f32(vp, vox, normals=False)
    T = np.zeros(len(nodes), L.DUALQUAT)
    T["real"][:, 3] = 1
    sp = L.solver_params(w.gn_iterations, w.pcg_iterations, 0.0, w.lambda_smooth, 1.0, w.lambda_damping)
    gn = L.gn_params(w.gn_iterations, w.update_scale, w.max_update_rot, w.max_update_trans, 1, w.max_dx_mean)
    return dict(w=w, vp=vp, cam=cam, pose=pose, frames=frames, vox=vox, nodes=nodes, knn=knn, knn_w=knn_w,
                verts=verts, T=T, sp=sp, gn=gn)


def test_track_frame_line_search_lowers_rmse(orc):
    s = _sphere_state(orc)
    V = len(s["vox"])
    depth = s["frames"][1].depth.reshape(-1).astype(np.float32)
    for ls in (1, 0):
        tp = L.track_params(min_valid_corrs=2000, line_search=ls)
        opt, optf = np.zeros(V, np.int32), np.zeros(V, np.int32)
        T, st, corrs, hit = orc.track_frame_f32(s["vp"], s["cam"], s["pose"], depth, s["verts"], s["nodes"], s["T"],
                                                s["knn"], s["knn_w"], s["gn"], s["sp"], tp, 1, opt, optf)
        print(f"[oracle track ls={ls}] {st}")
        assert st["gn_applied"] >= 1 and st["num_valid"][0] > 10000
        assert st["stop_reason"] in (0, 1, 2, 3, 4)
        if ls:
            for base, best, sc in zip(st["base_rmse"], st["best_rmse"], st["best_scale"]):
                assert (sc != 0) == (best < base)
            # the marked association is the one after the last applied step
            assert orc.corr_rmse(corrs) < st["base_rmse"][0]
        # marking: every voxel of a valid slot's 3x3x3 neighbourhood counted once
        v = hit[(corrs["valid"] != 0) & (corrs["weight"] > 0)]
        assert (opt[v] == 1).all() and opt.max() == 1 and (optf[opt > 0] == 1).all()
        # the nodes moved
        from cudafusion_amd.harness import transform_errors
        assert transform_errors(T, s["T"])[1] > 1e-4
