"""Formats and IO (cudafusion_amd/formats.py): the ASCII PLY of
include/fusion_app.h:901-943, the depth-sequence loaders of
include/depth_sequence.h:106-270 and the YAML keys of
include/yaml_helper.h:39-193.

The PNG decoder is pinned by round trips through an encoder written here
from the PNG specification (per-pixel scalar filters, every filter type on
some row, 8- and 16-bit grey / grey-alpha / RGB / RGBA); the loaders by
synthetic sequences whose frame selection (comments, start_frame,
max_frames, a missing depth file) and depth scaling are known; the YAML
loader by the reference's own configuration files (tests/golden/config_*.yaml,
copied as data) and a required key removed.
"""
import os
import struct
import zlib

import numpy as np
import pytest

from cudafusion_amd import formats as F


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _encode_png(path, img, depth):
    img = np.asarray(img)
    h, w = img.shape[:2]
    ch = 1 if img.ndim == 2 else img.shape[2]
    ctype = {1: 0, 2: 4, 3: 2, 4: 6}[ch]
    bpp = ch * depth // 8
    if depth == 16:
        rows = img.astype(">u2").reshape(h, -1).view(np.uint8).reshape(h, w * bpp)
    else:
        rows = img.astype(np.uint8).reshape(h, w * bpp)
    out = bytearray()
    prev = [0] * (w * bpp)
    for y in range(h):
        cur = [int(v) for v in rows[y]]
        ft = y % 5
        line = [ft]
        for i, x in enumerate(cur):
            a = cur[i - bpp] if i >= bpp else 0
            b = prev[i]
            c = prev[i - bpp] if i >= bpp else 0
            pred = [0, a, b, (a + b) // 2, _paeth(a, b, c)][ft]
            line.append((x - pred) & 255)
        out += bytes(line)
        prev = cur

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xffffffff)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, depth, ctype, 0, 0, 0))
                + chunk(b"IDAT", zlib.compress(bytes(out))) + chunk(b"IEND", b""))


@pytest.mark.parametrize("shape,depth", [((13, 17), 8), ((11, 9), 16), ((7, 10, 2), 8), ((9, 6, 3), 8),
                                         ((8, 5, 4), 16)])
def test_png_round_trip_every_filter(tmp_path, shape, depth):
    rng = np.random.default_rng(3)
    img = rng.integers(0, 1 << depth, shape).astype(np.uint16 if depth == 16 else np.uint8)
    img.reshape(-1)[:5] = [0, (1 << depth) - 1, 1, 2, 3]
    p = str(tmp_path / "a.png")
    _encode_png(p, img, depth)
    got = F.read_png(p)
    assert got.dtype == img.dtype and got.shape == img.shape
    np.testing.assert_array_equal(got, img)


def test_ply_text_and_round_trip(tmp_path):
    v = np.array([[0.1, -2.5, 3.0], [1e-7, 123456.7, -0.333333333], [0.5, 0.25, 0.125]], np.float32)
    n = np.array([[0, 0, 1]] * 3, np.float32)
    t = np.array([[0, 1, 2]], np.int32)
    txt = F.ply_text(v, n, t)
    lines = txt.splitlines()
    assert lines[:10] == ["ply", "format ascii 1.0", "element vertex 3", "property float x", "property float y",
                          "property float z", "property float nx", "property float ny", "property float nz",
                          "element face 1"]
    assert lines[10:12] == ["property list uchar int vertex_indices", "end_header"]
    # std::ostream << float: 6 significant digits
    assert lines[12] == "0.1 -2.5 3 0 0 1"
    assert lines[13] == "1e-07 123457 -0.333333 0 0 1"
    assert lines[15] == "3 0 1 2"
    p = str(tmp_path / "m.ply")
    F.save_mesh_ply(p, v, n, t)
    rv, rn, rt = F.read_ply(p)
    np.testing.assert_allclose(rv, v, rtol=1e-5)
    np.testing.assert_array_equal(rn, n)
    np.testing.assert_array_equal(rt, t)


def test_mesh_normals_of_a_soup():
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 0], [0, 0, 1], [1, 0, 0]], np.float32)
    n = F.mesh_normals(v)
    np.testing.assert_allclose(n[:3], [[0, 0, 1]] * 3)
    np.testing.assert_allclose(n[3:], [[0, 1, 0]] * 3)


def _depth_png(path, mm):
    _encode_png(path, np.asarray(mm, np.uint16), 16)


def test_tum_associations(tmp_path):
    d = tmp_path
    (d / "depth").mkdir()
    (d / "rgb").mkdir()
    lines = ["# comment", ""]
    for i in range(5):
        if i != 2:
            _depth_png(str(d / "depth" / f"{i}.png"), np.full((4, 6), 1000 + i, np.uint16))
        _encode_png(str(d / "rgb" / f"{i}.png"), np.full((4, 6, 3), [10 * i, 20, 30], np.uint8), 8)
        lines.append(f"{i}.5 rgb/{i}.png {i}.25 depth/{i}.png")
    (d / "associations.txt").write_text("\n".join(lines) + "\n")
    seq = F.DepthSequence(F.SequenceConfig(path=str(d), format=F.Format.TUM, depth_scale=0.001, start_frame=1,
                                           max_frames=3))
    # line 0 skipped by start_frame, line 2 has no depth file (counted, not kept)
    assert [f.index for f in seq.frames] == [1, 3, 4]
    assert [f.timestamp for f in seq.frames] == [1.25, 3.25, 4.25]
    np.testing.assert_array_equal(seq.frames[1].depth_m, np.float32(1003 * float(np.float32(0.001))))
    assert seq.frames[0].depth_m.dtype == np.float32 and seq.frames[0].color_gray.shape == (4, 6)
    assert int(seq.frames[0].color_gray[0, 0]) == int(np.rint(0.299 * 10 + 0.587 * 20 + 0.114 * 30))
    assert seq.total() == 3 and seq.next().index == 1 and seq.has_next()


def test_tum_folders_and_raw_directory(tmp_path):
    d = tmp_path / "tum"
    (d / "depth").mkdir(parents=True)
    (d / "color").mkdir()
    for i in range(3):
        _depth_png(str(d / "depth" / f"{i:03d}.png"), np.full((3, 4), 500 * (i + 1), np.uint16))
        _encode_png(str(d / "color" / f"{i:03d}.png"), np.full((6, 8), 7 * i, np.uint8), 8)  # resized to 3 x 4
    seq = F.DepthSequence(F.SequenceConfig(path=str(d), format=F.Format.TUM, depth_scale=0.002))
    assert [f.index for f in seq.frames] == [0, 1, 2] and seq.frames[2].color_gray.shape == (3, 4)
    assert float(seq.frames[1].depth_m[0, 0]) == pytest.approx(2.0)
    r = tmp_path / "raw"
    r.mkdir()
    _depth_png(str(r / "a.png"), np.array([[50, 150, 2000], [5999, 6001, 7000]], np.uint16))
    seq = F.DepthSequence(F.SequenceConfig(path=str(r), format=F.Format.RAW_PNG, depth_scale=0.001))
    dm = seq.frames[0].depth_m
    # the raw loader keeps 0.1 m < depth <= 6 m
    np.testing.assert_array_equal(dm > 0, [[False, True, True], [True, False, False]])
    g = seq.frames[0].color_gray
    assert g[0, 1] == 255 and g[1, 0] == 0 and g[0, 0] == 0
    with pytest.raises(NotImplementedError):
        F.DepthSequence(F.SequenceConfig(path=str(r), format=F.Format.RAW_EXR))


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_yaml_reference_configs():
    app = F.load_config(os.path.join(GOLDEN, "config_params.yaml"))
    df = app.df
    assert app.seq.format == F.Format.TUM and app.seq.depth_scale == pytest.approx(0.001)
    assert app.seq.max_frames == -1 and df.cam.fx == pytest.approx(570.342) and df.cam.width == 640
    assert df.tsdf.dims == (256, 256, 256) and df.tsdf.origin == (-0.7, -0.7, 0.5)
    assert df.tsdf.decay_alpha == pytest.approx(0.8) and df.tsdf.voxel_size == pytest.approx(0.005)
    assert (df.solver.gn_iterations, df.solver.pcg_iterations) == (2, 80)
    assert df.solver.lambda_smooth == 100.0 and df.solver.lambda_damping == 1.0
    assert df.solver.pcg_tolerance == pytest.approx(1e-4)  # not a YAML key: the solver's default
    assert df.max_nodes == 2048 and df.prune_min_component_size == 16 and df.update_scale == 0.0
    assert df.integrate_warped and df.integrate_min_optimized_count == 1
    assert (df.dist_threshold, df.angle_threshold, df.view_threshold, df.search_radius_px,
            df.min_valid_corrs) == (0.2, 0.0, 0.0, 50, 5000)
    assert df.rigid_tracking  # not in the file: the default
    assert df.use_sift and df.sift_weight == pytest.approx(0.4) and df.sift_max_history == 12
    assert df.profile_timings and not app.quiet and app.use_vis and not df.bbox.enabled
    cfg, tp = F.pipeline_config(app)
    assert tuple(cfg.volume.dims) == (256, 256, 256) and cfg.gn.gn_iterations == 2
    assert cfg.min_opt_count == 1 and tp.min_valid_corrs == 5000 and tp.assoc.search_radius_px == 50
    for name in ("config_upperbody.yaml", "config_boxing.yaml"):
        a = F.load_config(os.path.join(GOLDEN, name))
        assert a.df.tsdf.dims[0] > 0 and a.df.cam.height > 0


def test_yaml_missing_required_key(tmp_path):
    import yaml
    with open(os.path.join(GOLDEN, "config_params.yaml")) as f:
        cfg = yaml.safe_load(f)
    del cfg["icp"]["dist_threshold"]
    p = tmp_path / "c.yaml"
    p.write_text(yaml.safe_dump(cfg))
    with pytest.raises(KeyError):
        F.load_config(str(p))
    cfg = yaml.safe_load(open(os.path.join(GOLDEN, "config_params.yaml")))
    del cfg["tsdf"]["decay_alpha"]  # optional: the TSDFVolume::Params default
    p.write_text(yaml.safe_dump(cfg))
    assert F.load_config(str(p)).df.tsdf.decay_alpha == pytest.approx(0.9)
