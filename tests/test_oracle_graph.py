"""Pins of the oracle's graph maintenance (oracle/graph.cpp, restating
src/warp_field.cu:407-472, 488-694, 709-954, 986-1104) by closed forms, and
the host part of the library (nrf_graph_add_nodes, no device needed)
against it:

* greedy sampling of a line of points 0.3 min_dist apart keeps every 4th
  point (the first at or beyond min_dist of the last node), so nodes are
  1.2 min_dist apart; each new node's graph neighbours are the earlier nodes
  within 2.5 min_dist, nearest first; orthogonal normals cut the edges;
  every input point lies within min_dist of a node;
* a node added next to moved nodes starts from their blended transform;
* geodesic distances on a triangulated planar grid (unit cells, one
  diagonal per cell) are graph shortest paths, sqrt(2) min(i, j) + |i - j|
  cells between grid vertices; two parallel sheets that no triangle joins
  get no edges across, though they are within the Euclidean radius;
* pruning drops a node in observed empty space and a component below the
  minimum size, keeps supported ones, and remaps the neighbour lists.
"""
import ctypes as C

import numpy as np
import pytest

from cudafusion_amd import _lib as L


def _buf(n):
    nodes = np.zeros(n, L.NODE)
    T = np.zeros(n, L.DUALQUAT)
    return nodes, T


def test_line_sampling_spacing_and_edges(orc):
    md = 0.02
    pts = np.zeros((40, 3), np.float32)
    pts[:, 0] = np.arange(40, dtype=np.float32) * np.float32(0.3 * md)
    nodes, T = _buf(64)
    added, n, _, _ = orc.graph_add_nodes(nodes, T, 0, pts, None, md, 0.01)
    assert added == n == 10
    np.testing.assert_array_equal(nodes["pos"][:n, 0], pts[::4, 0])
    for i in range(n):
        exp = [j for j in (i - 1, i - 2) if j >= 0]
        assert nodes["num_neighbors"][i] == len(exp)
        assert list(nodes["neighbors"][i][: len(exp)]) == exp
        assert (nodes["neighbors"][i][len(exp):] == -1).all()
    # covering: every point within min_dist of a node
    d = np.abs(pts[:, None, 0] - nodes["pos"][None, :n, 0]).min(1)
    assert (d < md).all()
    assert (T["real"][:n, 3] == 1).all()
    # orthogonal normals on alternate points: no edges between them
    nrm = np.zeros_like(pts)
    nrm[:, 2] = 1.0
    nrm[4::8] = [0.0, 1.0, 0.0]
    nodes2, T2 = _buf(64)
    orc.graph_add_nodes(nodes2, T2, 0, pts, nrm, md, 0.01)
    for i in range(1, 10):
        for k in range(nodes2["num_neighbors"][i]):
            j = nodes2["neighbors"][i][k]
            assert np.dot(nodes2["normal"][i], nodes2["normal"][j]) >= 0.7


def test_new_node_transform_blend(orc):
    md = 0.02
    nodes, T = _buf(8)
    nodes["pos"][0] = (0, 0, 0)
    nodes["radius"][0] = 0.01
    T["real"][0] = (0, 0, np.sin(0.1), np.cos(0.1))   # rotation about z
    T["dual"][0] = (0.01, 0, 0, 0)
    added, n, _, _ = orc.graph_add_nodes(nodes, T, 1, np.array([[0.03, 0, 0]], np.float32), None, md, 0.01)
    assert added == 1 and n == 2
    assert nodes["neighbors"][1][0] == 0 and nodes["num_neighbors"][1] == 1
    np.testing.assert_allclose(T["real"][1], T["real"][0], atol=1e-7)
    np.testing.assert_allclose(T["dual"][1], T["dual"][0], atol=1e-7)


def _grid_mesh(nx, ny, h, z=0.0):
    """Triangle soup of a planar grid: cells (i, j) split along the (1, 1)
    diagonal, vertex (i, j) at (i h, j h, z)."""
    tri = []
    for j in range(ny - 1):
        for i in range(nx - 1):
            a, b, c, d = (i, j), (i + 1, j), (i + 1, j + 1), (i, j + 1)
            tri += [(a, b, c), (a, c, d)]
    v = np.array([[p[0] * h, p[1] * h, z] for t in tri for p in t], np.float32)
    return v, np.arange(len(v), dtype=np.int32).reshape(-1, 3)


def test_geodesic_distances_on_a_grid(orc):
    h = 0.005
    v, t = _grid_mesh(30, 30, h)
    nodes, T = _buf(256)
    md = 0.02
    added, n, gi, gd = orc.graph_add_nodes(nodes, T, 0, v, None, md, 0.03, tris=t)
    assert added == n > 10
    pos = nodes["pos"][:n]
    max_dist = max(3.5 * md, 1.5 * 0.03)
    checked = 0
    for s in range(n):
        ij = np.rint(pos[s, :2] / h).astype(int)
        for k in range(8):
            j = gi[s, k]
            if j < 0:
                continue
            off = np.rint(pos[j, :2] / h).astype(int) - ij
            d = np.abs(off)
            if off[0] * off[1] >= 0:
                # both offsets of one sign: the (1, 1) diagonals shorten the path
                ref = (np.sqrt(2) * d.min() + abs(d[0] - d[1])) * h
            else:
                ref = (d[0] + d[1]) * h
            assert gd[s, k] == pytest.approx(ref, rel=1e-5), (s, j)
            assert 0 < gd[s, k] <= max_dist
            checked += 1
        # neighbours sorted by distance, at most 8
        dd = gd[s][gi[s] >= 0]
        assert (np.diff(dd) >= 0).all()
    assert checked > 3 * n
    # edges follow the geodesic neighbours (same normals: all kept), weights exp(-d^2 / 2 r^2)
    for s in range(n):
        m = nodes["num_neighbors"][s]
        np.testing.assert_array_equal(nodes["neighbors"][s][:m], gi[s][:m])
        np.testing.assert_allclose(nodes["neighbor_w"][s][:m], np.exp(-(gd[s][:m] ** 2) / (2 * 0.03 ** 2)),
                                   rtol=1e-6)


def test_geodesic_separates_unjoined_sheets(orc):
    h = 0.005
    v1, t1 = _grid_mesh(20, 20, h, 0.0)
    v2, t2 = _grid_mesh(20, 20, h, 0.015)
    v = np.concatenate([v1, v2])
    t = np.concatenate([t1, t2 + len(v1)])
    nodes, T = _buf(256)
    md = 0.02
    _, n, gi, _ = orc.graph_add_nodes(nodes, T, 0, v, None, md, 0.03, tris=t)
    z = nodes["pos"][:n, 2]
    # the Euclidean sampling graph joins the sheets (1.5 cm < 2.5 min_dist) ...
    nodes_e, T_e = _buf(256)
    _, ne, _, _ = orc.graph_add_nodes(nodes_e, T_e, 0, v, None, md, 0.03)
    cross_e = sum(abs(nodes_e["pos"][i, 2] - nodes_e["pos"][nodes_e["neighbors"][i][k], 2]) > 0.01
                  for i in range(ne) for k in range(nodes_e["num_neighbors"][i]))
    assert cross_e > 0
    # ... the geodesic graph does not
    for s in range(n):
        for k in range(nodes["num_neighbors"][s]):
            assert abs(z[s] - z[nodes["neighbors"][s][k]]) < 1e-6


def test_prune_empty_and_small_components(orc):
    dims = (20, 20, 20)
    s = 0.01
    vp = L.volume_params(dims, s, 0.03, (0, 0, 0), 1.0, 50.0)
    vox = np.zeros(int(np.prod(dims)), L.VOXEL)
    vox["weight"] = 1.0
    vox["tsdf"] = 1.0                     # observed free space everywhere ...
    zc = (np.arange(dims[2]) + 0.5) * s
    surf = np.abs(zc - 0.1) < 0.02        # ... but a surface band around z = 0.1
    vox["tsdf"].reshape(dims[::-1])[surf] = 0.1
    nodes, T = _buf(8)
    P = [(0.05, 0.05, 0.1), (0.07, 0.05, 0.1), (0.09, 0.05, 0.1), (0.11, 0.05, 0.1), (0.13, 0.05, 0.1),
         (0.10, 0.15, 0.16), (0.05, 0.15, 0.10), (0.06, 0.16, 0.10)]
    for i, p in enumerate(P):
        nodes["pos"][i] = p
        nodes["radius"][i] = 0.02
        nodes["neighbors"][i] = -1
    chain = [(0, 1), (1, 2), (2, 3), (3, 4)]
    for a, b in chain + [(6, 7)]:
        k = nodes["num_neighbors"][a]
        nodes["neighbors"][a][k] = b
        nodes["neighbor_w"][a][k] = 0.5
        nodes["num_neighbors"][a] = k + 1
    T["dual"][:, 0] = np.arange(8)
    pp = L.prune_params(min_component_size=3)
    removed, n, keep = orc.graph_prune(vp, vox, nodes, T, 8, pp)
    # node 5 sits in observed free space, nodes 6-7 form a component of 2
    assert list(keep) == [1, 1, 1, 1, 1, 0, 1, 1]
    assert removed == 3 and n == 5
    np.testing.assert_array_equal(T["dual"][:5, 0], [0, 1, 2, 3, 4])
    assert list(nodes["neighbors"][0][:1]) == [1] and nodes["num_neighbors"][4] == 0


def test_library_sampling_matches_oracle(orc):
    """nrf_graph_add_nodes (host code of the library) against the oracle on
    the sphere scene's surface samples: identical nodes, edges, transforms."""
    from cudafusion_amd import scene as S
    lib = L.load()
    sc = S.make_scene("sphere", frames=1)
    p, nrm = sc.surface_samples(sc.w.node_spacing / 4)
    p = np.ascontiguousarray(p, np.float32)
    nrm = np.ascontiguousarray(nrm, np.float32)
    a_nodes, a_T = _buf(4096)
    b_nodes, b_T = _buf(4096)
    na = C.c_int(0)
    added = C.c_int(0)
    L.check(lib.nrf_graph_add_nodes(a_nodes.ctypes.data, a_T.ctypes.data, C.byref(na), 4096, p.ctypes.data,
                                    nrm.ctypes.data, len(p), sc.w.node_spacing, sc.w.node_spacing,
                                    C.byref(added)), "add_nodes")
    _, nb, _, _ = orc.graph_add_nodes(b_nodes, b_T, 0, p, nrm, sc.w.node_spacing, sc.w.node_spacing)
    assert na.value == nb > 100 and added.value == nb
    assert a_nodes[:nb].tobytes() == b_nodes[:nb].tobytes()
    assert a_T[:nb].tobytes() == b_T[:nb].tobytes()
