"""CPU checks of the host BVH builder: layout invariants, and a numpy traversal of the emitted
nodes reproduces the brute-force nearest hit of the oracle."""
import ctypes as C

import numpy as np

from snisr import _native as N
from snisr import scene as S


def build(vertices, faces):
    L = N.lib()
    nb, tb = C.c_int64(0), C.c_int64(0)
    assert L.snisr_bvh_build_bounds(faces.shape[0], C.byref(nb), C.byref(tb)) == 0
    nodes = np.empty(nb.value // 4, np.float32)
    tris = np.empty(max(tb.value, 64) // 4, np.float32)
    nn, depth = C.c_int32(0), C.c_int32(0)
    rc = L.snisr_bvh_build(vertices.ctypes.data_as(C.c_void_p), vertices.shape[0], faces.ctypes.data_as(C.c_void_p),
                           faces.shape[0], nodes.ctypes.data_as(C.c_void_p), nb.value,
                           tris.ctypes.data_as(C.c_void_p), tris.nbytes, C.byref(nn), C.byref(depth))
    assert rc == 0, L.snisr_last_error()
    return nodes[: 16 * nn.value].reshape(-1, 16), tris.reshape(-1, 16)[: faces.shape[0]], depth.value


def leaves(nodes):
    codes = nodes[:, 12:14].copy().view(np.int32)
    return codes


def test_bvh_layout_covers_every_triangle_once():
    m = S.room_mesh(3, (0, 0, 0), (2, 2, 1), 6000)
    nodes, tris, depth = build(m.vertices, m.faces)
    assert depth < 64
    ids = tris[:, 3].copy().view(np.int32)
    assert sorted(ids.tolist()) == list(range(m.faces.shape[0]))
    # leaf-ordered triangles carry their first vertex, the edges from it and the edge-cross normal
    # (v1 - v0) x (v2 - v1), each an FP32 difference of FP32 products as the device test forms it
    v = m.vertices[m.faces[ids]].astype(np.float32)
    e1, e2, f = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0], v[:, 2] - v[:, 1]
    n = np.stack([e1[:, 1] * f[:, 2] - e1[:, 2] * f[:, 1], e1[:, 2] * f[:, 0] - e1[:, 0] * f[:, 2],
                  e1[:, 0] * f[:, 1] - e1[:, 1] * f[:, 0]], 1)
    assert np.array_equal(tris[:, 0:3], v[:, 0]) and np.array_equal(tris[:, 4:7], e1)
    assert np.array_equal(tris[:, 8:11], e2) and np.array_equal(tris[:, 12:15], n)
    assert not np.any(tris[:, [7, 11, 15]].copy().view(np.int32))
    # walk the tree from the root: every triangle index is reached exactly once, child boxes contain them
    seen = np.zeros(m.faces.shape[0], np.int32)
    codes = leaves(nodes)
    stack = [0]
    while stack:
        n = stack.pop()
        for k in range(2):
            c = int(codes[n, k])
            box = nodes[n, 6 * k:6 * k + 6]
            if c >= 0:
                stack.append(c)
                sub = [c]
                # child box contains the grandchildren's boxes
                for kk in range(2):
                    gb = nodes[c, 6 * kk:6 * kk + 6]
                    if codes[c, kk] >= 0 or (~codes[c, kk]) & 7:
                        assert np.all(gb[:3] >= box[:3]) and np.all(gb[3:] <= box[3:])
            else:
                first, cnt = (~c) >> 3, (~c) & 7
                assert cnt <= 4
                seen[first:first + cnt] += 1
                for j in range(first, first + cnt):
                    pts = m.vertices[m.faces[int(ids[j])]]
                    assert np.all(pts >= box[:3]) and np.all(pts <= box[3:])
    assert np.all(seen == 1)


def _slab(box, o, inv):
    t0 = (box[:3] - o) * inv
    t1 = (box[3:] - o) * inv
    return np.max(np.minimum(t0, t1)), np.min(np.maximum(t0, t1))


def test_numpy_traversal_matches_oracle_brute_force(oracle_mod):
    m = S.room_mesh(4, (0, 0, 0), (2, 2, 1), 3000)
    nodes, tris, _ = build(m.vertices, m.faces)
    codes = leaves(nodes)
    rng = np.random.default_rng(0)
    n = 60
    origins = rng.uniform([0.2, 0.2, 0.2], [1.8, 1.8, 0.8], (n, 3)).astype(np.float32)
    dirs = oracle_mod.normalize(rng.standard_normal((n, 3)).astype(np.float32))
    t_ref, id_ref = oracle_mod.mesh_hit(m.vertices, m.faces, origins, dirs)
    V = m.vertices.astype(np.float64)
    for i in range(n):
        o, d = origins[i].astype(np.float64), dirs[i].astype(np.float64)
        inv = np.where(d != 0, 1 / np.where(d != 0, d, 1), np.copysign(1e30, d))
        best, bid = np.inf, -1
        stack = [0]
        while stack:
            nd = stack.pop()
            for k in range(2):
                tn, tf = _slab(nodes[nd, 6 * k:6 * k + 6].astype(np.float64), o, inv)
                if not (tn <= tf and tf >= 0 and tn <= best + 1e-9):
                    continue
                c = int(codes[nd, k])
                if c >= 0:
                    stack.append(c)
                    continue
                first, cnt = (~c) >> 3, (~c) & 7
                for j in range(first, first + cnt):
                    fid = int(tris[j, 3:4].copy().view(np.int32)[0])
                    tt, _ = oracle_mod.mesh_hit(m.vertices, m.faces[fid:fid + 1], origins[i:i + 1], dirs[i:i + 1])
                    if tt[0] > 0 and (tt[0] < best or (tt[0] == best and fid < bid)):
                        best, bid = float(tt[0]), fid
        assert bid == id_ref[i] and (best == t_ref[i] or (bid == -1 and np.isinf(t_ref[i])))
