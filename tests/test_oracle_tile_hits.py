"""The per-tile first / double mesh hits (ray_firstHit_inTile_kernel / ray_doubleHit_inTile_kernel,
training/src/cuda_func.cu:130-444) in the oracle, pinned by closed forms, and the host face-list
builder (snisr_tile_face_lists) pinned by point sampling:

* a box inside one tile: the first hit is the slab entry t (its face lies in the entry plane), the
  double hit is (entry, exit); rays that miss give face -1 / (-1, -1);
* on a random room mesh the first hit of the first visited tile holding any hit equals the global
  brute-force nearest hit (the oracle's mesh_hit);
* every point of every face lies in a block whose list holds that face."""
import ctypes as C

import numpy as np

from snisr import _native as N
from snisr import scene as S


def _slab(o, d, lo, hi):
    inv = 1.0 / d.astype(np.float64)
    t1, t2 = (lo - o) * inv, (hi - o) * inv
    return np.minimum(t1, t2).max(1), np.maximum(t1, t2).min(1)


def test_box_in_one_tile_closed_forms(oracle_mod):
    lo, hi = np.array([0.31, 0.27, 0.35]), np.array([0.62, 0.71, 0.66])
    box = S.box_mesh(lo[None], hi[None])
    centers = np.array([[0.5, 0.5, 0.5]], np.float32)
    rng = np.random.default_rng(0)
    n = 3000
    o = (np.array([0.5, 0.5, 0.5]) + rng.normal(0, 1, (n, 3)) * 2.0).astype(np.float32)
    tgt = rng.uniform(0.0, 1.0, (n, 3))
    d = tgt - o
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    vis = np.zeros((n, 1), np.int32)
    hits, face = oracle_mod.mesh_hit_tiles(box.vertices, box.faces, centers, 1.0, 8, o, d, vis, mode=0)
    dh = oracle_mod.mesh_hit_tiles(box.vertices, box.faces, centers, 1.0, 8, o, d, vis, mode=1)
    tn, tf = _slab(o, d, lo, hi)
    hit = (tn < tf) & (tn > 1e-4)
    # rays grazing an edge or corner are left out (the float test may go either way there)
    p = o + tn[:, None] * d
    margin = np.minimum(np.abs(p - lo), np.abs(p - hi))
    clear = np.sort(margin, axis=1)[:, 1] > 1e-4
    sel = hit & clear
    assert sel.sum() > 500 and (~hit).sum() > 300
    np.testing.assert_allclose(hits[sel, 0], tn[sel], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dh[sel, 0, 0], tn[sel], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(dh[sel, 0, 1], tf[sel], rtol=1e-5, atol=1e-6)
    # the reported face lies in the entry plane
    V, F = box.vertices.astype(np.float64), box.faces
    fv = V[F[face[sel, 0]]]
    ent_axis = np.argmax(np.abs(p[sel] - np.where(np.abs(p[sel] - lo) < np.abs(p[sel] - hi), lo, hi)) < 1e-5, axis=1)
    assert np.allclose(fv[np.arange(sel.sum()), :, ent_axis].std(axis=1), 0.0)
    miss = ~hit
    assert (face[miss, 0] == -1).all() and np.isnan(hits[miss, 0]).all()
    assert (dh[miss, 0] == -1).all()


def test_first_tile_hit_is_the_global_nearest_hit(oracle_mod):
    mesh = S.room_mesh(3, (0, 0, 0), (2, 2, 1), 600, size_range=(0.05, 0.2))
    grid = S.make_scene(0, tile_shape=(2, 2, 1), res=8, occupancy=0.0)
    centers = grid.tile_centers
    rng = np.random.default_rng(1)
    n = 400
    o = rng.uniform([0.1, 0.1, 0.1], [1.9, 1.9, 0.9], (n, 3)).astype(np.float32)
    d = rng.normal(0, 1, (n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    vis = np.stack([oracle_mod.trace_tiles(oracle_mod.OracleScene(grid), o[i], d[i:i + 1], 4)[0] for i in range(n)])
    hits, face = oracle_mod.mesh_hit_tiles(mesh.vertices, mesh.faces, centers, 1.0, 8, o, d, vis, mode=0)
    t_ref, id_ref = oracle_mod.mesh_hit(mesh.vertices, mesh.faces, o, d)
    checked = 0
    for i in range(n):
        k = np.flatnonzero(face[i] >= 0)
        if k.size == 0 or not np.isfinite(t_ref[i]):
            continue
        j = k[0]
        assert hits[i, j] == t_ref[i] and face[i, j] == id_ref[i], i
        checked += 1
    assert checked > 200


def test_face_lists_cover_every_face_point():
    mesh = S.room_mesh(5, (0, 0, 0), (2, 2, 1), 800, size_range=(0.02, 0.3))
    centers = np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.5, 1.5, 0.5], [1.5, 1.5, 0.5]], np.float32)
    L = N.lib()
    nb = 8
    V = np.ascontiguousarray(mesh.vertices, np.float32)
    F = np.ascontiguousarray(mesh.faces, np.int32)
    fnum = np.empty((4 * nb ** 3, 2), np.int32)
    total = C.c_int64(0)
    args = (V.ctypes.data_as(C.c_void_p), V.shape[0], F.ctypes.data_as(C.c_void_p), F.shape[0],
            centers.ctypes.data_as(C.c_void_p), 4, 1.0, nb, fnum.ctypes.data_as(C.c_void_p))
    assert L.snisr_tile_face_lists(*args, None, 0, C.byref(total)) == 0
    fidx = np.empty(total.value, np.int32)
    assert L.snisr_tile_face_lists(*args, fidx.ctypes.data_as(C.c_void_p), 10, C.byref(total)) == -5
    assert L.snisr_tile_face_lists(*args, fidx.ctypes.data_as(C.c_void_p), fidx.size, C.byref(total)) == 0
    assert fnum[0, 0] == 0 and (fnum[1:, 0] == np.cumsum(fnum[:-1, 1])).all() and fnum[:, 1].sum() == total.value
    rng = np.random.default_rng(2)
    bs = 1.0 / nb
    for f in rng.choice(F.shape[0], 300, replace=False):
        w = rng.dirichlet([1, 1, 1], 40)
        pts = w @ V[F[f]].astype(np.float64)
        for p in pts:
            for t, c in enumerate(centers.astype(np.float64)):
                rel = (p - (c - 0.5)) / bs
                if (rel < 0).any() or (rel >= nb).any():
                    continue
                b = rel.astype(int)
                bi = t * nb ** 3 + (b[0] * nb + b[1]) * nb + b[2]
                s, k = fnum[bi]
                assert f in fidx[s:s + k], (f, t, b)
    # bad vertex indices are refused
    Fb = F.copy()
    Fb[0, 0] = V.shape[0]
    args_b = (V.ctypes.data_as(C.c_void_p), V.shape[0], Fb.ctypes.data_as(C.c_void_p), F.shape[0],
              centers.ctypes.data_as(C.c_void_p), 4, 1.0, nb, fnum.ctypes.data_as(C.c_void_p))
    assert L.snisr_tile_face_lists(*args_b, None, 0, C.byref(total)) == -1
