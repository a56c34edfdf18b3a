"""The oracle's restatement of the training side's dense FP32 marcher (render_diffuse_kernel,
training/src/cuda_func.cu:1687-1802) pinned without the reference: the OVERLAP interpolation
equals torch's grid_sample on the dilated grid (align_corners=True; zeros padding inside the tile,
border padding where the reference degrades to lower-order interpolation at the upper grid edge),
and whole renders of constant fields match the closed forms T = exp(-sigma L), rgb = c (1 - T), the
first-hit depth, the group abort and unflagged cells."""
import math

import numpy as np
import torch
import torch.nn.functional as F

from snisr import scene as S


def _grid_sample(vox, idx, mode):
    """grid_sample of vox f32[nv, nv, nv, 4] (x, y, z, c) at fractional cell indices idx f32[M, 3]."""
    nv = vox.shape[0]
    inp = torch.from_numpy(np.ascontiguousarray(vox.transpose(3, 2, 1, 0)))[None].double()
    g = torch.from_numpy(idx.astype(np.float64) / (nv - 1) * 2 - 1)[None, None, None]
    out = F.grid_sample(inp, g, mode="bilinear", padding_mode=mode, align_corners=True)
    return out[0, :, 0, 0].T.numpy()


def test_overlap_interpolation_is_grid_sample(oracle_mod):
    ds = S.make_dense_scene(3, tile_shape=(1, 1, 1), num_voxel=16, block=4, block_occupancy=0.7)
    vox, c, ts, vs = ds.voxels[0], ds.centers[0], ds.tile_size, ds.voxel_size
    rng = np.random.default_rng(0)
    p = (c + rng.uniform(-0.5, 0.5, (4000, 3)) * ts).astype(np.float32)
    got = oracle_mod.dense_overlap(vox, c, ts, vs, p)
    idx = (p.astype(np.float64) - (c - ts / 2)) / vs + 0.5
    ref = _grid_sample(vox, idx, "zeros")
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-5 * np.abs(vox).max())
    # beyond the upper grid edge the reference interpolates only along the axes still inside: the
    # value of the edge plane, i.e. grid_sample with border padding
    nv = vox.shape[0]
    q = rng.uniform(0.5, nv - 1.0, (2000, 3))
    ax = rng.integers(0, 3, 2000)
    q[np.arange(2000), ax] = rng.uniform(nv - 1.0, nv - 0.01, 2000)
    pts = ((q - 0.5) * vs + (c - ts / 2)).astype(np.float32)
    got = oracle_mod.dense_overlap(vox, c, ts, vs, pts)
    qq = (pts.astype(np.float64) - (c - ts / 2)) / vs + 0.5
    ref = _grid_sample(vox, np.minimum(qq, nv - 1), "border")
    np.testing.assert_allclose(got, ref, rtol=2e-5, atol=2e-5 * np.abs(vox).max())


def _constant_tile(nv, rgb, sigma, flag=1):
    vox = np.zeros((1, nv, nv, nv, 4), np.float32)
    vox[..., :3] = rgb
    vox[..., 3] = sigma
    return vox, np.full((1, nv, nv, nv), flag, np.int16)


def test_constant_field_closed_forms(oracle_mod):
    nv, ts = 18, 1.0
    vs = ts / (nv - 2)
    step = 0.2 * vs
    sigma = 2.5
    c = np.array([0.3, 0.6, 0.9], np.float32)
    vox, nodes = _constant_tile(nv, c, sigma)
    center = np.array([[0.5, 0.5, 0.5]], np.float32)
    rng = np.random.default_rng(1)
    n = 500
    o = (np.array([0.5, 0.5, 0.5]) + rng.normal(0, 1, (n, 3)) * 3).astype(np.float32)
    tgt = rng.uniform(0.2, 0.8, (n, 3))
    d = (tgt - o)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    vis = np.zeros((n, 2), np.int32)
    vis[:, 1] = -1
    dif, dep, val = oracle_mod.render_dense(vox, nodes, np.array([0], np.int32), center, ts, vs, step, o, d, vis)
    # chord of the unit cube (float64 slab), T = exp(-sigma L), rgb = c (1 - T)
    inv = 1.0 / d.astype(np.float64)
    t1, t2 = (0.0 - o) * inv, (1.0 - o) * inv
    near = np.maximum(np.minimum(t1, t2).max(1), 0.0)
    far = np.maximum(t1, t2).min(1)
    T = np.exp(-sigma * (far - near))
    np.testing.assert_allclose(dif, c[None] * (1 - T)[:, None], rtol=2e-4, atol=1e-6)
    # first sample at or after the crossing of T = 0.5; invalid (depth -1) when the chord is too short
    k = np.ceil(math.log(2) / (sigma * step) - 1e-9)
    reach = near + k * step < far
    assert np.array_equal(val == 1, reach)
    np.testing.assert_allclose(dep[reach], (near + k * step)[reach], atol=2e-5)
    assert (dep[~reach] == -1).all()


def test_group_abort_and_unflagged_cells(oracle_mod):
    nv, ts = 10, 1.0
    vs = ts / (nv - 2)
    vox, nodes = _constant_tile(nv, [1.0, 1.0, 1.0], 50.0)
    centers = np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5]], np.float32)
    o = np.array([[-1.0, 0.5, 0.5]] * 3, np.float32)
    d = np.array([[1.0, 0.0, 0.0]] * 3, np.float32)
    vis = np.array([[1, 0], [0, 1], [0, -1]], np.int32)
    # tile 1 is outside the group: a ray whose first visited tile is tile 1 stops there, invalid
    dif, dep, val = oracle_mod.render_dense(vox, nodes, np.array([0, -1], np.int32), centers, ts, vs, 0.02, o, d, vis)
    assert val.tolist() == [0, 1, 1] and dep[0] == -1 and (dif[0] == 0).all()
    assert dif[1].min() > 0.99 and abs(dep[1] - 1.0) < 0.05
    # every cell flagged 0: nothing is sampled
    dif, dep, val = oracle_mod.render_dense(vox, np.zeros_like(nodes), np.array([0, -1], np.int32), centers, ts, vs,
                                            0.02, o, d, vis)
    assert (dif == 0).all() and (val == 0).all() and (dep == -1).all()
