"""The oracle's subset form of the FRender frame (frender_subset) against the same oracle chain run
over the whole frame, with torch's bilinear x2 upsampling (align_corners=True) as the library
routine for the reflection layer: a subset of pixels must reproduce the full-frame chain."""
import numpy as np
import torch
import torch.nn.functional as F

from snisr import scene as S


def _grid_scene(osc):
    sc = S.make_scene(0, tile_shape=osc.tile_shape, tile_size=osc.tile_size, scene_min=tuple(osc.scene_min), res=8,
                      occupancy=0.0)
    sc.tile_index_map = osc.tile_index_map.copy()
    return sc


def test_frender_subset_equals_full_frame_chain(oracle_mod):
    H, W = 48, 80
    osc = S.make_octree_scene(9, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=6)
    params, centers = S.make_reflection_groups(19, 6, (0, 0, 0), (4, 4, 1))
    cam = S.room_camera(9, (0, 0, 0), (4, 4, 1), H, W, focal=48.0)
    far = float(np.linalg.norm(np.asarray(osc.tile_shape, np.float64) * osc.tile_size))
    O = oracle_mod
    dirs = O.compute_rays(H, W, cam.intrinsic, cam.c2w)
    grid = O.OracleScene(_grid_scene(osc))
    vis = O.trace_tiles(grid, cam.origin, dirs, 30)
    diff, inv, net, _ = O.render_octree(osc, cam.origin, dirs, vis, trans_th=0.8, early_stop=0.01)
    DH, DW = H // 2, W // 2
    inv_near = inv.reshape(H, W)[::2, ::2].copy().ravel()
    net_lo = net.reshape(H, W)[::2, ::2].ravel().copy()
    qpi, bnet, bps = S.reflection_blocks(net_lo)
    sample_far = (np.float32(1.0) / inv_near + np.float32(far)).astype(np.float32)
    inv_bound = ((np.float32(1.0) / sample_far - inv_near) / np.float32(63)).astype(np.float32)
    d_lo = np.ascontiguousarray(dirs.reshape(H, W, 3)[::2, ::2])
    T = np.ones(DH * DW, np.float32)
    col = np.zeros((DH * DW, 3), np.float32)
    for step in range(0, 64, 16):
        pts, st = O.inverse_z_sampling(cam.origin, d_lo.reshape(-1, 3), inv_near, inv_bound, step, 16)
        res = np.zeros((DH * DW * 16, 4), np.float32)
        O.inference_mlp(16, bnet, bps, qpi, T, centers, pts, d_lo, params, 1.0 / far, 0.01, res)
        O.integrate_points(res, st, np.float32(0.01), T, col)
    up = F.interpolate(torch.from_numpy(col.reshape(1, DH, DW, 3)).permute(0, 3, 1, 2), scale_factor=2,
                       mode="bilinear", align_corners=True)[0].permute(1, 2, 0).numpy().reshape(-1, 3)
    frame = np.clip(diff.astype(np.float64) + up, 0, 1)
    assert (net_lo >= 0).mean() > 0.3 and np.abs(col).max() > 1e-2   # the chain is not trivial
    hi = np.sort(np.random.default_rng(3).choice(H * W, 300, replace=False))
    hi = np.concatenate([hi, [0, W - 1, (H - 1) * W, H * W - 1]])       # image corners
    sub = O.frender_subset(osc, grid, params, centers, cam.origin, dirs, H, W, far, hi)
    assert np.array_equal(sub["diffuse"], diff[hi])
    assert np.array_equal(sub["spec_lo"], col[sub["lo_idx"]])
    assert np.abs(sub["frame"] - frame[hi]).max() <= 1e-6
