"""GPU parity of the dense FP32 marcher (render_diffuse_kernel, training/src/cuda_func.cu:1687-1802)
with the oracle: diffuse colours, depths and valid flags bit-exact, on a multi-tile scene in the
training tiles' format with dropped tiles, a tile outside the group, camera rays from inside and
outside the scene and axis-parallel rays."""
import numpy as np
import pytest
import torch

from snisr import dense
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _grid(ds):
    sc = S.make_scene(0, tile_shape=ds.tile_shape, tile_size=ds.tile_size, scene_min=tuple(ds.scene_min), res=8,
                      occupancy=0.0)
    sc.tile_index_map = ds.tile_index_map.copy()
    return sc


@pytest.mark.parametrize("eye,seed", [((-0.7, -0.4, 0.6), 1), ((1.3, 0.9, 0.5), 2)])
def test_render_dense_bit_exact(oracle_mod, eye, seed):
    ds = S.make_dense_scene(seed, tile_shape=(3, 2, 1), num_voxel=32, block=4, block_occupancy=0.3, drop_tiles=1)
    T = ds.tile_ids.size
    s2g = np.arange(T, dtype=np.int32)
    s2g[T - 1] = -1                                   # one tile outside the group
    H, W = 48, 64
    cam = S.look_at_camera(np.asarray(eye, np.float64), np.array([1.5, 1.0, 0.5]), H, W, 40.0)
    dirs = S.camera_rays_np(cam)
    dirs[:40] = np.array([1.0, 0.0, 0.0], np.float32)  # axis-parallel rays
    dirs = (dirs / np.linalg.norm(dirs, axis=1, keepdims=True)).astype(np.float32)
    vis = oracle_mod.trace_tiles(oracle_mod.OracleScene(_grid(ds)), cam.origin, dirs, 8)
    n = dirs.shape[0]
    o = np.repeat(cam.origin[None].astype(np.float32), n, 0)
    step = 0.2 * ds.voxel_size
    ref = oracle_mod.render_dense(ds.voxels, ds.nodes, s2g, ds.centers, ds.tile_size, ds.voxel_size, step, o, dirs,
                                  vis)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    dif = torch.empty((n, 3), dtype=torch.float32, device=DEV)
    dep = torch.empty(n, dtype=torch.float32, device=DEV)
    val = torch.empty(n, dtype=torch.int32, device=DEV)
    dense.render_diffuse(t(o), t(dirs), t(vis), t(s2g), ds.tile_size, ds.num_voxel + 2, ds.voxel_size, step,
                         t(ds.centers), t(ds.voxels), t(ds.nodes), dif, dep, val)
    g = dif.cpu().numpy(), dep.cpu().numpy(), val.cpu().numpy()
    assert np.array_equal(g[0].view(np.uint32), ref[0].view(np.uint32))
    assert np.array_equal(g[1].view(np.uint32), ref[1].view(np.uint32))
    assert np.array_equal(g[2], ref[2])
    assert 0.2 < ref[2].mean() < 0.95 and (ref[0] > 0).any()     # hits and misses both present
