"""GPU parity of the reference-format octree renderer (rendering_diffuse_octree_fp16) against the
FP32 oracle: per-ray sample counts, first-hit inverse depth and group bit-exact, pixels within max
abs error 1e-4, through the FASTRENDERING stand-in exactly as FRender.render calls it
(rendering/src/__init__.py:247-268)."""
import numpy as np
import pytest
import torch

from snisr import fastrendering as FR
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
RGB_TOL = 1e-4


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return t if dtype is None else t.to(dtype)


def render_both(oracle_mod, osc, cam, K=30, soft_hit=False, trans_th=0.8, early_stop=0.01):
    H, W = cam.H, cam.W
    rays_d = torch.full((H, W, 3), -1.0, device=DEV)
    FR.compute_rays(to_dev(cam.intrinsic), to_dev(cam.c2w), 256, rays_d)
    vis = torch.full((H, W, K), -1, device=DEV, dtype=torch.int32)
    scene_size = torch.tensor(np.asarray(osc.tile_shape, np.float32) * np.float32(osc.tile_size))
    FR.tracing_tiles(rays_d, to_dev(osc.tile_index_map), to_dev(cam.origin), torch.tensor(osc.tile_shape),
                     torch.from_numpy(osc.scene_min), scene_size, osc.tile_size, 256, vis)
    frame = torch.full((H, W, 3), -1.0, device=DEV)
    inv = torch.full((H, W), 1e-8, device=DEV)
    net = torch.full((H, W), -1, device=DEV, dtype=torch.int16)
    ns = torch.zeros((H, W), device=DEV, dtype=torch.int32)
    FR.rendering_diffuse_octree_fp16(rays_d, vis, to_dev(cam.origin), to_dev(osc.group_map), to_dev(osc.data_voxels),
                                     to_dev(osc.block_index_map), to_dev(osc.nodes_index_map),
                                     to_dev(osc.nodes_sample_flag), to_dev(osc.voxels_start), to_dev(osc.centers),
                                     osc.tile_size, osc.voxel_size, osc.sample_step, osc.num_voxel, 256, trans_th,
                                     early_stop, soft_hit, frame, inv, net, n_samples=ns)
    rf, ri, rn, rs = oracle_mod.render_octree(osc, cam.origin, rays_d.cpu().numpy(), vis.cpu().numpy(),
                                              trans_th=trans_th, early_stop=early_stop, soft_hit=soft_hit)
    return (frame.cpu().numpy().reshape(-1, 3), inv.cpu().numpy().ravel(), net.cpu().numpy().ravel(),
            ns.cpu().numpy().ravel()), (rf, ri, rn, rs)


@pytest.mark.parametrize("soft_hit", [False, True])
def test_octree_room_matches_oracle(oracle_mod, soft_hit):
    osc = S.make_octree_scene(7, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=5)
    cam = S.room_camera(7, (0, 0, 0), (4, 4, 1), 192, 320, focal=190.0)
    (f, i, n, s), (rf, ri, rn, rs) = render_both(oracle_mod, osc, cam, soft_hit=soft_hit)
    assert s.sum() > 100000
    assert np.array_equal(s, rs)
    assert np.array_equal(i.view(np.uint32), ri.view(np.uint32))
    assert np.array_equal(n, rn)
    assert np.abs(f - rf).max() <= RGB_TOL
    assert (n >= 0).mean() > 0.5


def test_octree_closed_form_scene_matches_oracle(oracle_mod):
    """The linear-field opaque scene of the oracle's closed-form pin, seen from outside the room."""
    def field(t, ijk):
        v = np.empty(ijk.shape[:-1] + (4,), np.float32)
        v[..., :3] = (ijk + 0.5) / 128.0
        v[..., 3] = 60000.0
        return v
    osc = S.make_octree_scene(8, tile_shape=(2, 1, 1), block_occupancy=0.6, field=field, flag_border=True)
    cam = S.look_at_camera((-1.5, 0.5, 0.6), (1.0, 0.5, 0.5), 64, 96, 60.0)
    (f, i, n, s), (rf, ri, rn, rs) = render_both(oracle_mod, osc, cam, K=4)
    assert np.array_equal(s, rs) and np.array_equal(n, rn)
    assert np.array_equal(i.view(np.uint32), ri.view(np.uint32))
    assert np.abs(f - rf).max() <= RGB_TOL
