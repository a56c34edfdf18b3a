"""Export of training outputs (per-tile Voxel-T / Node-T dense grids, per-group Coeffi-TShared state
dicts; training/scripts/tile/tile.py:415-432, tileGroup.py:531-552) into renderData.npz
(rendering/src/__init__.py:98-135):

* every node cell of the octree tables addresses the value and flag of the dense grid cell it stands
  for, and a block is a node exactly when it holds a sampled voxel;
* at points of sampled cells, the octree trilinear read through the exported tables (the reference's
  block-local convention, raydiffuse_render_octree_fp16.cu:76-96 with half_ope.h:246-284: empty
  cells read zero) equals torch's grid_sample of the training tile's grid (its query_voxel,
  tile.py:225-236), with the voxel values rounded to FP16 as the file stores them;
* the packed reflection MLPs equal pack_reflection_mlp of the saved state dicts, and the group centres
  come from group_center.npy or the mean of the group's tile centres;
* rendering the exported scene with the octree oracle approximates the dense FP32 marcher of the
  same grids (different sample positions: a per-block restart against a per-tile march)."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from snisr import render_data as RD
from snisr import scene as S


def _write_checkpoint(tmp_path, ds, epoch=7):
    g0 = tmp_path / "group0"
    g1 = tmp_path / "group1"
    g0.mkdir()
    g1.mkdir()
    torch.manual_seed(0)
    groups = []
    for gi, gdir in enumerate((g0, g1)):
        mlp = torch.nn.Sequential(*[torch.nn.Linear(i, o) for i, o in S.REF_MLP_DIMS])
        torch.save(mlp.state_dict(), gdir / f"Coeffi-TShared-E{epoch}.pt")
        tiles = []
        for j in np.flatnonzero(ds.groups == gi):
            fidx = 100 + int(j)               # the training's own tile numbering
            np.save(gdir / f"Voxel-T{fidx}-E{epoch}.npy", ds.voxels[j])
            np.save(gdir / f"Node-T{fidx}-E{epoch}.npy", ds.nodes[j].astype(np.int32))
            tiles.append((fidx, int(ds.tile_ids[j])))
        if gi == 1:
            np.save(gdir / "group_center.npy", np.array([9.0, 8.0, 7.0], np.float32))
        groups.append(dict(name=f"g{gi}", logdir=str(gdir), tile_dir=str(gdir), tiles=tiles))
    return groups


@pytest.fixture(scope="module")
def exported(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("ckpt")
    ds = S.make_dense_scene(4, tile_shape=(3, 2, 1), num_voxel=32, block=4, block_occupancy=0.3, drop_tiles=1)
    groups = _write_checkpoint(tmp, ds)
    out = tmp / "renderData.npz"
    RD.export_checkpoint(out, groups, ds.tile_shape, ds.tile_size, ds.scene_min, epoch=7)
    return ds, groups, RD.load_render_data(out), tmp


def test_tables_address_the_dense_cells(exported):
    ds, _, rd, _ = exported
    osc = rd.scene
    NV = ds.num_voxel
    bs, M = NV // 4, NV // 4 + 2
    assert osc.num_voxel == M and osc.tile_shape == ds.tile_shape
    assert np.array_equal(osc.tile_index_map, ds.tile_index_map)
    assert np.allclose(osc.centers, ds.centers)
    half = ds.voxels.astype(np.float16)
    for t in range(ds.tile_ids.size):
        f = ds.nodes[t]
        for b in range(64):
            bx, by, bz = b // 16, (b // 4) % 4, b % 4
            node = osc.block_index_map[t * 64 + b]
            inner = f[1 + bs * bx:1 + bs * (bx + 1), 1 + bs * by:1 + bs * (by + 1), 1 + bs * bz:1 + bs * (bz + 1)]
            assert (node >= 0) == bool((inner == 1).any())
            if node < 0:
                continue
            win = (slice(bs * bx, bs * bx + M), slice(bs * by, bs * by + M), slice(bs * bz, bs * bz + M))
            idx = osc.nodes_index_map[node * M ** 3:(node + 1) * M ** 3]
            flg = osc.nodes_sample_flag[node * M ** 3:(node + 1) * M ** 3]
            wf = f[win].reshape(-1)
            assert np.array_equal(idx >= 0, wf >= 0) and np.array_equal(flg, wf == 1)
            got = osc.data_voxels[osc.voxels_start[node] + idx[idx >= 0]]
            assert np.array_equal(got, half[t][win].reshape(-1, 4)[wf >= 0])


def test_octree_reads_equal_grid_sample(exported):
    ds, _, rd, _ = exported
    osc = rd.scene
    rng = np.random.default_rng(2)
    vs = np.float64(osc.voxel_size)
    M = osc.num_voxel
    bsz = (M - 2) * vs
    for t in range(ds.tile_ids.size):
        cells = np.argwhere(ds.nodes[t][1:-1, 1:-1, 1:-1] == 1)         # sampled interior voxels
        pick = cells[rng.choice(len(cells), min(300, len(cells)), replace=False)]
        local = (pick + rng.uniform(0.0, 1.0, pick.shape)) * vs          # points in the tile frame
        # the training tile's query_voxel: grid_sample(align_corners=True) on the dilated grid, FP16 values
        grid = ds.voxels[t].astype(np.float16).astype(np.float64)
        nv = grid.shape[0]
        gidx = local / vs + 0.5
        inp = torch.from_numpy(np.ascontiguousarray(grid.transpose(3, 2, 1, 0)))[None]
        g = torch.from_numpy(gidx / (nv - 1) * 2 - 1)[None, None, None]
        ref = F.grid_sample(inp, g, mode="bilinear", padding_mode="zeros", align_corners=True)[0, :, 0, 0].T.numpy()
        # the exported tables read the reference's way: block of the point, node, cell indices, 8 corners
        blk = np.minimum((local / bsz).astype(np.int64), 3)
        node = osc.block_index_map[t * 64 + blk[:, 0] * 16 + blk[:, 1] * 4 + blk[:, 2]]
        assert (node >= 0).all()
        fidx = (local - (blk * bsz - vs)) / vs - 0.5
        i0 = fidx.astype(np.int64)
        w = fidx - i0
        out = np.zeros((len(local), 4))
        for cx in (0, 1):
            for cy in (0, 1):
                for cz in (0, 1):
                    cell = ((i0[:, 0] + cx) * M + (i0[:, 1] + cy)) * M + (i0[:, 2] + cz)
                    vi = osc.nodes_index_map[node * M ** 3 + cell]
                    val = np.where(vi[:, None] >= 0,
                                   osc.data_voxels[osc.voxels_start[node] + np.maximum(vi, 0)].astype(np.float64), 0.0)
                    wt = (w[:, 0] if cx else 1 - w[:, 0]) * (w[:, 1] if cy else 1 - w[:, 1]) * \
                         (w[:, 2] if cz else 1 - w[:, 2])
                    out += wt[:, None] * val
        np.testing.assert_allclose(out, ref, rtol=1e-9, atol=1e-9 * np.abs(grid).max())


def test_mlps_and_group_centres(exported):
    ds, groups, rd, tmp = exported
    for gi, grp in enumerate(groups):
        sd = torch.load(f"{grp['logdir']}/Coeffi-TShared-E7.pt", weights_only=True)
        assert np.array_equal(rd.net_params[gi * S.REF_MLP_PARAMS:(gi + 1) * S.REF_MLP_PARAMS],
                              RD.pack_reflection_mlp(sd))
    assert np.allclose(rd.group_centers[0], ds.centers[ds.groups == 0].mean(0))
    assert np.array_equal(rd.group_centers[1], np.array([9.0, 8.0, 7.0], np.float32))
    assert list(rd.group_names) == ["g0", "g1"]
    assert np.array_equal(rd.scene.group_map, ds.groups)


def test_exported_scene_renders_like_the_dense_grids(exported, oracle_mod):
    ds, _, rd, _ = exported
    osc = rd.scene
    cam = S.look_at_camera(np.array([-0.6, -0.5, 0.7]), np.array([1.5, 1.0, 0.5]), 60, 80, 50.0)
    d = S.camera_rays_np(cam)
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    grid = S.make_scene(0, tile_shape=ds.tile_shape, tile_size=ds.tile_size, res=8, occupancy=0.0)
    grid.tile_index_map = ds.tile_index_map.copy()
    vis = oracle_mod.trace_tiles(oracle_mod.OracleScene(grid), cam.origin, d, 8)
    step = np.float32(osc.sample_step)
    oct_rgb, _, _, ns = oracle_mod.render_octree(osc, cam.origin, d, vis, trans_th=0.8, early_stop=1e-5)
    s2g = np.arange(ds.tile_ids.size, dtype=np.int32)
    o = np.repeat(cam.origin[None], d.shape[0], 0)
    dense_rgb, _, _ = oracle_mod.render_dense(ds.voxels.astype(np.float16).astype(np.float32), ds.nodes, s2g,
                                              ds.centers, ds.tile_size, ds.voxel_size, step, o, d, vis)
    mse = np.mean((oct_rgb.astype(np.float64) - dense_rgb) ** 2)
    psnr = 10 * np.log10(1.0 / mse)
    print(f"exported octree vs dense grids: PSNR {psnr:.2f} dB over {d.shape[0]} rays")
    assert (ns > 0).mean() > 0.3 and psnr >= 30.0
