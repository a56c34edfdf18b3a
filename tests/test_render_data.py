"""renderData.npz (the reference's scene file, read by FRender.render_prepare,
rendering/src/__init__.py:98-135): round trip of every array and derived scalar, the reference's
derived-scalar arithmetic, the reflection-MLP packing of a trained state dict, and the format checks."""
import numpy as np
import pytest
import torch

from snisr import render_data as RD
from snisr import scene as S


def _scene():
    osc = S.make_octree_scene(31, tile_shape=(3, 2, 1), block_occupancy=0.4, n_groups=3)
    params, centers = S.make_reflection_groups(31, 3, (0, 0, 0), (3, 2, 1))
    return osc, params, centers


def test_round_trip(tmp_path):
    osc, params, centers = _scene()
    f = tmp_path / "renderData.npz"
    RD.save_render_data(f, osc, params, centers, ["a", "b", "c"])
    rd = RD.load_render_data(f, sample_step_scale=0.2)
    for k in ("tile_index_map", "centers", "group_map", "block_index_map", "nodes_index_map", "nodes_sample_flag",
              "voxels_start", "data_voxels", "scene_min"):
        a, b = getattr(osc, k), getattr(rd.scene, k)
        assert a.dtype == b.dtype and np.array_equal(a, b), k
    assert rd.scene.tile_shape == osc.tile_shape and rd.scene.num_voxel == osc.num_voxel == 18
    assert np.array_equal(rd.net_params, params) and np.array_equal(rd.group_centers, centers)
    assert list(rd.group_names) == ["a", "b", "c"]
    # derived scalars as render_prepare computes them: tile_size = num_voxel * voxel_size (64 voxels),
    # sample_step = 0.2 * voxel_size, in double precision from the stored float32 voxel size
    vs = np.float64(np.float32(osc.voxel_size))
    assert rd.scene.voxel_size == vs
    assert rd.scene.tile_size == 64 * vs == osc.tile_size
    assert rd.scene.sample_step == 0.2 * vs
    assert np.float32(rd.scene.sample_step) == np.float32(osc.sample_step)


def test_rejects_other_octree_depths_and_inconsistent_files(tmp_path):
    osc, params, centers = _scene()
    f = tmp_path / "a.npz"
    RD.save_render_data(f, osc, params, centers)
    d = dict(np.load(f, allow_pickle=False))
    d["octree_depth"] = np.array([3], np.int64)
    np.savez(tmp_path / "b.npz", **d)
    with pytest.raises(ValueError, match="octree_depth"):
        RD.load_render_data(tmp_path / "b.npz")
    d = dict(np.load(f, allow_pickle=False))
    d["net_params"] = d["net_params"][:-1]
    np.savez(tmp_path / "c.npz", **d)
    with pytest.raises(ValueError, match="net_params"):
        RD.load_render_data(tmp_path / "c.npz")
    d = dict(np.load(f, allow_pickle=False))
    d["block_IndexMap"] = np.full_like(d["block_IndexMap"], d["voxels_start"].size)
    np.savez(tmp_path / "d.npz", **d)
    with pytest.raises(ValueError, match="block_IndexMap"):
        RD.load_render_data(tmp_path / "d.npz")


def _reference_style_mlp(seed):
    """The reference's MLP(hiden_depth=6, hiden_width=64, in_channel=3, L=10, num_out=28) layout,
    built here from torch.nn: a Sequential of Linear / LeakyReLU under the name mlp.mlp."""
    torch.manual_seed(seed)
    layers = []
    for li, (i, o) in enumerate(S.REF_MLP_DIMS):
        layers.append(torch.nn.Linear(i, o))
        if li < len(S.REF_MLP_DIMS) - 1:
            layers.append(torch.nn.LeakyReLU(0.01))
    m = torch.nn.Module()
    m.mlp = torch.nn.Module()
    m.mlp.mlp = torch.nn.Sequential(*layers)
    return m


def test_pack_reflection_mlp_layout():
    m = _reference_style_mlp(0)
    flat = RD.pack_reflection_mlp(m.state_dict())
    assert flat.dtype == np.float16 and flat.size == S.REF_MLP_PARAMS
    off = 0
    for lin in [l for l in m.mlp.mlp if isinstance(l, torch.nn.Linear)]:
        o, i = lin.weight.shape
        assert np.array_equal(flat[off:off + o], lin.bias.detach().numpy().astype(np.float16))
        w = flat[off + o:off + o + i * o].reshape(i, o)       # [in][out]
        assert np.array_equal(w, lin.weight.detach().numpy().T.astype(np.float16))
        off += o + i * o
    assert off == S.REF_MLP_PARAMS


def test_pack_reflection_mlp_evaluates_like_the_module(oracle_mod):
    """The oracle's inference_mlp on the packed weights equals the torch module (on the FP16-rounded
    weights) applied to the positional encoding."""
    m = _reference_style_mlp(1)
    flat = RD.pack_reflection_mlp(m.state_dict())
    with torch.no_grad():   # module on the FP16-rounded weights, in float64
        for p in m.parameters():
            p.copy_(p.half().float())
        m = m.double()
    rng = np.random.default_rng(2)
    H, W, spr = 2, 3, 2
    samples = rng.uniform(-1.0, 1.0, (H * W * spr, 3)).astype(np.float32)
    rays = rng.standard_normal((H, W, 3)).astype(np.float32)
    centers = np.zeros((1, 3), np.float32)
    net = np.zeros(1, np.int16)
    ps = np.array([0, H * W], np.int32)
    qpi = np.arange(H * W, dtype=np.int32)
    T = np.ones(H * W, np.float32)
    res = np.zeros((H * W * spr, 4), np.float32)
    oracle_mod.inference_mlp(spr, net, ps, qpi, T, centers, samples, rays, flat, 1.0, 0.0, res)
    x = torch.from_numpy(samples.astype(np.float64))
    enc = torch.cat([x] + [f(2.0 ** k * x) for k in range(10) for f in (torch.sin, torch.cos)], -1)
    with torch.no_grad():
        out = m.mlp.mlp(enc).numpy()
    assert np.allclose(res[:, 3], np.maximum(out[:, 27], 0.0), rtol=1e-6, atol=1e-6)
