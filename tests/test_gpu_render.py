"""GPU parity of the FP32 render path against the oracle (bit-exact traversal, sample counts,
T and depth; rgb within max abs 1e-4)."""
import numpy as np
import pytest
import torch

from snisr import device as D
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
RGB_TOL = 1e-4  # FP32 pixels vs the FP32 oracle (max abs error)


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return t if dtype is None else t.to(dtype)


def assert_parity(out, ref, idx=None, brick_cap=0):
    g = {k: v.cpu().numpy() for k, v in out.items() if v is not None}
    if idx is not None:
        g = {k: v[idx] for k, v in g.items()}
    assert np.array_equal(g["n_samples"], ref["n_samples"]), np.flatnonzero(g["n_samples"] != ref["n_samples"])[:10]
    assert np.array_equal(g["T"].view(np.uint32), ref["T"].view(np.uint32))
    assert np.array_equal(g["depth"].view(np.uint32), ref["depth"].view(np.uint32))
    if brick_cap:
        assert np.array_equal(g["bricks"], ref["bricks"])
    if "rgb" in g and ref.get("rgb") is not None:
        err = np.abs(g["rgb"] - ref["rgb"]).max()
        assert err <= RGB_TOL, err


@pytest.fixture(scope="module")
def c1():
    sc, cam, _ = S.config_c1()
    return sc, cam, D.DeviceScene(sc)


def test_compute_rays_bit_exact(oracle_mod, c1):
    sc, cam, _ = c1
    out = D.compute_rays(to_dev(cam.intrinsic), to_dev(cam.c2w), cam.H, cam.W)
    ref = oracle_mod.compute_rays(cam.H, cam.W, cam.intrinsic, cam.c2w)
    assert np.array_equal(out.cpu().numpy().reshape(-1, 3).view(np.uint32), ref.view(np.uint32))


def test_trace_tiles_bit_exact(oracle_mod):
    sc = S.make_scene(21, tile_shape=(6, 5, 3), res=8, occupancy=0.0)
    sc.tile_index_map = np.where(np.random.default_rng(0).random(90) < 0.5, np.arange(90), -1).astype(np.int32)
    ds = D.DeviceScene(sc)
    osc = oracle_mod.OracleScene(sc)
    rng = np.random.default_rng(1)
    dirs = rng.standard_normal((20000, 3)).astype(np.float32)
    dirs[:100, 1:] = 0.0   # axis-parallel
    dirs[100:200, 0] = -0.0
    for origin in (np.array([3.1, 2.2, 1.3], np.float32), np.array([-4.0, 9.0, -2.0], np.float32)):
        for K in (1, 4, 30, 64, 130):   # K > 127: lanes write their rows directly (no LDS staging)
            got = D.trace_tiles(ds, to_dev(origin), to_dev(dirs), K).cpu().numpy()
            ref = oracle_mod.trace_tiles(osc, origin, dirs, K)
            assert np.array_equal(got, ref), (origin, K)


def test_render_fp32_c1_full_frame(oracle_mod, c1):
    sc, cam, ds = c1
    dirs = S.camera_rays_np(cam)
    out = D.render_fp32(ds, to_dev(cam.origin), to_dev(dirs), brick_cap=32)
    ref = oracle_mod.render(oracle_mod.OracleScene(sc), cam.origin, dirs, brick_debug=32)
    assert ref["n_samples"].sum() > 100000
    assert_parity(out, ref, brick_cap=32)


def test_render_fp32_visited_list_equals_inline_dda(oracle_mod, c1):
    sc, cam, ds = c1
    dirs = S.camera_rays_np(cam)
    dn = oracle_mod.normalize(dirs)
    vis = D.trace_tiles(ds, to_dev(cam.origin), to_dev(dn), sc.max_tiles)
    a = D.render_fp32(ds, to_dev(cam.origin), to_dev(dirs))
    b = D.render_fp32(ds, to_dev(cam.origin), to_dev(dirs), visited=vis)
    for k in ("n_samples", "T", "depth", "rgb"):
        assert torch.equal(a[k], b[k]), k


def test_render_fp32_c2_fp16_features(oracle_mod):
    sc, cam, _ = S.config_c2(n_triangles=0, H=360, W=640)
    cam.intrinsic[2] *= 2.0  # same field of view at half resolution
    ds = D.DeviceScene(sc)
    dirs = S.camera_rays_np(cam)
    out = D.render_fp32(ds, to_dev(cam.origin), to_dev(dirs), image_width=cam.W, brick_cap=16)
    idx = np.random.default_rng(2).choice(dirs.shape[0], 4000, replace=False)
    ref = oracle_mod.render(oracle_mod.OracleScene(sc), cam.origin, dirs, ray_index=idx, brick_debug=16)
    assert ref["n_samples"].sum() > 10000
    assert_parity(out, ref, idx=idx, brick_cap=16)
    # full-frame invariant: colours are convex combinations weighted by 1 - T
    rgb = out["rgb"].cpu().numpy()
    T = out["T"].cpu().numpy()
    assert np.all(rgb >= 0) and np.all(rgb <= (1 - T)[:, None] + 1e-5)


def test_render_fp32_per_ray_origins_and_t_hit(oracle_mod):
    sc, _, _ = S.config_c2(n_triangles=0, H=8, W=8)
    ds = D.DeviceScene(sc)
    rng = np.random.default_rng(3)
    n = 5000
    origins = rng.uniform([0.1, 0.1, 0.1], [3.9, 3.9, 0.9], (n, 3)).astype(np.float32)
    dirs = rng.standard_normal((n, 3)).astype(np.float32)
    t_hit = rng.uniform(0.05, 3.0, n).astype(np.float32)
    t_hit[::5] = np.inf
    out = D.render_fp32(ds, to_dev(origins), to_dev(dirs), t_hit=to_dev(t_hit))
    ref = oracle_mod.render(oracle_mod.OracleScene(sc), origins, dirs, t_hit=t_hit)
    assert_parity(out, ref)


def test_render_fp32_edge_cases(oracle_mod):
    sc = S.make_scene(31, tile_shape=(2, 1, 1), res=16, occupancy=1.0, feat_dim=16, mlp_width=64, mlp_layers=3,
                      sample_step=1.0 / 64)
    sc.sigma[: sc.n_bricks // 2] = 1e5      # first tile opaque: immediate termination
    ds = D.DeviceScene(sc)
    osc = oracle_mod.OracleScene(sc)
    origins = np.array([[-1.0, 0.5, 0.5], [5.0, 5.0, 5.0], [0.5, 0.5, 0.5], [1.5, 0.5, 0.5], [0.0, 0.0, 0.0]],
                       np.float32)
    dirs = np.array([[1.0, 0, 0], [1.0, 0, 0], [0, 0, 1.0], [0, 1.0, 0], [1.0, 1.0, 1.0]], np.float32)
    for K in (0, 1, 30):
        out = D.render_fp32(ds, to_dev(origins), to_dev(dirs), max_tiles=K, brick_cap=8)
        ref = oracle_mod.render(osc, origins, dirs, max_tiles=K, brick_debug=8)
        assert_parity(out, ref, brick_cap=8)
    # sigma-only march (no rgb requested) reports the same counts
    out = D.render_fp32(ds, to_dev(origins), to_dev(dirs), rgb=False)
    ref = oracle_mod.render(osc, origins, dirs, with_mlp=False)
    ref.pop("rgb")
    assert_parity(out, ref)
