"""The FRender.render-form frame driver (rendering/src/__init__.py:247-341) on an octree room with
per-group reflection MLPs, against the oracle chain: diffuse layer within 1e-4 (samples, inverse
depth and groups bit-exact), reflection layer within the FP16-MLP tolerance of the reflection-pass
test (2e-2 max, 1e-3 mean)."""
import numpy as np
import pytest
import torch
import torch.nn.functional as F

from snisr import frender
from snisr import scene as S

pytestmark = pytest.mark.gpu


def test_frender_frame_matches_oracle_chain(oracle_mod):
    osc = S.make_octree_scene(9, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=6)
    params, centers = S.make_reflection_groups(19, 6, (0, 0, 0), (4, 4, 1))
    H, W = 96, 160
    cam = S.room_camera(9, (0, 0, 0), (4, 4, 1), H, W, focal=95.0)
    fr = frender.FRender(osc, params, centers, H=H, W=W)
    fr.set_camera(cam)
    frame = fr.render().cpu().numpy()
    diffuse = fr.frame_diffuse.cpu().numpy().reshape(-1, 3)
    spec_up = fr.frame_speculars.cpu().numpy()
    # diffuse layer through the oracle (rays and tile lists are bit-exact, tested elsewhere)
    dirs = oracle_mod.compute_rays(H, W, cam.intrinsic, cam.c2w)
    vis = oracle_mod.trace_tiles(oracle_mod.OracleScene(_grid_scene(osc)), cam.origin, dirs, 30)
    rf, ri, rn, rs = oracle_mod.render_octree(osc, cam.origin, dirs, vis, trans_th=0.8, early_stop=0.01)
    assert np.abs(diffuse - rf).max() <= 1e-4
    # reflection layer: the oracle chain on the downscaled grid
    DH, DW = H // 2, W // 2
    inv_near = ri.reshape(H, W)[::2, ::2].copy()
    net = rn.reshape(H, W)[::2, ::2].ravel().copy()
    qpi, bnet, bps = S.reflection_blocks(net)
    far = np.float32(fr.far)
    inv_bound = ((np.float32(1.0) / (np.float32(1.0) / inv_near + far) - inv_near) / np.float32(63)).astype(np.float32)
    dirs_ds = np.ascontiguousarray(dirs.reshape(H, W, 3)[::2, ::2])
    T = np.ones(DH * DW, np.float32)
    col = np.zeros((DH * DW, 3), np.float32)
    for step in range(0, 64, 16):
        pts, st = oracle_mod.inverse_z_sampling(cam.origin, dirs_ds.reshape(-1, 3), inv_near, inv_bound, step, 16)
        res = np.zeros((DH * DW * 16, 4), np.float32)
        oracle_mod.inference_mlp(16, bnet, bps, qpi, T, centers, pts, dirs_ds, params, 1.0 / fr.far, 0.01, res,
                                 quant16=True)
        oracle_mod.integrate_points(res, st, np.float32(0.01), T, col)
    up = F.interpolate(torch.from_numpy(col.reshape(1, DH, DW, 3)).permute(0, 3, 1, 2), scale_factor=2,
                       mode="bilinear", align_corners=True)[0].permute(1, 2, 0).numpy()
    err = np.abs(spec_up - up)
    assert err.max() <= 2e-2 and err.mean() <= 1e-3, (err.max(), err.mean())
    assert np.array_equal(frame, np.clip(diffuse.reshape(H, W, 3) + spec_up, 0, 1))
    assert fr.n_reflection_blocks > 0 and (net >= 0).mean() > 0.3


def _grid_scene(osc):
    """A brick-format scene with the octree scene's tile grid, for the oracle's tile tracer."""
    sc = S.make_scene(0, tile_shape=osc.tile_shape, tile_size=osc.tile_size, scene_min=tuple(osc.scene_min), res=8,
                      occupancy=0.0)
    sc.tile_index_map = osc.tile_index_map.copy()
    return sc


def test_frender_from_render_data_file(tmp_path):
    """A scene written to renderData.npz and read back renders the identical frame."""
    from snisr import render_data as RD
    osc = S.make_octree_scene(10, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=4)
    params, centers = S.make_reflection_groups(20, 4, (0, 0, 0), (4, 4, 1))
    H, W = 64, 96
    cam = S.room_camera(10, (0, 0, 0), (4, 4, 1), H, W, focal=70.0)
    a = frender.FRender(osc, params, centers, H=H, W=W)
    a.set_camera(cam)
    fa = a.render().cpu()
    f = tmp_path / "renderData.npz"
    RD.save_render_data(f, osc, params, centers)
    b = frender.FRender.from_render_data(f, H=H, W=W)
    b.set_camera(cam)
    fb = b.render().cpu()
    assert torch.equal(fa, fb)
    assert len(b.group_names) == 4


def test_frender_cnn_refinement_matches_fp32_module():
    """With the refinement network enabled (FP16 on the device), the frame equals the FP32 module
    applied on the CPU to the same 5-channel input within FP16 tolerance."""
    from snisr import cnn as CN
    osc = S.make_octree_scene(11, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=4)
    params, centers = S.make_reflection_groups(21, 4, (0, 0, 0), (4, 4, 1))
    H, W = 64, 96
    cam = S.room_camera(11, (0, 0, 0), (4, 4, 1), H, W, focal=70.0)
    fr = frender.FRender(osc, params, centers, H=H, W=W)
    fr.set_camera(cam)
    base = fr.render().clone()
    torch.manual_seed(0)
    ref_net = CN.CNN().eval()
    fr.load_cnn(ref_net.state_dict())
    out = fr.render()
    x = torch.cat([base.permute(2, 0, 1), fr.large_inverse_near.reshape(1, H, W),
                   fr.boundary.reshape(1, H, W).float()], 0)[None].cpu()
    with torch.no_grad():
        ref = ref_net(x)[0].permute(1, 2, 0).clamp(0, 1)
    err = (out.cpu() - ref).abs()
    assert err.max() < 2e-2 and err.mean() < 2e-3, (err.max(), err.mean())
    net = fr.large_netIdx
    assert bool(fr.boundary.any()) == bool((net != net.flatten()[0]).any())   # flags iff the groups vary


@pytest.mark.parametrize("n,groups,seed", [(230400, 9, 0), (4097, 3, 1), (1, 1, 2), (640, 40, 3)])
def test_regroup_pixels_matches_host_regrouping(n, groups, seed):
    """snisr_regroup_pixels against the reference's regrouping in numpy (stable sort, balanced
    <= 64-pixel blocks per group, group -1 included): bit-exact query, block groups and starts."""
    from snisr import fastrendering as FR
    rng = np.random.default_rng(seed)
    net = rng.integers(-1, groups - 1, size=n, endpoint=True).astype(np.int16)
    q_ref, bnet_ref, bps_ref = S.reflection_blocks(net)
    dev = torch.device("cuda:0")
    from snisr import _native as N
    mb = FR.regroup_max_blocks(n, 64)
    query = torch.empty(n, dtype=torch.int32, device=dev)
    bnet = torch.full((mb,), 7, dtype=torch.int16, device=dev)
    bps = torch.full((mb + 1,), 7, dtype=torch.int32, device=dev)
    nbl = torch.zeros(1, dtype=torch.int32, device=dev)
    ws = torch.empty(int(N.lib().snisr_regroup_workspace_size(n)), dtype=torch.uint8, device=dev)
    FR.regroup_pixels(torch.from_numpy(net).to(dev), 64, query, bnet, bps, nbl, ws)
    k = int(nbl)
    assert k == bnet_ref.size
    assert np.array_equal(query.cpu().numpy(), q_ref)
    assert np.array_equal(bnet[:k].cpu().numpy(), bnet_ref)
    assert np.array_equal(bps[:k + 1].cpu().numpy(), bps_ref)
    assert bool((bnet[k:] == -1).all()) and bool((bps[k:] == n).all())


@pytest.mark.parametrize("H,W", [(96, 160), (720, 1280)])
def test_frender_device_and_graphed_frames_equal_reference_form(H, W):
    """render_device (no host synchronisation) and its HIP-graph replay give the frame of render(),
    bit for bit, across camera changes; the replays interleave with eager work.  At 720p the
    regrouping sorts 230,400 pixels, where the radix sort takes its multi-pass path."""
    print(f"frender graphed {H}x{W}", flush=True)
    osc = S.make_octree_scene(9, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=6)
    params, centers = S.make_reflection_groups(19, 6, (0, 0, 0), (4, 4, 1))
    cams = [S.room_camera(sd, (0, 0, 0), (4, 4, 1), H, W, focal=95.0 * W / 160) for sd in (9, 10)]
    fr = frender.FRender(osc, params, centers, H=H, W=W)
    fr.load_cnn()
    eager = []
    for cam in cams:
        fr.set_camera(cam)
        eager.append(fr.render().clone())
        assert torch.equal(fr.render_device(), eager[-1])
    for rep in range(3):
        for k, cam in enumerate(cams):
            fr.set_camera(cam)
            out = fr.render_graphed()
            torch.cuda.synchronize()
            assert torch.equal(out, eager[k]), (rep, k)
            _ = torch.randn(1000, device="cuda").sum().item()  # eager work between replays


def _psnr(a, b):
    mse = np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)
    return 10 * np.log10(1.0 / max(mse, 1e-30))


@pytest.mark.timeout(600)
def test_frender_720p_frame_psnr_vs_fp64_chain(oracle_mod):
    """The benched reference-form frame (1280x720, octree room, 8 reflection groups with FP16 MFMA
    MLPs) against the oracle chain on 600 seeded pixels: the diffuse pass in the reference's FP32
    arithmetic (within 1e-4), the reflection MLPs and compositing in FP64 without FP16 rounding of the
    activations, the x2 bilinear upsampling and the clamp.  Frame and low-resolution reflection layer
    >= 50 dB PSNR; the FP16-quantised oracle chain stays within the reflection pass tolerance
    (2e-2 max, 1e-3 mean).  (rendering/src/__init__.py:247-334, rayreflection_render.cu:111-301)"""
    print("frender 720p PSNR vs FP64", flush=True)
    H, W = 720, 1280
    osc = S.make_octree_scene(12, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=8)
    params, centers = S.make_reflection_groups(13, 8, (0, 0, 0), (4, 4, 1))
    cam = S.room_camera(2, (0, 0, 0), (4, 4, 1), H, W)
    fr = frender.FRender(osc, params, centers, H=H, W=W)
    fr.set_camera(cam)
    frame = fr.render().cpu().numpy().reshape(-1, 3)
    spec_lo = fr.frame_speculars_lowres.cpu().numpy().reshape(-1, 3)
    diffuse = fr.frame_diffuse.cpu().numpy().reshape(-1, 3)
    hi = np.sort(np.random.default_rng(31).choice(H * W, 600, replace=False))
    dirs = oracle_mod.compute_rays(H, W, cam.intrinsic, cam.c2w)
    grid = oracle_mod.OracleScene(_grid_scene(osc))
    ref = oracle_mod.frender_subset(osc, grid, params, centers, cam.origin, dirs, H, W, fr.far, hi)
    assert np.abs(diffuse[hi] - ref["diffuse"]).max() <= 1e-4
    lo = ref["lo_idx"]
    assert (ref["net_lo"] >= 0).mean() > 0.3             # most sampled pixels hit a surface with a group
    p_spec = _psnr(spec_lo[lo], ref["spec_lo"])
    p_frame = _psnr(frame[hi], ref["frame"])
    print(f"720p reference-form frame: PSNR vs FP64 chain frame {p_frame:.2f} dB, reflection {p_spec:.2f} dB",
          flush=True)
    assert p_frame >= 50.0 and p_spec >= 50.0
    q16 = oracle_mod.frender_subset(osc, grid, params, centers, cam.origin, dirs, H, W, fr.far, hi, quant16=True)
    err = np.abs(spec_lo[lo] - q16["spec_lo"])
    assert err.max() <= 2e-2 and err.mean() <= 1e-3, (err.max(), err.mean())


def test_frender_render_modes():
    """Render modes 2 (diffuse), 3 (reflection), 4 (depth) and 1 (full frame) as the reference's
    FRender.render returns them (rendering/src/__init__.py:296-303, 330-334), eager and graphed."""
    H, W = 96, 160
    osc = S.make_octree_scene(9, tile_shape=(4, 4, 1), block_occupancy=0.3, n_groups=6)
    params, centers = S.make_reflection_groups(19, 6, (0, 0, 0), (4, 4, 1))
    cam = S.room_camera(9, (0, 0, 0), (4, 4, 1), H, W, focal=95.0)
    fr = frender.FRender(osc, params, centers, H=H, W=W)
    fr.set_camera(cam)
    full = fr.render().clone()
    diffuse, spec = fr.frame_diffuse.clone(), fr.frame_speculars.clone()
    depth = (1 / fr.large_inverse_near[..., None] / 4.0).repeat(1, 1, 3)   # max_depth = max(4, 4, 1) * 1
    assert torch.equal(full, torch.clamp(diffuse + spec, 0, 1))
    for mode, want in ((2, diffuse), (3, spec), (4, depth), (1, full)):
        fr.set_mode(mode)
        assert torch.equal(fr.render(), want), mode
        assert torch.equal(fr.render_device(), want), mode
        assert torch.equal(fr.render_graphed(), want), mode
    with pytest.raises(ValueError):
        fr.set_mode(5)


def test_frender_renders_an_exported_training_checkpoint(oracle_mod, tmp_path):
    """Training outputs (dense Voxel-T / Node-T grids and Coeffi state dicts) exported to
    renderData.npz and rendered by FRender.from_render_data: the diffuse layer equals the octree
    oracle on the exported scene (within 1e-4; inverse depth and groups bit-exact)."""
    from snisr import render_data as RD
    from test_checkpoint_export import _write_checkpoint
    ds = S.make_dense_scene(4, tile_shape=(3, 2, 1), num_voxel=64, block=8, block_occupancy=0.3, drop_tiles=1)
    groups = _write_checkpoint(tmp_path, ds)
    f = tmp_path / "renderData.npz"
    RD.export_checkpoint(f, groups, ds.tile_shape, ds.tile_size, ds.scene_min, epoch=7)
    H, W = 96, 160
    fr = frender.FRender.from_render_data(str(f), H=H, W=W)
    cam = S.look_at_camera(np.array([-0.6, -0.5, 0.7]), np.array([1.5, 1.0, 0.5]), H, W, 95.0)
    fr.set_camera(cam)
    frame = fr.render()
    assert torch.isfinite(frame).all()
    rd = RD.load_render_data(str(f))
    dirs = oracle_mod.compute_rays(H, W, cam.intrinsic, cam.c2w)
    grid = oracle_mod.OracleScene(_grid_scene(rd.scene))
    vis = oracle_mod.trace_tiles(grid, cam.origin, dirs, 30)
    rf, ri, rn, rs = oracle_mod.render_octree(rd.scene, cam.origin, dirs, vis, trans_th=0.8, early_stop=0.01)
    assert np.abs(fr.frame_diffuse.cpu().numpy().reshape(-1, 3) - rf).max() <= 1e-4
    assert np.array_equal(fr.large_inverse_near.cpu().numpy().reshape(-1).view(np.uint32), ri.view(np.uint32))
    assert np.array_equal(fr.large_netIdx.cpu().numpy().reshape(-1), rn)
    assert (rs > 0).mean() > 0.2
