"""The measured workloads at their full sizes, through the same graph-replayed steps that bench.py
times, against the FP32 path and the oracle:

* C2: the 1280x720 frame of the 16-tile room with the 200k-triangle proxy mesh (camera rays, BVH
  nearest hit bounding the march, FP16 MFMA render), one HIP graph;
* C4: 64 tiles with grids from {64,96,128}^3, 4M rays with Zipf(1.2) per-tile counts, BF16 MFMA,
  the autotuned schedule, one HIP graph;
* C5: 256 tiles, 2M-triangle mesh, 4 views of 1920x1080 in one batch, FP16 MFMA, one HIP graph.

Each asserts per-ray sample counts, T and depth equal to the plain FP32 path on the whole batch;
counts and T bit-exact against the FP32 oracle (given the GPU's mesh bound) and triangle ids and t
bit-exact against a brute-force hit on a seeded ray subset; and PSNR against the FP64 oracle on that
subset (>= 50 dB with FP16 operands, >= 40 dB with BF16).  The MLP is the reference's
rendering/src/raydiffuse_render_octree_fp16.cu:141-235 per-ray loop generalised to feature scenes."""
import numpy as np
import pytest
import torch

from snisr import _native as N
from snisr import autotune
from snisr import device as D
from snisr import graph as G
from snisr import scene as S

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]
DEV = torch.device("cuda:0")
PSNR_F16, PSNR_BF16 = 50.0, 40.0
N_SUB = 2000


def psnr(a, b):
    mse = np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2)
    return 10 * np.log10(1.0 / max(mse, 1e-30))


def check_against_fp32_path(ds, origins, dirs, t_hit, image_width, out):
    f32 = D.render_fp32(ds, origins, dirs, t_hit=t_hit, image_width=image_width, rgb=False)
    for k in ("n_samples", "T", "depth"):
        assert torch.equal(out[k], f32[k]), k


def check_subset(oracle_mod, sc, origins_np, dirs_np, t_hit_np, out, thr, seed, label):
    n = dirs_np.shape[0]
    idx = np.sort(np.random.default_rng(seed).choice(n, N_SUB, replace=False))
    osc = oracle_mod.OracleScene(sc)
    r32 = oracle_mod.render(osc, origins_np, dirs_np, t_hit=t_hit_np, ray_index=idx)
    g_ns = out["n_samples"].cpu().numpy()[idx]
    g_T = out["T"].cpu().numpy()[idx]
    assert np.array_equal(g_ns, r32["n_samples"])
    assert np.array_equal(g_T.view(np.uint32), r32["T"].view(np.uint32))
    r64 = oracle_mod.render(osc, origins_np, dirs_np, t_hit=t_hit_np, ray_index=idx, precision="f64")
    p = psnr(out["rgb"].cpu().numpy()[idx], r64["rgb"])
    print(f"{label}: PSNR vs FP64 on {N_SUB} rays = {p:.2f} dB (bound {thr})", flush=True)
    assert p >= thr
    return idx


def check_mesh_subset(oracle_mod, mesh, origins_np, dirs_np, t_hit, tri, idx):
    o = origins_np if origins_np.size == 3 else origins_np[idx]
    t_ref, id_ref = oracle_mod.mesh_hit(mesh.vertices, mesh.faces, o, oracle_mod.normalize(dirs_np[idx]))
    assert np.array_equal(tri.cpu().numpy()[idx], id_ref)
    assert np.array_equal(t_hit.cpu().numpy()[idx].view(np.uint32), t_ref.view(np.uint32))


def test_c2_frame_as_benched(oracle_mod):
    print("C2 1280x720 + 200k-triangle mesh, FP16 MFMA, graph replay", flush=True)
    sc, cam, mesh = S.config_c2()
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dm = D.DeviceMesh(mesh.vertices, mesh.faces)
    fg = G.FrameGraph(ds, dm, [cam])
    out = fg.replay()
    torch.cuda.synchronize()
    assert not fg.overflowed()
    rays = fg.rays.reshape(-1, 3)
    check_against_fp32_path(ds, fg.origin, rays, fg.t_hit, cam.W, out)
    dirs_np = rays.cpu().numpy()
    th_np = fg.t_hit.cpu().numpy()
    assert np.array_equal(dirs_np, oracle_mod.compute_rays(cam.H, cam.W, cam.intrinsic, cam.c2w))
    idx = check_subset(oracle_mod, sc, cam.origin, dirs_np, th_np, out, PSNR_F16, 21, "C2")
    check_mesh_subset(oracle_mod, mesh, cam.origin, dirs_np, fg.t_hit, fg.tri, idx[:500])
    assert np.isfinite(th_np).mean() > 0.5   # the proxy mesh bounds most rays


def test_c4_full_batch_as_benched(oracle_mod):
    print("C4 64 tiles, 4M Zipf rays, BF16 MFMA, autotuned graph replay", flush=True)
    sc, o_np, d_np, counts = S.config_c4()
    assert sc.n_tiles == 64 and int(counts.sum()) == o_np.shape[0] == 4_000_000
    assert set(np.unique(sc.tile_res).tolist()) <= {64, 96, 128}
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_BF16)
    o, d = torch.from_numpy(o_np).to(DEV), torch.from_numpy(d_np).to(DEV)
    best, _ = autotune.tune_render(ds, o, d, reps=1)
    g = G.RenderGraph(ds, o, d, tuning=best)
    out = g.replay()
    torch.cuda.synchronize()
    assert not g.overflowed()
    check_against_fp32_path(ds, o, d, None, 0, out)
    check_subset(oracle_mod, sc, o_np, d_np, None, out, PSNR_BF16, 41, "C4")


def test_c5_four_views_as_benched(oracle_mod):
    print("C5 256 tiles, 2M-triangle mesh, 4 x 1920x1080, FP16 MFMA, graph replay", flush=True)
    sc, cams, mesh = S.config_c5()
    assert sc.n_tiles == 256 and len(cams) == 4 and mesh.faces.shape[0] >= 1_900_000
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dm = D.DeviceMesh(mesh.vertices, mesh.faces)
    fg = G.FrameGraph(ds, dm, cams)
    out = fg.replay()
    torch.cuda.synchronize()
    assert not fg.overflowed()
    rays = fg.rays.reshape(-1, 3)
    check_against_fp32_path(ds, fg.origin, rays, fg.t_hit, cams[0].W, out)
    o_np = fg.origin.cpu().numpy()
    d_np = rays.cpu().numpy()
    idx = check_subset(oracle_mod, sc, o_np, d_np, fg.t_hit.cpu().numpy(), out, PSNR_F16, 51, "C5")
    check_mesh_subset(oracle_mod, mesh, o_np, d_np, fg.t_hit, fg.tri, idx[:200])
