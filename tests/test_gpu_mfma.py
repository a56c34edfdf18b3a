"""GPU checks of the MFMA render path: traversal outputs bit-identical to the FP32 path and the
oracle; pixels within PSNR >= 50 dB (FP16 operands) of the FP64 oracle."""
import numpy as np
import pytest
import torch

from snisr import _native as N
from snisr import device as D
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
PSNR_F16 = 50.0   # FP16 MFMA operands, FP32 accumulation, vs the FP64 oracle
PSNR_BF16 = 40.0  # BF16 operands (8-bit mantissa), vs the FP64 oracle


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 10 * np.log10(1.0 / max(mse, 1e-30))


@pytest.mark.parametrize("dtype,thr", [(N.SNISR_F16, PSNR_F16), (N.SNISR_BF16, PSNR_BF16)])
def test_mfma_c1_vs_fp64_oracle(oracle_mod, dtype, thr):
    sc, cam, _ = S.config_c1()
    ds = D.DeviceScene(sc, mfma_dtype=dtype)
    dirs = S.camera_rays_np(cam)
    out = D.render_mfma(ds, to_dev(cam.origin), to_dev(dirs), brick_cap=16)
    ref64 = oracle_mod.render(oracle_mod.OracleScene(sc), cam.origin, dirs, precision="f64")
    ref32 = oracle_mod.render(oracle_mod.OracleScene(sc), cam.origin, dirs, brick_debug=16)
    g = {k: v.cpu().numpy() for k, v in out.items() if isinstance(v, torch.Tensor)}
    assert np.array_equal(g["n_samples"], ref32["n_samples"])
    assert np.array_equal(g["T"].view(np.uint32), ref32["T"].view(np.uint32))
    assert np.array_equal(g["bricks"], ref32["bricks"])
    assert out["total_samples"] == int(ref32["n_samples"].sum())
    p = psnr(g["rgb"], ref64["rgb"])
    print(f"C1 MFMA dtype={dtype} PSNR vs FP64 = {p:.2f} dB")
    assert p >= thr


def test_mfma_c2_matches_fp32_path_and_fp64(oracle_mod):
    sc, cam, _ = S.config_c2(n_triangles=0, H=360, W=640)
    cam.intrinsic[2] *= 2.0
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dirs = S.camera_rays_np(cam)
    o, d = to_dev(cam.origin), to_dev(dirs)
    a = D.render_mfma(ds, o, d, image_width=cam.W)
    b = D.render_fp32(ds, o, d, image_width=cam.W)
    for k in ("n_samples", "T", "depth"):
        assert torch.equal(a[k], b[k]), k
    idx = np.random.default_rng(5).choice(dirs.shape[0], 3000, replace=False)
    ref64 = oracle_mod.render(oracle_mod.OracleScene(sc), cam.origin, dirs, ray_index=idx, precision="f64")
    p = psnr(a["rgb"].cpu().numpy()[idx], ref64["rgb"])
    p32 = psnr(b["rgb"].cpu().numpy()[idx], ref64["rgb"])
    print(f"C2 PSNR vs FP64: MFMA f16 {p:.2f} dB, FP32 path {p32:.2f} dB")
    assert p >= PSNR_F16
    # full frame: MFMA vs the FP32 path
    assert psnr(a["rgb"].cpu().numpy(), b["rgb"].cpu().numpy()) >= PSNR_F16


def test_mfma_capacity_overflow_is_detected_and_recovered():
    sc, cam, _ = S.config_c1()
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dirs = S.camera_rays_np(cam)
    o, d = to_dev(cam.origin), to_dev(dirs)
    ws = D.MfmaWorkspace()
    small = D.render_mfma(ds, o, d, workspace=ws, capacity=2, check=False)   # 2 records per ray
    total, over, mx = D.mfma_stats(ws.buf, with_max=True)
    assert over == 1 and mx > 2 and total > 2
    ref = D.render_mfma(ds, o, d)
    fits = small["n_samples"] <= 2
    assert torch.isnan(small["rgb"][~fits]).all()
    assert torch.equal(small["rgb"][fits], ref["rgb"][fits])   # rays within their slots are unaffected
    assert int(ref["n_samples"].max()) == mx and int(ref["n_samples"].sum()) == total
    full = D.render_mfma(ds, o, d, workspace=ws, capacity=2, check=True)
    assert not torch.isnan(full["rgb"]).any()
    assert torch.equal(full["rgb"], ref["rgb"])


def test_mfma_sample_capacity_overflow_is_detected_and_recovered():
    sc, cam, _ = S.config_c1()
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dirs = S.camera_rays_np(cam)
    o, d = to_dev(cam.origin), to_dev(dirs)
    ref = D.render_mfma(ds, o, d)
    ws = D.MfmaWorkspace()
    small = D.render_mfma(ds, o, d, workspace=ws, capacity=128, sample_capacity=1000, check=False)
    total, over = D.mfma_stats(ws.buf)
    assert over == 1 and total == int(ref["n_samples"].sum())
    assert torch.isnan(small["rgb"]).all()          # the frame is skipped, not partially rendered
    assert torch.equal(small["n_samples"], ref["n_samples"]) and torch.equal(small["T"], ref["T"])
    full = D.render_mfma(ds, o, d, workspace=ws, check=True)
    assert torch.equal(full["rgb"], ref["rgb"])


def test_mfma_deterministic():
    sc, cam, _ = S.config_c1()
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dirs = S.camera_rays_np(cam)
    o, d = to_dev(cam.origin), to_dev(dirs)
    a = D.render_mfma(ds, o, d)
    b = D.render_mfma(ds, o, d)
    assert torch.equal(a["rgb"], b["rgb"])


def test_flattened_march_matches_nested_march_with_bounds_and_visit_lists(oracle_mod):
    """Arbitrary rays (no image width) take the flattened march loop; camera packets and the FP32
    path take the nested one.  With per-ray t_hit bounds and precomputed visit lists both loops give
    the same samples, transmittance, depth and brick lists, and the FP32 oracle agrees."""
    sc, cam, _ = S.config_c1()
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dirs = S.camera_rays_np(cam)
    rng = np.random.default_rng(17)
    t_hit = rng.uniform(0.3, 3.0, dirs.shape[0]).astype(np.float32)
    t_hit[::11] = np.inf
    o, d, th = to_dev(cam.origin), to_dev(dirs), to_dev(t_hit)
    vis = D.trace_tiles(ds, o, to_dev(oracle_mod.normalize(dirs)), sc.max_tiles)
    flat = D.render_mfma(ds, o, d, t_hit=th, visited=vis, brick_cap=16)                      # flattened
    nested = D.render_fp32(ds, o, d, t_hit=th, visited=vis, image_width=cam.W, brick_cap=16)  # nested
    for k in ("n_samples", "T", "depth", "bricks"):
        assert torch.equal(flat[k], nested[k]), k
    ref = oracle_mod.render(oracle_mod.OracleScene(sc), cam.origin, dirs, t_hit=t_hit, brick_debug=16)
    assert np.array_equal(flat["n_samples"].cpu().numpy(), ref["n_samples"])
    assert np.array_equal(flat["T"].cpu().numpy().view(np.uint32), ref["T"].view(np.uint32))
