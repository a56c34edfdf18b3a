"""Batched uneven tiles: launch schedules (ray order, MLP workgroup size, sort grid) never change the
render, the autotuner picks among them, and the HIP-graph replay of the whole render equals the
eager render on every replay (with eager work and synchronisations between replays, and after the
inputs are updated in place)."""
import numpy as np
import pytest
import torch

from snisr import autotune
from snisr import device as D
from snisr import graph as G
from snisr import scene as S
from snisr import _native as N

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
PSNR_BF16 = 40.0


def small_c4(n_rays=60000):
    sc, o, d, counts = S.config_c4(seed=14, n_rays=n_rays, tile_shape=(2, 2, 2))
    return sc, torch.from_numpy(o).to(DEV), torch.from_numpy(d).to(DEV), counts


@pytest.fixture(scope="module")
def c4():
    sc, o, d, counts = small_c4()
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_BF16)
    return sc, ds, o, d


def test_every_schedule_gives_identical_render(c4, oracle_mod):
    sc, ds, o, d = c4
    ws = D.MfmaWorkspace()
    ref = D.render_mfma(ds, o, d, workspace=ws, check=True)
    for cand in autotune.CANDIDATES:
        out = D.render_mfma(ds, o, d, workspace=ws, check=True, tuning=cand)
        assert torch.equal(out["rgb"], ref["rgb"]), cand
        assert torch.equal(out["n_samples"], ref["n_samples"]) and torch.equal(out["T"], ref["T"]), cand
    # counts equal the FP32 path's, pixels close to the FP64 oracle
    f32 = D.render_fp32(ds, o, d)
    assert torch.equal(f32["n_samples"], ref["n_samples"])
    idx = np.sort(np.random.default_rng(0).choice(d.shape[0], 1500, replace=False))
    r64 = oracle_mod.render(oracle_mod.OracleScene(sc), o.cpu().numpy(), d.cpu().numpy(), ray_index=idx,
                            precision="f64")
    mse = np.mean((ref["rgb"].cpu().numpy()[idx].astype(np.float64) - r64["rgb"]) ** 2)
    assert 10 * np.log10(1.0 / max(mse, 1e-30)) >= PSNR_BF16


def test_autotuner_returns_fastest_valid_schedule(c4, tmp_path):
    sc, ds, o, d = c4
    path = str(tmp_path / "tune.json")
    best, table = autotune.tune_render(ds, o, d, reps=2, cache_path=path)
    assert len(table) == len(autotune.CANDIDATES)
    assert table[min(table, key=table.get)] == min(table.values())
    assert best in autotune.CANDIDATES
    autotune._memo.clear()
    best2, _ = autotune.tune_render(ds, o, d, cache_path=path)   # served from the file
    assert best2 == best


@pytest.mark.parametrize("tuning", [None, dict(ray_order=1, mlp_threads=256, sort_grid=256),
                                    dict(ray_order=0, mlp_threads=256, sort_grid=256, chunks=3)])
def test_graph_replays_match_eager(c4, tuning):
    sc, ds, o, d = c4
    eager = D.render_mfma(ds, o, d, check=True, tuning=tuning)
    g = G.RenderGraph(ds, o.clone(), d.clone(), tuning=tuning)
    for k in range(3):
        out = g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out["rgb"], eager["rgb"]), k
        assert torch.equal(out["n_samples"], eager["n_samples"]), k
        # unrelated eager work on the device between replays
        x = torch.randn(1 << 20, device=DEV)
        x.mul_(2).add_(1)
        torch.cuda.synchronize()
    assert not g.overflowed()
    # inputs updated in place: the same rays in reverse order
    perm = torch.arange(d.shape[0] - 1, -1, -1, device=DEV)
    g.origins.copy_(o[perm])
    g.dirs.copy_(d[perm])
    out = g.replay()
    torch.cuda.synchronize()
    assert not g.overflowed()
    assert torch.equal(out["rgb"], eager["rgb"][perm])


@pytest.mark.parametrize("chunks", [2, 5, 64])
def test_chunked_pixel_frame_equals_one_pass(chunks):
    """Chunks of whole 8-row packet bands (a 256x96 image has 12 bands: 5 chunks split them
    unevenly, 64 is clamped to 12) give the one-pass render bit for bit, and so does a ray count
    that is not a whole number of 64-ray wavefronts."""
    sc, cam, _ = S.config_c2(n_triangles=0, H=96, W=256)
    cam.intrinsic[2] *= 5.0
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    dirs = torch.from_numpy(S.camera_rays_np(cam)).to(DEV)
    o = torch.from_numpy(cam.origin).to(DEV)
    ref = D.render_mfma(ds, o, dirs, image_width=cam.W, check=True)
    out = D.render_mfma(ds, o, dirs, image_width=cam.W, check=True, tuning=dict(chunks=chunks, mlp_threads=256))
    for k in ("rgb", "n_samples", "T", "depth"):
        assert torch.equal(out[k], ref[k]), k
    assert out["total_samples"] == ref["total_samples"]
    # no packets: 64-ray chunks over 24565 rays (the last wavefront partly filled)
    flat = dirs.reshape(-1, 3)[:24565].contiguous()
    ref = D.render_mfma(ds, o, flat, check=True)
    out = D.render_mfma(ds, o, flat, check=True, tuning=dict(chunks=chunks))
    for k in ("rgb", "n_samples", "T", "depth"):
        assert torch.equal(out[k], ref[k]), k


def test_chunked_schedule_rejects_ray_sort_and_bad_counts(c4):
    sc, ds, o, d = c4
    with pytest.raises(N.SnisrError) as e:
        D.render_mfma(ds, o, d, check=False, tuning=dict(ray_order=1, chunks=2))
    assert e.value.status == N.SNISR_ERR_ARG
    with pytest.raises(N.SnisrError) as e:
        D.render_mfma(ds, o, d, check=False, tuning=dict(chunks=65))
    assert e.value.status == N.SNISR_ERR_ARG


@pytest.mark.parametrize("records", ["tight", "loose", 96])
def test_record_schedules_recover_from_a_small_first_workspace(c4, records):
    """A checked render that overflows its first workspace (8 slots per ray) is repeated with room
    for the schedule's own slot count, which may exceed the largest count + 25%."""
    sc, ds, o, d = c4
    ref = D.render_mfma(ds, o, d, check=True)
    ws = D.MfmaWorkspace()
    out = D.render_mfma(ds, o, d, workspace=ws, capacity=8, check=True, tuning=dict(records=records))
    assert torch.equal(out["rgb"], ref["rgb"]) and torch.equal(out["n_samples"], ref["n_samples"])
    assert ws.cpr >= D.schedule_records(dict(records=records), ws.max_count)
