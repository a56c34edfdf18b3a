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


@pytest.mark.parametrize("tuning", [None, dict(ray_order=1, mlp_threads=256, sort_grid=256)])
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
