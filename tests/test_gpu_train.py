"""GPU training path: forward identical to the render path and the oracle; gradients of sigma,
features and MLP parameters within 1e-3 relative error of the oracle's FP64 adjoint (which the
CPU tests pin with FP64 finite differences); the autograd.Function wiring."""
import numpy as np
import pytest
import torch

from snisr import device as D
from snisr import scene as S
from snisr import train as TR

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
GRAD_REL = 1e-3


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)


def case(sc, n, seed, lo, hi, tlo, thi):
    rng = np.random.default_rng(seed)
    origins = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
    targets = rng.uniform(tlo, thi, (n, 3)).astype(np.float32)
    dirs = (targets - origins).astype(np.float32)
    g = rng.standard_normal((n, 3)).astype(np.float32)
    return origins, dirs, g


@pytest.mark.parametrize("width,early", [(32, 0.0), (64, 1e-3)])
def test_train_grads_match_oracle_small(oracle_mod, width, early):
    sc = S.make_scene(5, tile_shape=(2, 1, 1), res=16, occupancy=0.6, feat_dim=16, mlp_width=width, mlp_layers=3,
                      sample_step=1.0 / 32, early_stop=early)
    origins, dirs, g = case(sc, 3000, 6, [-0.5, -0.5, -0.5], [2.5, 1.5, 1.5], [0.1, 0.1, 0.1], [1.9, 0.9, 0.9])
    ds = D.DeviceScene(sc)
    out, ws = TR.train_forward(ds, to_dev(origins), to_dev(dirs))
    ref_r = oracle_mod.render(oracle_mod.OracleScene(sc), origins, dirs)
    assert np.array_equal(out["n_samples"].cpu().numpy(), ref_r["n_samples"])
    assert np.abs(out["rgb"].cpu().numpy() - ref_r["rgb"]).max() <= 1e-4
    dsig, dfeat, dmlp = TR.train_backward(ds, to_dev(origins), to_dev(dirs), to_dev(g), ws)
    ref = oracle_mod.train_grad(oracle_mod.OracleScene(sc), origins, dirs, g, precision="f64")
    for name, got, want in (("sigma", dsig, ref["dsigma"]), ("feats", dfeat, ref["dfeats"]),
                            ("mlp", dmlp, ref["dmlp"])):
        e = rel(got.cpu().numpy(), want)
        print(f"width {width} {name}: rel err {e:.2e}")
        assert e <= GRAD_REL, (name, e)


def test_train_grads_match_oracle_c3_scene(oracle_mod):
    sc, cam, _ = S.config_c2(n_triangles=0, H=8, W=8, feat_half=False)
    rng = np.random.default_rng(7)
    n_per = 256
    origins, dirs = [], []
    for t in range(sc.n_tiles):
        c = sc.tile_centers[t]
        o = rng.uniform([0.1, 0.1, 0.1], [3.9, 3.9, 0.9], (n_per, 3))
        tgt = c + rng.uniform(-0.5, 0.5, (n_per, 3))
        origins.append(o)
        dirs.append(tgt - o)
    origins = np.concatenate(origins).astype(np.float32)
    dirs = np.concatenate(dirs).astype(np.float32)
    g = rng.standard_normal(origins.shape).astype(np.float32)
    ds = D.DeviceScene(sc)
    out, ws = TR.train_forward(ds, to_dev(origins), to_dev(dirs))
    dsig, dfeat, dmlp = TR.train_backward(ds, to_dev(origins), to_dev(dirs), to_dev(g), ws)
    ref = oracle_mod.train_grad(oracle_mod.OracleScene(sc), origins, dirs, g, precision="f64")
    assert np.abs(out["rgb"].cpu().numpy() - ref["rgb"]).max() <= 1e-4
    for name, got, want in (("sigma", dsig, ref["dsigma"]), ("feats", dfeat, ref["dfeats"]),
                            ("mlp", dmlp, ref["dmlp"])):
        e = rel(got.cpu().numpy(), want)
        print(f"C3 scene {name}: rel err {e:.2e}")
        assert e <= GRAD_REL, (name, e)


def test_autograd_function_and_gpu_finite_difference():
    sc = S.make_scene(8, tile_shape=(2, 1, 1), res=16, occupancy=0.6, feat_dim=16, mlp_width=32, mlp_layers=3,
                      sample_step=1.0 / 32, early_stop=0.0)
    origins, dirs, g = case(sc, 2000, 9, [-0.5, -0.5, -0.5], [2.5, 1.5, 1.5], [0.1, 0.1, 0.1], [1.9, 0.9, 0.9])
    ds = D.DeviceScene(sc)
    sigma = ds.sigma.clone().requires_grad_(True)
    feats = ds.feats.clone().requires_grad_(True)
    mlp = ds.mlp_params.clone().requires_grad_(True)
    o, d, gt = to_dev(origins), to_dev(dirs), to_dev(g)
    rgb = TR.render(sigma, feats, mlp, o, d, ds)
    (rgb * gt).sum().backward()
    # the same gradients as the direct calls
    out, ws = TR.train_forward(ds, o, d)
    dsig, dfeat, dmlp = TR.train_backward(ds, o, d, gt, ws)
    # gradient atomics may sum in a different order between calls: compare in norm
    for a, b in ((sigma.grad, dsig), (feats.grad, dfeat), (mlp.grad, dmlp)):
        assert rel(a.cpu().numpy(), b.cpu().numpy()) <= 1e-5
    # finite differences on the GPU forward, reduced in FP64, on the largest MLP gradients
    idx = torch.topk(mlp.grad.abs().flatten(), 5).indices
    eps = 1e-3
    for i in idx.tolist():
        p = mlp.detach().clone().flatten()
        p[i] += eps
        lp = (TR.render(sigma.detach(), feats.detach(), p.view_as(mlp), o, d, ds).double() * gt.double()).sum()
        p[i] -= 2 * eps
        lm = (TR.render(sigma.detach(), feats.detach(), p.view_as(mlp), o, d, ds).double() * gt.double()).sum()
        num = float((lp - lm) / (2 * eps))
        an = float(mlp.grad.flatten()[i])
        assert abs(num - an) <= 1e-2 * abs(an) + 1e-4, (i, num, an)


def test_workspace_reuse_before_backward_raises_and_steps_do_not_synchronise():
    """One TrainWorkspace shared by two pending forwards: the first forward's backward would read the
    second forward's records, so it raises.  After the first (sizing) forward, forwards on the same
    ray count leave the capacity check to the caller (no host synchronisation per step)."""
    sc = S.make_scene(5, tile_shape=(2, 1, 1), res=16, occupancy=0.6, feat_dim=16, mlp_width=32, mlp_layers=3,
                      sample_step=1.0 / 32, early_stop=0.0)
    origins, dirs, g = case(sc, 2000, 7, [-0.5, -0.5, -0.5], [2.5, 1.5, 1.5], [0.1, 0.1, 0.1], [1.9, 0.9, 0.9])
    ds = D.DeviceScene(sc)
    sigma = ds.sigma.clone().requires_grad_(True)
    feats = ds.feats.clone().requires_grad_(True)
    mlp = ds.mlp_params.clone().requires_grad_(True)
    o, d, gt = to_dev(origins), to_dev(dirs), to_dev(g)
    ws = TR.TrainWorkspace()
    a = TR.render(sigma, feats, mlp, o, d, ds, workspace=ws)
    assert ws.sized_for == o.shape[0]
    b = TR.render(sigma, feats, mlp, o, d, ds, workspace=ws)
    with pytest.raises(RuntimeError, match="another forward"):
        (a * gt).sum().backward()
    (b * gt).sum().backward()          # the last forward's backward is fine
    assert not ws.overflowed()
    assert torch.equal(a.detach(), b.detach())
    assert ws.samples_last() > 0
