"""GPU parity of the training-form compositing (utils.volume_rendering, cumprod of 1 - alpha + 1e-10
with an initial transparency) and of the specular head (sigmoid of degree-2 SH), forward and
adjoint, through their autograd Functions, against the FP64 oracle; and one specular-branch
training step of the reference's form (PE -> MLP -> SH + sigmoid, relu sigma -> composite ->
clamp(diffuse + specular) -> MSE) whose parameter gradients agree with an FP64 evaluation of the
same chain within 1e-3 relative."""
import numpy as np
import pytest
import torch

from snisr import train_ref as TR

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def dev(a):
    return torch.tensor(np.asarray(a, np.float32), device=DEV)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("B,n,with_init", [(1000, 64, False), (333, 7, True), (5, 1, True)])
def test_volume_rendering_matches_oracle(oracle_mod, B, n, with_init):
    rng = np.random.default_rng(B + n)
    rgb, sig = rng.uniform(0, 1, (B, n, 3)), rng.uniform(0, 8, (B, n, 1))
    dst = rng.uniform(0, 0.2, (B, n, 1))
    dst[:, -1] = 1e10
    init = rng.uniform(0.1, 1, (B, 1)) if with_init else None
    g_out, g_tr = rng.normal(size=(B, 3)), rng.normal(size=(B, n, 1))
    t_rgb, t_sig = dev(rgb).requires_grad_(), dev(sig).requires_grad_()
    t_init = dev(init).requires_grad_() if with_init else None
    out, T, alpha = TR.volume_rendering(t_rgb, t_sig, dev(dst), t_init)
    ((out * dev(g_out)).sum() + (T * dev(g_tr)).sum()).backward()
    r = oracle_mod.volume_render(rgb, sig[..., 0], dst[..., 0], init_T=init, g_out=g_out, g_trans=g_tr[..., 0])
    np.testing.assert_allclose(out.detach().cpu().numpy(), r["out"], rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(T.detach().cpu().numpy()[..., 0], r["trans"], rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(alpha.detach().cpu().numpy()[..., 0], r["alpha"], rtol=1e-5, atol=1e-7)
    assert rel(t_rgb.grad.cpu().numpy(), r["g_rgb"]) < 1e-5
    assert rel(t_sig.grad.cpu().numpy()[..., 0], r["g_sigma"]) < 1e-4
    if with_init:
        assert rel(t_init.grad.cpu().numpy()[:, 0], r["g_init"]) < 1e-5


def test_background_pass_chained_behind_the_tile_pass(oracle_mod):
    """tile.py:703-704: the background composite starts at the in-tile pass's last transparency;
    the gradient of the sum reaches both passes."""
    rng = np.random.default_rng(7)
    B, n1, n2 = 257, 16, 8
    r1, s1, d1 = rng.uniform(0, 1, (B, n1, 3)), rng.uniform(0, 3, (B, n1, 1)), rng.uniform(0, 0.2, (B, n1, 1))
    r2, s2, d2 = rng.uniform(0, 1, (B, n2, 3)), rng.uniform(0, 3, (B, n2, 1)), rng.uniform(0, 0.5, (B, n2, 1))
    d2[:, -1] = 1e10
    g = rng.normal(size=(B, 3))
    t = [dev(a).requires_grad_() for a in (r1, s1, r2, s2)]
    o1, T1, _ = TR.volume_rendering(t[0], t[1], dev(d1))
    o2, _, _ = TR.volume_rendering(t[2], t[3], dev(d2), T1[:, -1, :])
    ((o1 + o2) * dev(g)).sum().backward()
    # oracle: pass 2 adjoint gives dL/dinit, which enters pass 1 as dL/dtrans of its last sample
    a1 = oracle_mod.volume_render(r1, s1[..., 0], d1[..., 0])
    a2 = oracle_mod.volume_render(r2, s2[..., 0], d2[..., 0], init_T=a1["trans"][:, -1], g_out=g)
    g_tr1 = np.zeros((B, n1))
    g_tr1[:, -1] = a2["g_init"]
    a1 = oracle_mod.volume_render(r1, s1[..., 0], d1[..., 0], g_out=g, g_trans=g_tr1)
    np.testing.assert_allclose((o1 + o2).detach().cpu().numpy(), a1["out"] + a2["out"], rtol=1e-5, atol=1e-6)
    for got, want in ((t[0], a1["g_rgb"]), (t[2], a2["g_rgb"])):
        assert rel(got.grad.cpu().numpy(), want) < 1e-5
    for got, want in ((t[1], a1["g_sigma"]), (t[3], a2["g_sigma"])):
        assert rel(got.grad.cpu().numpy()[..., 0], want) < 1e-4


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_sh_sigmoid_matches_oracle(oracle_mod, deg):
    rng = np.random.default_rng(11 + deg)
    B, S = 300, 16
    q = rng.normal(size=(B, S, 28))
    dirs = rng.normal(size=(B, 3))
    g = rng.normal(size=(B, S, 3))
    tq = dev(q).requires_grad_()
    spec = TR.sh_sigmoid(tq, dev(dirs), deg)
    (spec * dev(g)).sum().backward()
    ref, gq = oracle_mod.sh_sigmoid(q.reshape(-1, 28), dirs, deg, g_spec=g.reshape(-1, 3))
    np.testing.assert_allclose(spec.detach().cpu().numpy().reshape(-1, 3), ref, rtol=1e-5, atol=1e-6)
    assert rel(tq.grad.cpu().numpy().reshape(-1, 28), gq) < 1e-5


def _pe(x, L=10):
    parts = [x]
    for i in range(L):
        parts += [torch.sin(2.0 ** i * x), torch.cos(2.0 ** i * x)]
    return torch.cat(parts, -1)


def _mlp(dtype, device, seed=3):
    torch.manual_seed(seed)
    layers, w = [torch.nn.Linear(63, 64)], 64
    for _ in range(4):
        layers += [torch.nn.LeakyReLU(0.01), torch.nn.Linear(w, w)]
    layers += [torch.nn.LeakyReLU(0.01), torch.nn.Linear(w, 28)]
    return torch.nn.Sequential(*layers).to(device=device, dtype=dtype)


def test_specular_branch_step_gradients(oracle_mod):
    """One step of the reference's specular branch (tile.py:745-760 with train_FLAG 1): the MLP runs in
    torch, the SH head and the composite on the HIP adjoints; the MLP's parameter gradients equal
    those of the same chain evaluated in FP64 (composite and SH by their formulas, Y from the oracle)
    within 1e-3 relative."""
    rng = np.random.default_rng(13)
    B, S = 512, 64
    o = rng.uniform(-1, 1, (B, 3))
    d = rng.normal(size=(B, 3))
    z = np.sort(rng.uniform(0.1, 4.0, (B, S)), 1)
    pts = (o[:, None, :] + z[..., None] * d[:, None, :]) * 0.25          # a normalised point cloud
    dists = np.concatenate([z[:, 1:] - z[:, :-1], np.full((B, 1), 1e10)], 1)[..., None]
    diffuse, gt = rng.uniform(0, 0.5, (B, 3)), rng.uniform(0, 1, (B, 3))
    mask = (rng.uniform(0, 1, (B, 1)) > 0.1).astype(np.float64)

    m32 = _mlp(torch.float32, DEV)
    raw = m32(_pe(dev(pts)))
    spec = TR.sh_sigmoid(raw, dev(d), 2)
    sigma = torch.relu(raw[..., 27:])
    rgb_vd, _, _ = TR.volume_rendering(spec, sigma, dev(dists))
    pred = torch.clamp(dev(diffuse) + dev(mask) * rgb_vd, 0, 1)
    torch.mean((pred - dev(gt)) ** 2).backward()

    m64 = _mlp(torch.float64, "cpu")
    raw64 = m64(_pe(torch.tensor(pts)))
    u = d / np.linalg.norm(d, axis=1, keepdims=True)
    Y = torch.tensor(oracle_mod.sh2_basis(u))                               # [B, 9]
    z64 = torch.einsum("bj,bsjc->bsc", Y, raw64[..., :27].reshape(B, S, 9, 3))
    spec64 = torch.sigmoid(z64)
    sig64 = torch.relu(raw64[..., 27])
    alpha = 1 - torch.exp(-sig64 * torch.tensor(dists[..., 0]))
    T = torch.cumprod(torch.cat([torch.ones(B, 1, dtype=torch.float64), 1 - alpha + 1e-10], 1), 1)[:, :-1]
    rgb64 = (alpha[..., None] * T[..., None] * spec64).sum(1)
    pred64 = torch.clamp(torch.tensor(diffuse) + torch.tensor(mask) * rgb64, 0, 1)
    torch.mean((pred64 - torch.tensor(gt)) ** 2).backward()

    np.testing.assert_allclose(pred.detach().cpu().numpy(), pred64.detach().numpy(), atol=1e-4)
    for p32, p64 in zip(m32.parameters(), m64.parameters()):
        assert rel(p32.grad.cpu().numpy(), p64.grad.numpy()) < 1e-3
