"""GPU parity of the training-form compositing (utils.volume_rendering, cumprod of 1 - alpha + 1e-10
with an initial transparency) and of the specular head (sigmoid of degree-2 SH), forward and
adjoint, through their autograd Functions, against the FP64 oracle; networks.MLP (positional
encoding + Linear/LeakyReLU, the PEMLP module) forward and parameter gradient against the oracle and
against torch's own layers; and one specular-branch training step of the reference's form (PE -> MLP -> SH + sigmoid, relu sigma -> composite ->
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


@pytest.mark.parametrize("B,n,with_init", [(1000, 64, False), (333, 7, True), (5, 1, True), (77, 150, True), (40, 129, False)])
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


@pytest.mark.parametrize("mlp", ["hip", "torch"])
def test_specular_branch_step_gradients(oracle_mod, mlp):
    """One step of the reference's specular branch (tile.py:745-760 with train_FLAG 1): the MLP (the
    HIP PEMLP, or torch layers), the SH head and the composite on the HIP adjoints; the MLP's parameter
    gradients equal those of the same chain evaluated in FP64 (composite and SH by their formulas, Y
    from the oracle) within 1e-3 relative."""
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
    if mlp == "hip":
        state = {f"mlp.mlp.{k}": v for k, v in m32.state_dict().items()}
        pe = TR.PEMLP(6, 28, 0.01, device=DEV).load_linear_state(state)
        raw = pe(dev(pts))
    else:
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
    if mlp == "hip":
        g64 = torch.cat([q.grad.reshape(-1) for q in m64.parameters()]).numpy()
        assert rel(pe.params.grad.cpu().numpy(), g64) < 1e-3
        return
    for p32, p64 in zip(m32.parameters(), m64.parameters()):
        assert rel(p32.grad.cpu().numpy(), p64.grad.numpy()) < 1e-3


def _pemlp_params(rng, depth, n_out):
    parts = []
    for k in range(depth):
        ni, no = (63 if k == 0 else 64), (n_out if k == depth - 1 else 64)
        parts += [rng.normal(0, 1.2 / np.sqrt(ni), no * ni), rng.normal(0, 0.1, no)]
    return np.concatenate(parts)


def _min_abs_preactivation(p, pts, depth, slope):
    """Per sample, the smallest |pre-activation| over the hidden layers (numpy FP64)."""
    x = np.concatenate([pts] + [f(2.0 ** i * pts) for i in range(10) for f in (np.sin, np.cos)], 1)
    off, m = 0, np.full(len(pts), np.inf)
    for k in range(depth - 1):
        ni = 63 if k == 0 else 64
        z = x @ p[off:off + 64 * ni].reshape(64, ni).T + p[off + 64 * ni:off + 64 * ni + 64]
        off += 64 * ni + 64
        m = np.minimum(m, np.abs(z).min(1))
        x = np.where(z > 0, z, slope * z)
    return m


@pytest.mark.parametrize("depth,n_out,slope,n", [(6, 28, 0.01, 70001), (8, 4, 0.0, 4099), (8, 4, 0.0, 1),
                                                 (2, 28, 0.2, 300)])
def test_pemlp_matches_oracle(oracle_mod, depth, n_out, slope, n):
    """networks.MLP (L = 10, width 64) forward and parameter gradient on the HIP kernels against the
    FP64 oracle: outputs within 2e-5 absolute + 1e-4 relative, the gradient within 2e-5 relative in
    norm.  Samples with a hidden pre-activation within 1e-4 of zero get no output gradient: there an
    FP32 evaluation (the reference's torch one included) may take the other branch of the
    activation's derivative, a discrete difference that is not a rounding error."""
    rng = np.random.default_rng(depth * 1000 + n)
    p = _pemlp_params(rng, depth, n_out)
    pts = rng.uniform(-1, 1, (n, 3))
    g_out = rng.normal(size=(n, n_out)) / np.sqrt(n)
    g_out[_min_abs_preactivation(p.astype(np.float32).astype(np.float64), pts.astype(np.float32), depth, slope)
          < 1e-4] = 0.0
    m = TR.PEMLP(depth, n_out, slope, device=DEV)
    with torch.no_grad():
        m.params.copy_(dev(p))
    x = dev(pts)
    out = m(x)
    out.backward(dev(g_out))
    p32 = m.params.detach().double().cpu().numpy()            # the FP32-rounded weights the kernels used
    ro, rg = oracle_mod.pemlp(p32, pts.astype(np.float32), depth, n_out, slope, g_out=g_out.astype(np.float32))
    np.testing.assert_allclose(out.detach().cpu().numpy(), ro, rtol=1e-4, atol=2e-5)
    assert rel(m.params.grad.cpu().numpy(), rg) < 2e-5
    with torch.no_grad():
        np.testing.assert_allclose(m(x).cpu().numpy(), out.detach().cpu().numpy(), rtol=0, atol=0)


def test_pemlp_loads_the_reference_module_state():
    """load_linear_state takes networks.MLP's state dict (mlp.mlp.{2k}.weight / bias) and reproduces
    the torch module's forward and parameter gradients."""
    torch.manual_seed(9)

    class Ref(torch.nn.Module):                 # networks.MLP's module tree: MLP.mlp (GeneralMLP).mlp
        def __init__(self):
            super().__init__()
            self.mlp = torch.nn.Module()
            self.mlp.mlp = _mlp(torch.float32, DEV, seed=9)

        def forward(self, x):
            return self.mlp.mlp(_pe(x))

    ref = Ref()
    m = TR.PEMLP(6, 28, 0.01, device=DEV).load_linear_state(ref.state_dict())
    x = torch.rand((2000, 3), device=DEV) * 2 - 1
    g = torch.randn((2000, 28), device=DEV)
    a, b = m(x), ref(x)
    np.testing.assert_allclose(a.detach().cpu().numpy(), b.detach().cpu().numpy(), rtol=1e-4, atol=1e-4)
    a.backward(g)
    b.backward(g)
    gref = torch.cat([t.grad.reshape(-1) for t in ref.parameters()]).cpu().numpy()
    assert rel(m.params.grad.cpu().numpy(), gref.astype(np.float64)) < 1e-4


def test_pemlp_rejects_bad_shapes():
    with pytest.raises(ValueError):
        TR.PEMLP(9, 4, 0.0, device=DEV)
    with pytest.raises(ValueError):
        TR.PEMLP(6, 27, 0.0, device=DEV)


def _flags(rng, D, p):
    return np.where(rng.uniform(0, 1, (D, D, D)) < p, 1, -1).astype(np.int16)


@pytest.mark.parametrize("D,p", [(18, 0.4), (130, 0.3), (3, 0.0), (5, 1.0)])
def test_flag_cells_and_bake_points(oracle_mod, D, p):
    """Ordered compaction of the occupied cells (more than 1024 workgroups at D = 130) and their MLP
    inputs, bit-exact against the oracle (itself equal to the reference's torch expressions)."""
    rng = np.random.default_rng(D)
    flags = _flags(rng, D, p)
    cells = TR.flag_cells(torch.tensor(flags, device=DEV))
    ref = oracle_mod.flag_cells(flags)
    np.testing.assert_array_equal(cells.cpu().numpy(), ref)
    if ref.size == 0:
        return
    mc, vs, c, s = np.array([0.37, -1.25, 2.5], np.float32), 0.0625, np.array([0.5, -0.75, 2.0], np.float32), 3.7
    pts = torch.empty((ref.size, 3), dtype=torch.float32, device=DEV)
    from snisr import _native as NN
    from snisr.device import _ptr
    NN.check(NN.lib().snisr_bake_points(_ptr(cells), ref.size, D, *mc.tolist(), vs, *c.tolist(), s, _ptr(pts), None),
             "bake_points")
    np.testing.assert_array_equal(pts.cpu().numpy(), oracle_mod.bake_points(ref, D, mc, vs, c, s))


def test_diffuse_head_and_voxel_query_match_oracle(oracle_mod):
    rng = np.random.default_rng(12)
    raw = rng.normal(size=(5000, 4)) * 3
    g = rng.normal(size=(5000, 4))
    t = dev(raw).requires_grad_(True)
    out = TR._DiffuseHead.apply(t, None, 5000)
    out.backward(dev(g))
    ro, rg = oracle_mod.diffuse_head(raw.astype(np.float32), g.astype(np.float32))
    np.testing.assert_allclose(out.detach().cpu().numpy(), ro, rtol=1e-6, atol=1e-7)
    # s (1 - s) in FP32 (torch's sigmoid backward form) cancels for |raw| > 8: absolute bound
    np.testing.assert_allclose(t.grad.cpu().numpy(), rg, rtol=1e-5, atol=1e-6)

    D, vs = 11, 0.07
    mc = np.array([-0.3, 0.2, 1.1], np.float32)
    vox = rng.uniform(0, 1, (D, D, D, 4))
    pts = mc + rng.uniform(-0.6, D + 0.6, (20000, 3)) * vs
    g_rgb, g_sig = rng.normal(size=(20000, 3)), rng.normal(size=(20000, 1))
    tv = dev(vox).requires_grad_(True)
    rgb, sig = TR.query_voxel(tv, dev(pts), mc, vs)
    ((rgb * dev(g_rgb)).sum() + (sig * dev(g_sig)).sum()).backward()
    r = oracle_mod.voxel_query(vox.astype(np.float32), pts.astype(np.float32), mc, vs, g_rgb=g_rgb.astype(np.float32),
                               g_sigma=g_sig.astype(np.float32))
    np.testing.assert_allclose(rgb.detach().cpu().numpy(), r["rgb"], atol=2e-5)
    np.testing.assert_allclose(sig.detach().cpu().numpy()[:, 0], r["sigma"], atol=2e-5)
    assert rel(tv.grad.cpu().numpy(), r["g_vox"]) < 1e-5


def test_diffuse_training_step_through_the_bake():
    """One diffuse training step of the reference's form (tile.py:686-710, train_FLAG 0): the surface
    MLP baked into the dilated grid (compute_voxels), the in-tile samples read from it (query_voxel),
    the background samples through the MLP (inference_diffuse), both composited with the background
    chained behind the tile's leftover transparency, MSE.  Everything runs on the HIP kernels; the
    loss and the MLP's parameter gradients equal those of the same step in FP64 torch (nn.Linear,
    F.grid_sample, the compositing formula) within 1e-3 relative."""
    rng = np.random.default_rng(31)
    NV, vs = 16, 0.05
    D, tile_size = NV + 2, NV * vs
    tc = np.array([0.2, -0.1, 0.4])
    mc = (tc - tile_size / 2 - vs).astype(np.float32)
    gc, gs = np.array([0.0, 0.0, 0.3], np.float32), 2.0
    flags = _flags(rng, D, 0.4)
    B, S, NB = 512, 32, 8
    o = tc + rng.uniform(-0.4, 0.4, (B, 3)) * tile_size
    d = rng.normal(size=(B, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    z = np.concatenate([np.sort(rng.uniform(0, 0.8, (B, S)), 1), np.sort(rng.uniform(0.9, 2.0, (B, NB)), 1)], 1)
    dists = np.concatenate([z[:, 1:] - z[:, :-1], np.full((B, 1), 1e10)], 1)
    pts = (o[:, None] + z[..., None] * d[:, None]).astype(np.float32)
    gt = rng.uniform(0, 1, (B, 3))
    bg_mask = (rng.uniform(0, 1, (B, 1)) < 0.8).astype(np.float64)
    p = _pemlp_params(rng, 8, 4)

    m = TR.PEMLP(8, 4, 0.0, device=DEV)
    with torch.no_grad():
        m.params.copy_(dev(p))
    cells = TR.flag_cells(torch.tensor(flags, device=DEV))
    vox = TR.bake_voxels(m, cells, D, mc, vs, gc, gs)
    rgb, sigma = TR.query_voxel(vox, dev(pts[:, :S]), mc, vs)
    bg_rgb, bg_sigma = TR.inference_diffuse(m, dev(pts[:, S:]), gc, gs)
    d_in, trans, _ = TR.volume_rendering(rgb, sigma, dev(dists[:, :S, None]))
    d_bg, _, _ = TR.volume_rendering(bg_rgb, bg_sigma, dev(dists[:, S:, None]), trans[:, -1:])
    loss = torch.mean((d_in + d_bg * dev(bg_mask) - dev(gt)) ** 2)
    loss.backward()

    seq = torch.nn.Sequential(*sum([[torch.nn.Linear(63 if k == 0 else 64, 4 if k == 7 else 64).double()]
                                    + ([torch.nn.ReLU()] if k < 7 else []) for k in range(8)], []))
    off = 0
    with torch.no_grad():
        for lin in seq[::2]:
            no, ni = lin.weight.shape
            lin.weight.copy_(torch.tensor(p.astype(np.float32)[off:off + no * ni].astype(np.float64)).reshape(no, ni))
            lin.bias.copy_(torch.tensor(p.astype(np.float32)[off + no * ni:off + no * ni + no].astype(np.float64)))
            off += no * ni + no
    Zs, Ys, Xs = torch.meshgrid(torch.arange(D), torch.arange(D), torch.arange(D), indexing="ij")
    regular = (torch.stack([Xs, Ys, Zs], -1).double() + 0.5) * vs + torch.tensor(mc, dtype=torch.float64)
    ne = torch.where(torch.tensor(flags) != -1)
    raw = seq(_pe((regular[ne] - torch.tensor(gc, dtype=torch.float64)) / gs))
    voxels = torch.zeros((D, D, D, 4), dtype=torch.float64)
    voxels[ne] = torch.cat([torch.sigmoid(raw[:, :3]), torch.relu(raw[:, 3:])], 1)
    grids = (((torch.tensor(pts[:, :S], dtype=torch.float64) - torch.tensor(mc, dtype=torch.float64)) / vs - 0.5)
             / (D - 1) * 2 - 1)[None, None]
    rgba = torch.nn.functional.grid_sample(voxels[None].permute(0, 4, 1, 2, 3), grids, padding_mode="zeros",
                                           align_corners=True)[0, :, 0].permute(1, 2, 0)
    raw_bg = seq(_pe((torch.tensor(pts[:, S:], dtype=torch.float64) - torch.tensor(gc, dtype=torch.float64)) / gs))

    def comp(c, s, dd, init):
        a = 1 - torch.exp(-s * torch.tensor(dd))
        T = torch.cumprod(torch.cat([init, 1 - a + 1e-10], 1), 1)[:, :-1]
        return (a[..., None] * T[..., None] * c).sum(1), T

    c_in, T_in = comp(rgba[..., :3], rgba[..., 3], dists[:, :S], torch.ones(B, 1, dtype=torch.float64))
    c_bg, _ = comp(torch.sigmoid(raw_bg[..., :3]), torch.relu(raw_bg[..., 3]), dists[:, S:], T_in[:, -1:])
    loss64 = torch.mean((c_in + c_bg * torch.tensor(bg_mask) - torch.tensor(gt)) ** 2)
    loss64.backward()
    g64 = torch.cat([q.grad.reshape(-1) for q in seq.parameters()]).numpy()
    assert abs(loss.item() - loss64.item()) <= 1e-5 * loss64.item()
    assert rel(m.params.grad.cpu().numpy(), g64) < 1e-3


def test_diffuse_head_scatter_skips_out_of_range_cells(oracle_mod):
    """The bake's scatter writes row cells[k] of a zeroed grid; indices outside the grid are skipped
    and read a zero gradient back."""
    rng = np.random.default_rng(3)
    raw = rng.normal(size=(6, 4))
    cells = torch.tensor([0, 7, -1, 3, 8, 5], dtype=torch.int32, device=DEV)   # -1 and 8 lie outside 8 rows
    t = dev(raw).requires_grad_(True)
    out = TR._DiffuseHead.apply(t, cells, 8)
    g = rng.normal(size=(8, 4))
    out.backward(dev(g))
    ro, _ = oracle_mod.diffuse_head(raw.astype(np.float32), np.zeros((6, 4)))
    exp = np.zeros((8, 4))
    for k, c in enumerate([0, 7, -1, 3, 8, 5]):
        if 0 <= c < 8:
            exp[c] = ro[k]
    np.testing.assert_allclose(out.detach().cpu().numpy(), exp, rtol=1e-6, atol=1e-7)
    gr = t.grad.cpu().numpy()
    assert np.all(gr[2] == 0) and np.all(gr[4] == 0) and np.any(gr[0] != 0)
