"""Pins of the oracle's training-form compositing (utils.volume_rendering) and specular head
(sigmoid of degree-2 SH) and of their adjoints: closed forms, torch autograd in FP64 as an
independent gradient, and central finite differences."""
import numpy as np
import pytest
import torch


def _ref_volume_rendering(rgb, sigma, dists, init_T):
    """torch FP64 evaluation of the composite (exclusive cumprod of 1 - alpha + 1e-10)."""
    alpha = 1.0 - torch.exp(-sigma * dists)
    B = rgb.shape[0]
    lead = init_T.reshape(B, 1)
    T = torch.cumprod(torch.cat([lead, 1.0 - alpha + 1e-10], 1), 1)[:, :-1]
    return (alpha[..., None] * T[..., None] * rgb).sum(1), T, alpha


def test_constant_sigma_closed_form(oracle_mod):
    B, n = 4, 9
    rng = np.random.default_rng(1)
    sig = np.repeat(rng.uniform(0.5, 4.0, (B, 1)), n, 1)
    dst = np.full((B, n), 0.07)
    init = rng.uniform(0.3, 1.0, B)
    col = rng.uniform(0, 1, (B, 3))
    rgb = np.repeat(col[:, None, :], n, 1)
    r = oracle_mod.volume_render(rgb, sig, dst, init_T=init)
    e = np.exp(-sig[:, 0] * 0.07)
    T = init[:, None] * (e[:, None] + 1e-10) ** np.arange(n)[None, :]
    np.testing.assert_allclose(r["trans"], T, rtol=1e-12)
    np.testing.assert_allclose(r["alpha"], np.repeat((1 - e)[:, None], n, 1), rtol=1e-12)
    np.testing.assert_allclose(r["out"], col * ((1 - e) * T.sum(1))[:, None], rtol=1e-12)


@pytest.mark.parametrize("with_init", [False, True])
def test_volume_render_adjoint_vs_torch_autograd(oracle_mod, with_init):
    B, n = 5, 12
    rng = np.random.default_rng(2 + with_init)
    rgb, sig, dst = rng.uniform(0, 1, (B, n, 3)), rng.uniform(0, 6, (B, n)), rng.uniform(0, 0.4, (B, n))
    dst[:, -1] = 1e10      # the reference's last interval
    init = rng.uniform(0.1, 1.0, B) if with_init else np.ones(B)
    g_out, g_tr, g_al = rng.normal(size=(B, 3)), rng.normal(size=(B, n)), rng.normal(size=(B, n))
    r = oracle_mod.volume_render(rgb, sig, dst, init_T=init if with_init else None, g_out=g_out, g_trans=g_tr,
                                 g_alpha=g_al)
    t = [torch.tensor(a, dtype=torch.float64, requires_grad=True) for a in (rgb, sig, init)]
    out, T, alpha = _ref_volume_rendering(t[0], t[1], torch.tensor(dst), t[2])
    L = (out * torch.tensor(g_out)).sum() + (T * torch.tensor(g_tr)).sum() + (alpha * torch.tensor(g_al)).sum()
    L.backward()
    np.testing.assert_allclose(r["out"], out.detach().numpy(), rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(r["g_rgb"], t[0].grad.numpy(), rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(r["g_sigma"], t[1].grad.numpy(), rtol=1e-9, atol=1e-12)
    if with_init:
        np.testing.assert_allclose(r["g_init"], t[2].grad.numpy(), rtol=1e-9, atol=1e-12)


def test_volume_render_adjoint_vs_finite_differences(oracle_mod):
    B, n = 2, 6
    rng = np.random.default_rng(4)
    rgb, sig, dst = rng.uniform(0, 1, (B, n, 3)), rng.uniform(0.1, 3, (B, n)), rng.uniform(0.05, 0.3, (B, n))
    init = rng.uniform(0.2, 1.0, B)
    g_out = rng.normal(size=(B, 3))
    r = oracle_mod.volume_render(rgb, sig, dst, init_T=init, g_out=g_out)
    L = lambda rg, sg, it: float((oracle_mod.volume_render(rg, sg, dst, init_T=it)["out"] * g_out).sum())
    eps = 1e-6
    for (b, i) in [(0, 0), (0, 3), (1, 5)]:
        sp, sm = sig.copy(), sig.copy()
        sp[b, i] += eps
        sm[b, i] -= eps
        fd = (L(rgb, sp, init) - L(rgb, sm, init)) / (2 * eps)
        assert abs(fd - r["g_sigma"][b, i]) <= 1e-7 * max(1.0, abs(fd))
    for b in range(B):
        ip, im = init.copy(), init.copy()
        ip[b] += eps
        im[b] -= eps
        fd = (L(rgb, sig, ip) - L(rgb, sig, im)) / (2 * eps)
        assert abs(fd - r["g_init"][b]) <= 1e-7 * max(1.0, abs(fd))


@pytest.mark.parametrize("deg", [0, 1, 2])
def test_sh_sigmoid_forward_and_adjoint(oracle_mod, deg):
    rng = np.random.default_rng(5 + deg)
    n_rays, spr, K = 3, 4, 28
    q = rng.normal(size=(n_rays * spr, K))
    dirs = rng.normal(size=(n_rays, 3))
    g = rng.normal(size=(n_rays * spr, 3))
    spec, gq = oracle_mod.sh_sigmoid(q, dirs, deg, g_spec=g)
    u = dirs / np.linalg.norm(dirs, axis=1, keepdims=True)
    Y = oracle_mod.sh2_basis(np.repeat(u, spr, 0))      # pinned against scipy's complex SH
    nb = (deg + 1) ** 2
    z = np.einsum("sb,sbc->sc", Y[:, :nb], q[:, :27].reshape(-1, 9, 3)[:, :nb])
    np.testing.assert_allclose(spec, 1 / (1 + np.exp(-z)), rtol=1e-12)
    assert not gq[:, 27].any() and not gq[:, 3 * nb:27].any()
    eps = 1e-6
    for (s, j) in [(0, 0), (5, 3 * nb - 1), (11, 4)]:
        if j >= 3 * nb:
            continue
        qp, qm = q.copy(), q.copy()
        qp[s, j] += eps
        qm[s, j] -= eps
        fd = ((oracle_mod.sh_sigmoid(qp, dirs, deg) - oracle_mod.sh_sigmoid(qm, dirs, deg)) * g).sum() / (2 * eps)
        assert abs(fd - gq[s, j]) <= 1e-7 * max(1.0, abs(fd))


def _torch_pemlp(params, pts, depth, n_out, slope):
    """networks.MLP in FP64 from torch's own layers (nn.Linear, LeakyReLU / ReLU, sin / cos), its
    weights taken from the flat vector in nn.Linear order; returns (out, the Linear layers)."""
    layers, off = [], 0
    for k in range(depth):
        ni, no = (63 if k == 0 else 64), (n_out if k == depth - 1 else 64)
        lin = torch.nn.Linear(ni, no).double()
        with torch.no_grad():
            lin.weight.copy_(torch.tensor(params[off:off + no * ni]).reshape(no, ni))
            lin.bias.copy_(torch.tensor(params[off + no * ni:off + no * ni + no]))
        off += no * ni + no
        layers.append(lin)
    assert off == params.size
    x = torch.tensor(pts)
    h = torch.cat([x] + [f(2.0 ** i * x) for i in range(10) for f in (torch.sin, torch.cos)], -1)
    act = torch.nn.ReLU() if slope == 0 else torch.nn.LeakyReLU(slope)
    for k, lin in enumerate(layers):
        h = lin(h)
        if k < depth - 1:
            h = act(h)
    return h, layers


def _pemlp_params(rng, depth, n_out):
    parts = []
    for k in range(depth):
        ni, no = (63 if k == 0 else 64), (n_out if k == depth - 1 else 64)
        parts += [rng.normal(0, 1.2 / np.sqrt(ni), no * ni), rng.normal(0, 0.1, no)]
    return np.concatenate(parts)


@pytest.mark.parametrize("depth,n_out,slope", [(6, 28, 0.01), (8, 4, 0.0), (2, 4, 0.2)])
def test_pemlp_vs_torch_layers_and_autograd(oracle_mod, depth, n_out, slope):
    rng = np.random.default_rng(depth * 31 + n_out)
    n = 37
    p = _pemlp_params(rng, depth, n_out)
    pts = rng.uniform(-1, 1, (n, 3))
    g_out = rng.normal(size=(n, n_out))
    out, gp = oracle_mod.pemlp(p, pts, depth, n_out, slope, g_out=g_out)
    ref, layers = _torch_pemlp(p, pts, depth, n_out, slope)
    (ref * torch.tensor(g_out)).sum().backward()
    gref = np.concatenate([t.grad.numpy().reshape(-1) for lin in layers for t in (lin.weight, lin.bias)])
    np.testing.assert_allclose(out, ref.detach().numpy(), rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(oracle_mod.pemlp(p, pts, depth, n_out, slope), out, rtol=0, atol=0)
    np.testing.assert_allclose(gp, gref, rtol=1e-9, atol=1e-11)


def test_pemlp_gradient_vs_finite_differences(oracle_mod):
    rng = np.random.default_rng(77)
    depth, n_out, slope = 3, 4, 0.05
    p = _pemlp_params(rng, depth, n_out)
    pts = rng.uniform(-1, 1, (5, 3))
    g_out = rng.normal(size=(5, n_out))
    _, gp = oracle_mod.pemlp(p, pts, depth, n_out, slope, g_out=g_out)
    h = 1e-6
    for k in rng.choice(p.size, 24, replace=False):
        e = np.zeros_like(p)
        e[k] = h
        fd = ((oracle_mod.pemlp(p + e, pts, depth, n_out, slope) - oracle_mod.pemlp(p - e, pts, depth, n_out, slope))
              * g_out).sum() / (2 * h)
        assert abs(fd - gp[k]) <= 1e-6 * max(1.0, abs(fd))


def test_pemlp_linear_closed_form(oracle_mod):
    """slope 1 makes the network affine: out = W_eff PE(x) + b_eff with W_eff the product of the
    layers' matrices."""
    rng = np.random.default_rng(5)
    depth, n_out = 4, 28
    p = _pemlp_params(rng, depth, n_out)
    pts = rng.uniform(-1, 1, (9, 3))
    W_eff, b_eff, off = np.eye(63), np.zeros(63), 0
    for k in range(depth):
        ni, no = (63 if k == 0 else 64), (n_out if k == depth - 1 else 64)
        Wk, bk = p[off:off + no * ni].reshape(no, ni), p[off + no * ni:off + no * ni + no]
        W_eff, b_eff = Wk @ W_eff, Wk @ b_eff + bk
        off += no * ni + no
    f = np.concatenate([pts] + [fn(2.0 ** i * pts) for i in range(10) for fn in (np.sin, np.cos)], 1)
    np.testing.assert_allclose(oracle_mod.pemlp(p, pts, depth, n_out, 1.0), f @ W_eff.T + b_eff, rtol=1e-10,
                               atol=1e-11)


def _torch_query_voxel(vox_cl, pts, min_corner, voxel_size):
    """Tile.query_voxel's torch form (tile.py:225-236) in FP64 on the channel-last grid's permuted view."""
    D = vox_cl.shape[0]
    voxels = vox_cl[None].permute(0, 4, 1, 2, 3)
    indices = (pts - min_corner) / voxel_size - 0.5
    grids = (indices / (D - 1) * 2 - 1)[None, None, None]
    rgba = torch.nn.functional.grid_sample(voxels, grids, padding_mode="zeros", align_corners=True)
    rgba = rgba[0, :, 0, 0].permute(1, 0)
    return rgba[:, :3], rgba[:, 3]


def test_voxel_query_vs_grid_sample_and_autograd(oracle_mod):
    rng = np.random.default_rng(21)
    D, vs = 7, 0.13
    mc = np.array([-0.3, 0.2, 1.1])
    vox = rng.uniform(0, 1, (D, D, D, 4))
    # inside, near the faces and outside (zero padding), plus exact cell centres
    pts = mc + rng.uniform(-0.5, D + 0.5, (300, 3)) * vs
    pts[:20] = mc + (rng.integers(0, D, (20, 3)) + 0.5) * vs
    g_rgb, g_sig = rng.normal(size=(300, 3)), rng.normal(size=300)
    r = oracle_mod.voxel_query(vox, pts, mc, vs, g_rgb=g_rgb, g_sigma=g_sig)
    tv = torch.tensor(vox, requires_grad=True)
    rgb, sig = _torch_query_voxel(tv, torch.tensor(pts), torch.tensor(mc), vs)
    ((rgb * torch.tensor(g_rgb)).sum() + (sig * torch.tensor(g_sig)).sum()).backward()
    np.testing.assert_allclose(r["rgb"], rgb.detach().numpy(), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(r["sigma"], sig.detach().numpy(), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(r["g_vox"], tv.grad.numpy(), rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(r["rgb"][:20], vox[tuple(((pts[:20] - mc) / vs - 0.5).round().astype(int)[:, ::-1].T)][:, :3],
                               rtol=1e-12)


def test_voxel_query_reproduces_a_linear_field(oracle_mod):
    """Trilinear interpolation is exact on an affine field: grid values a . centre + b give a . x + b
    at any point between the outermost cell centres."""
    D, vs = 6, 0.25
    mc = np.array([1.0, -2.0, 0.5])
    a, b = np.array([[0.3, -1.2, 2.0], [1.0, 0.0, 0.0], [0.0, 0.5, 0.5], [-0.7, 0.1, 0.9]]), np.array([0.1, 2.0, -1.0, 0.0])
    zz, yy, xx = np.meshgrid(np.arange(D), np.arange(D), np.arange(D), indexing="ij")
    centres = mc + (np.stack([xx, yy, zz], -1) + 0.5) * vs
    vox = centres @ a.T + b
    rng = np.random.default_rng(3)
    pts = mc + (0.5 + rng.uniform(0, D - 1, (200, 3))) * vs
    r = oracle_mod.voxel_query(vox, pts, mc, vs)
    exp = pts @ a.T + b
    np.testing.assert_allclose(r["rgb"], exp[:, :3], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(r["sigma"], exp[:, 3], rtol=1e-12, atol=1e-12)


def test_bake_points_and_cells_follow_the_reference_expressions(oracle_mod):
    """Tile.compute_voxels' inputs evaluated the reference's way in torch FP32 (create_meshgrid,
    (index + 0.5) * voxel_size + min_corner, torch.where(nodes_flag != -1), norm_diffuse_point)."""
    rng = np.random.default_rng(8)
    D, vs = 9, 0.0625
    mc = np.array([0.37, -1.25, 2.5], np.float32)
    center, scale = np.array([0.5, -0.75, 2.0], np.float32), 3.7
    flags = np.where(rng.uniform(0, 1, (D, D, D)) < 0.3, 1, -1).astype(np.int16)
    Zs, Ys, Xs = torch.meshgrid(torch.arange(D), torch.arange(D), torch.arange(D), indexing="ij")
    regular = (torch.stack([Xs, Ys, Zs], -1).float() + 0.5) * vs + torch.tensor(mc)
    ref = (regular[torch.where(torch.tensor(flags) != -1)] - torch.tensor(center)) / scale
    cells = oracle_mod.flag_cells(flags)
    assert cells.dtype == np.int32 and np.all(np.diff(cells) > 0)
    pts = oracle_mod.bake_points(cells, D, mc, vs, center, scale)
    np.testing.assert_array_equal(pts, ref.numpy())


def test_diffuse_head_adjoint_vs_autograd(oracle_mod):
    rng = np.random.default_rng(4)
    raw, g = rng.normal(size=(50, 4)) * 3, rng.normal(size=(50, 4))
    out, graw = oracle_mod.diffuse_head(raw, g)
    t = torch.tensor(raw, requires_grad=True)
    o = torch.cat([torch.sigmoid(t[:, :3]), torch.relu(t[:, 3:])], 1)
    (o * torch.tensor(g)).sum().backward()
    np.testing.assert_allclose(out, o.detach().numpy(), rtol=1e-14)
    np.testing.assert_allclose(graw, t.grad.numpy(), rtol=1e-13, atol=1e-15)
