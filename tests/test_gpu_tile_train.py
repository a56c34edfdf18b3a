"""The reference's tile training step (Tile.train_batch, training/scripts/tile/tile.py:631-770) on the
HIP ops through ``snisr.tile_train``: samplers, compute_voxels bake, grid samples, background
samples, chained composites, reflection samples through the reflection MLP with SH, and the clamp.
Its prediction and both MLPs' parameter gradients (of a masked MSE) are compared with the same step
evaluated in FP64 torch (nn.Linear layers, F.grid_sample, the compositing formula, SH from the
oracle's basis) on the samplers' own depths (sampling runs without gradients in the reference)."""
import numpy as np
import pytest
import torch

from snisr import tile_train as TT
from snisr import train_ref as TR

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def _pe(x, L=10):
    return torch.cat([x] + [f(2.0 ** i * x) for i in range(L) for f in (torch.sin, torch.cos)], -1)


def _params(rng, depth, n_out):
    parts = []
    for k in range(depth):
        ni, no = (63 if k == 0 else 64), (n_out if k == depth - 1 else 64)
        parts += [rng.normal(0, 1.2 / np.sqrt(ni), no * ni), rng.normal(0, 0.1, no)]
    return np.concatenate(parts).astype(np.float32)


def _seq64(p, depth, n_out, act):
    layers, off = [], 0
    for k in range(depth):
        ni, no = (63 if k == 0 else 64), (n_out if k == depth - 1 else 64)
        lin = torch.nn.Linear(ni, no).double()
        with torch.no_grad():
            lin.weight.copy_(torch.tensor(p[off:off + no * ni], dtype=torch.float64).reshape(no, ni))
            lin.bias.copy_(torch.tensor(p[off + no * ni:off + no * ni + no], dtype=torch.float64))
        off += no * ni + no
        layers += [lin] + ([act] if k < depth - 1 else [])
    return torch.nn.Sequential(*layers)


def _batch(rng, tc, tile_size, B, P):
    n = B * P * P
    o = tc + rng.normal(size=(n, 3)) / np.sqrt(3) * 1.6 * tile_size
    tgt = tc + rng.uniform(-0.45, 0.45, (n, 3)) * tile_size
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # every ray has a background depth: with -1 (no background) the reference's background depths stay
    # -1 and its interval check (tile.py:701) fails unless the ray takes no tile sample either
    bgdepth = np.linalg.norm(tgt - o, axis=1, keepdims=True) + tile_size * rng.uniform(0.6, 1.5, (n, 1))
    gt = rng.uniform(0, 1, (n, 3))
    gtd = rng.uniform(0, 1, (n, 3))
    mask = np.ones((n, 1))
    return np.concatenate([o, d, bgdepth, gt, gtd, mask], 1).reshape(B, P, P, 14).astype(np.float32)


@pytest.mark.parametrize("train_flag", [2, 0, 1])
def test_tile_train_step_matches_fp64(oracle_mod, train_flag):
    rng = np.random.default_rng(40 + train_flag)
    NV, vs, B, P, NS, NB, RS, far = 16, 0.05, 2, 16, 32, 16, 32, 4.0
    D, tile_size = NV + 2, NV * vs
    tc = np.array([0.3, -0.2, 0.5])
    flags = np.where(rng.uniform(0, 1, (D, D, D)) < 0.35, 1, -1).astype(np.int16)
    tile = TT.TileState(nodes_flag=torch.tensor(flags, device=DEV), num_voxel=NV, voxel_size=vs, tile_center=tuple(tc),
                        group_center=(0.0, 0.1, 0.2), group_size=2.0, far=far)
    ps, pr = _params(rng, 8, 4), _params(rng, 6, 28)
    surf, refl = TR.PEMLP(8, 4, 0.0, device=DEV), TR.PEMLP(6, 28, 0.01, device=DEV)
    with torch.no_grad():
        surf.params.copy_(torch.tensor(ps, device=DEV))
        refl.params.copy_(torch.tensor(pr, device=DEV))
    batch = torch.tensor(_batch(rng, tc, tile_size, B, P), device=DEV)

    # the samplers' depths, exactly as the step draws them (no gradients in the reference either)
    rays_o = batch[..., :3].reshape(-1, 3).contiguous()
    rays_d = batch[..., 3:6].reshape(-1, 3).contiguous()
    bgdepth = batch[..., 6:7].reshape(-1, 1).contiguous()
    inner = tile.nodes_flag[1:-1, 1:-1, 1:-1]
    z, sd = TT.sparse_sampling(rays_o, rays_d, inner, NV, tc, tile_size, NS)
    bz = TT.sampling_background(rays_o, rays_d, bgdepth, tc, tile_size, NB, 0.5)

    out = TT.train_batch_forward(tile, surf, refl, batch, train_flag, NS, NB, RS, 0.5, P)
    assert out is not None
    m = out["mask"]
    assert float(m.sum()) > 100
    loss = (((out["pred"] - out["gtcolor"]) ** 2) * m).sum() / m.sum()
    loss.backward()
    rz = None
    if train_flag != 0:
        rz = TT.reflection_sampling(rays_o, rays_d, inner, tile.voxels, NV, tc, tile_size, RS, far, 1.0)

    # ---- FP64 evaluation of the same step
    f64 = lambda t: t.detach().double().cpu()  # noqa: E731
    o64, d64 = f64(rays_o), f64(rays_d)
    invalid = (m.reshape(-1) == 0).cpu()
    surf64 = _seq64(ps, 8, 4, torch.nn.ReLU())
    refl64 = _seq64(pr, 6, 28, torch.nn.LeakyReLU(0.01))
    mc = torch.tensor(np.asarray(tile.min_corner), dtype=torch.float64)
    gc, gs = torch.tensor(tile.group_center, dtype=torch.float64), tile.group_size
    n = o64.shape[0]

    def comp(c, s, dd, init):
        a = 1 - torch.exp(-s * dd)
        T = torch.cumprod(torch.cat([init, 1 - a + 1e-10], 1), 1)[:, :-1]
        return (a[..., None] * T[..., None] * c).sum(1), T

    Zs, Ys, Xs = torch.meshgrid(torch.arange(D), torch.arange(D), torch.arange(D), indexing="ij")
    regular = (torch.stack([Xs, Ys, Zs], -1).double() + 0.5) * vs + mc
    ne = torch.where(torch.tensor(flags) != -1)
    raw = surf64(_pe((regular[ne] - gc) / gs))
    vox = torch.zeros((D, D, D, 4), dtype=torch.float64)
    vox[ne] = torch.cat([torch.sigmoid(raw[:, :3]), torch.relu(raw[:, 3:])], 1)
    smp = o64[:, None] + f64(z)[..., None] * d64[:, None]
    grids = (((smp - mc) / vs - 0.5) / (D - 1) * 2 - 1)[None, None]
    rgba = torch.nn.functional.grid_sample(vox[None].permute(0, 4, 1, 2, 3), grids, padding_mode="zeros",
                                           align_corners=True)[0, :, 0].permute(1, 2, 0)
    bgs = o64[:, None] + f64(bz)[..., None] * d64[:, None]
    rb = surf64(_pe((bgs - gc) / gs))
    tz = torch.cat([f64(z), f64(bz)], 1)
    dists = torch.cat([tz[:, 1:] - tz[:, :-1], torch.full((n, 1), 1e10, dtype=torch.float64)], 1)
    dists[:, :NS] = f64(sd)
    dists[invalid] = 0
    c_in, T_in = comp(rgba[..., :3], rgba[..., 3], dists[:, :NS], torch.ones(n, 1, dtype=torch.float64))
    c_bg, _ = comp(torch.sigmoid(rb[..., :3]), torch.relu(rb[..., 3]), dists[:, NS:], T_in[:, -1:])
    diffuse = c_in + c_bg * (f64(bgdepth) != -1)
    if train_flag == 0:
        pred = diffuse
    else:
        rs = o64[:, None] + f64(rz)[..., None] * d64[:, None]
        rr = refl64(_pe((rs - gc) / far))
        u = rays_d.double().cpu().numpy()
        u = u / np.linalg.norm(u, axis=1, keepdims=True)
        Y = torch.tensor(oracle_mod.sh2_basis(u))
        spec = torch.sigmoid(torch.einsum("bj,bsjc->bsc", Y, rr[..., :27].reshape(n, RS, 9, 3)))
        rd = torch.cat([f64(rz)[:, 1:] - f64(rz)[:, :-1], torch.full((n, 1), 1e10, dtype=torch.float64)], 1)
        rd[invalid] = 0
        rgb_vd, _ = comp(spec, torch.relu(rr[..., 27]), rd, torch.ones(n, 1, dtype=torch.float64))
        rmask = torch.all(f64(rz) != -1, -1, keepdim=True)
        base = diffuse.detach() if train_flag == 1 else diffuse
        pred = torch.clamp(base + rmask * rgb_vd, 0, 1)
    m64 = f64(m).reshape(n, 1)
    loss64 = (((pred - f64(out["gtcolor"]).reshape(n, 3)) ** 2) * m64).sum() / m64.sum()
    loss64.backward()

    np.testing.assert_allclose(out["pred"].detach().cpu().numpy().reshape(n, 3), pred.detach().numpy(), atol=2e-4)
    assert abs(loss.item() - loss64.item()) <= 1e-4 * loss64.item()
    if train_flag in (0, 2):
        g = torch.cat([q.grad.reshape(-1) for q in surf64.parameters()]).numpy()
        assert _rel(surf.params.grad.cpu().numpy(), g) < 1e-3
    else:
        assert surf.params.grad is None or float(surf.params.grad.abs().sum()) == 0.0
    if train_flag in (1, 2):
        g = torch.cat([q.grad.reshape(-1) for q in refl64.parameters()]).numpy()
        assert _rel(refl.params.grad.cpu().numpy(), g) < 1e-3


def test_tile_state_recomputes_cells_when_flags_change():
    flags = torch.full((6, 6, 6), -1, dtype=torch.int16, device=DEV)
    flags[1, 2, 3] = 1
    tile = TT.TileState(nodes_flag=flags, num_voxel=4, voxel_size=0.1, tile_center=(0, 0, 0), group_center=(0, 0, 0),
                        group_size=1.0, far=2.0)
    assert tile.cells().cpu().tolist() == [1 * 36 + 2 * 6 + 3]
    flags[4, 0, 1] = 0                                    # written in place
    assert tile.cells().cpu().tolist() == [1 * 36 + 2 * 6 + 3, 4 * 36 + 0 * 6 + 1]
    fresh = torch.full((6, 6, 6), -1, dtype=torch.int16, device=DEV)
    fresh[5, 5, 5] = 1
    tile.nodes_flag = fresh                               # replaced, as the reference's pruning does
    assert tile.cells().cpu().tolist() == [215]
