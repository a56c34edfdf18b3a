"""GPU parity of the reference-format reflection MLP (inference_mlp on FP16 MFMA) and of the whole
reflection pass (inverse_z_sampling -> inference_mlp -> integrate_points, 4 chunks of 16 samples)
against the FP64 oracle.

Tolerances: the oracle with quant16 rounds the 63 inputs and every hidden activation to FP16 where
the kernel's MFMA operands are FP16, so the two differ only by FP32 vs FP64 accumulation and the
FP32 sin/cos recurrence (~1e-6), which flip an occasional FP16 rounding by one unit (~1e-3
relative); rgb (sigmoid outputs) within 5e-3 absolute, sigma within 5e-3 (1 + sigma); the mean
error at least 20x below those bounds."""
import numpy as np
import pytest
import torch

from snisr import fastrendering as FR
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
RGB_TOL = 5e-3
SIG_TOL = 5e-3
SENTINEL = -7.0


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return t if dtype is None else t.to(dtype)


def check_close(got, ref, live):
    g, r = got[live], ref[live]
    e_rgb = np.abs(g[:, :3] - r[:, :3])
    e_sig = np.abs(g[:, 3] - r[:, 3]) / (1.0 + np.abs(r[:, 3]))
    assert e_rgb.max() <= RGB_TOL, e_rgb.max()
    assert e_sig.max() <= SIG_TOL, e_sig.max()
    assert e_rgb.mean() <= RGB_TOL / 20 and e_sig.mean() <= SIG_TOL / 20, (e_rgb.mean(), e_sig.mean())
    assert (got[~live] == SENTINEL).all()


def run_mlp(fr, qpi, bnet, bps, T, samples, spr, rays_view):
    DH, DW = fr.net_idx.shape
    res = torch.full((DH, DW, spr, 4), SENTINEL, device=DEV)
    FR.inference_mlp(spr, len(bnet), 256, 1.0 / fr.far, fr.trans_stop, to_dev(bnet), to_dev(T), to_dev(qpi),
                     to_dev(bps), to_dev(fr.centers), to_dev(samples), rays_view, to_dev(fr.params), res)
    return res.cpu().numpy().reshape(-1, 4)


def live_mask(fr, bnet, bps, qpi, T, spr):
    live = np.zeros(T.size * spr, bool)
    for b in range(len(bnet)):
        if bnet[b] < 0:
            continue
        ori = qpi[bps[b]:bps[b + 1]]
        ori = ori[~(T[ori] < fr.trans_stop)]
        live[(ori[:, None] * spr + np.arange(spr)).ravel()] = True
    return live


@pytest.mark.parametrize("H,W,groups,spr", [(96, 160, 6, 16), (40, 70, 3, 5), (2, 2, 1, 16)])
def test_inference_mlp_vs_oracle(oracle_mod, H, W, groups, spr):
    fr = S.reflection_frame(seed=H + W, H=H, W=W, n_groups=groups)
    DH, DW = fr.net_idx.shape
    rng = np.random.default_rng(H * W)
    qpi, bnet, bps = S.reflection_blocks(fr.net_idx)
    T = rng.uniform(0.0, 1.0, DH * DW).astype(np.float32)      # some pixels below trans_stop
    samples = (fr.origin + rng.uniform(-4, 4, (DH * DW * spr, 3))).astype(np.float32)
    rays_full = to_dev(fr.rays_d)
    got = run_mlp(fr, qpi, bnet, bps, T, samples, spr, rays_full[::2, ::2])
    ref = np.full_like(got, SENTINEL)
    oracle_mod.inference_mlp(spr, bnet, bps, qpi, T, fr.centers, samples, fr.rays_d[::2, ::2], fr.params,
                             1.0 / fr.far, fr.trans_stop, ref, quant16=True)
    live = live_mask(fr, bnet, bps, qpi, T, spr)
    assert live.any()
    check_close(got, ref, live)


def test_inference_mlp_full_frame_sampled_blocks(oracle_mod):
    """720p frame (360x640 downscaled, 48 groups): the GPU evaluates every block; the oracle checks
    60 seeded blocks (blocks are independent)."""
    fr = S.reflection_frame()
    DH, DW = fr.net_idx.shape
    spr = 16
    rng = np.random.default_rng(11)
    qpi, bnet, bps = S.reflection_blocks(fr.net_idx)
    T = np.where(rng.random(DH * DW) < 0.1, 0.0, 1.0).astype(np.float32)
    samples = (fr.origin + rng.uniform(-3, 3, (DH * DW * spr, 3))).astype(np.float32)
    got = run_mlp(fr, qpi, bnet, bps, T, samples, spr, to_dev(fr.rays_d)[::2, ::2])
    pick = np.sort(rng.choice(np.flatnonzero(bnet >= 0), 60, replace=False))
    sub_net = bnet[pick]
    sub_ps = np.empty(2 * len(pick), np.int32)
    # block k of the subset covers [bps[pick[k]], bps[pick[k]+1]); lay the pairs out as 2k, 2k+1
    # with an empty dummy block (net -1) in between
    sub_ps[0::2] = bps[pick]
    sub_ps[1::2] = bps[pick + 1]
    net2 = np.full(2 * len(pick) - 1, -1, np.int16)
    net2[0::2] = sub_net
    ref = np.full_like(got, SENTINEL)
    oracle_mod.inference_mlp(spr, net2, sub_ps, qpi, T, fr.centers, samples, fr.rays_d[::2, ::2], fr.params,
                             1.0 / fr.far, fr.trans_stop, ref, quant16=True)
    chk = np.zeros(DH * DW * spr, bool)
    for b in pick:
        ori = qpi[bps[b]:bps[b + 1]]
        ori = ori[~(T[ori] < fr.trans_stop)]
        chk[(ori[:, None] * spr + np.arange(spr)).ravel()] = True
    assert chk.sum() > 20000
    g, r = got[chk], ref[chk]
    assert np.abs(g[:, :3] - r[:, :3]).max() <= RGB_TOL
    assert (np.abs(g[:, 3] - r[:, 3]) / (1 + np.abs(r[:, 3]))).max() <= SIG_TOL
    # every live sample of the frame was written, every skipped one left alone
    all_live = live_mask(fr, bnet, bps, qpi, T, spr)
    assert (got[all_live][:, :3] >= 0).all() and (got[~all_live] == SENTINEL).all()


def test_reflection_pass_matches_oracle(oracle_mod):
    """The second half of FRender.render (rendering/src/__init__.py:272-321) at 192x320 (96x160
    downscaled) through the FASTRENDERING stand-in: sort_by_key, assign_pixels, then 4 x
    (inverse_z_sampling, inference_mlp, integrate_points); the oracle runs the same chain."""
    fr = S.reflection_frame(seed=21, H=192, W=320, n_groups=8)
    DH, DW = fr.net_idx.shape
    npx = DH * DW
    rays_full = to_dev(fr.rays_d)
    inv_near = fr.inverse_near
    # pixel regrouping on the GPU (bit-exact against the oracle in test_gpu_fastrendering)
    net = to_dev(fr.net_idx.ravel())
    q = torch.arange(npx, dtype=torch.int32, device=DEV)
    ps = torch.arange(npx, dtype=torch.int32, device=DEV)
    m = FR.sort_by_key(net, q, ps)
    tail = torch.full((1,), npx, dtype=torch.int32, device=DEV)
    per = torch.cat([ps[1:m], tail]) - ps[:m]
    pst = torch.cat([ps[:m], tail])
    bst = torch.cat([torch.zeros(1, dtype=torch.int32, device=DEV), torch.cumsum(torch.ceil(per / 64), 0).int()])
    nb = int(bst[-1])
    new_net = torch.empty(nb, dtype=torch.int16, device=DEV)
    new_ps = torch.empty(nb, dtype=torch.int32, device=DEV)
    FR.assign_pixels(net[:m], pst, bst, 256, new_net, new_ps)
    new_ps = torch.cat([new_ps, tail])
    qpi_h, bnet_h, bps_h = S.reflection_blocks(fr.net_idx)
    assert np.array_equal(q.cpu().numpy(), qpi_h) and np.array_equal(new_net.cpu().numpy(), bnet_h)
    assert np.array_equal(new_ps.cpu().numpy(), bps_h)
    far = np.float32(fr.far)
    sample_far = (np.float32(1.0) / inv_near + far).astype(np.float32)
    inv_bound = ((np.float32(1.0) / sample_far - inv_near) / np.float32(63)).astype(np.float32)
    T = torch.ones((DH, DW, 1), device=DEV)
    col = torch.zeros((DH, DW, 3), device=DEV)
    T_ref = np.ones(npx, np.float32)
    col_ref = np.zeros((npx, 3), np.float32)
    dirs_ds = np.ascontiguousarray(fr.rays_d[::2, ::2]).reshape(-1, 3)
    for step in range(0, 64, 16):
        pts = torch.full((DH, DW, 16, 3), -1.0, device=DEV)
        st = torch.zeros((DH, DW, 16), device=DEV)
        FR.inverse_z_sampling(to_dev(fr.origin), rays_full[::2, ::2], to_dev(inv_near), to_dev(inv_bound), 256,
                              step, 16, pts, st)
        res = torch.zeros((DH, DW, 16, 4), device=DEV)
        FR.inference_mlp(16, nb, 256, 1.0 / fr.far, fr.trans_stop, new_net, T, q, new_ps, to_dev(fr.centers), pts,
                         rays_full[::2, ::2], to_dev(fr.params), res)
        FR.integrate_points(res, st, 512, fr.trans_stop, T, col)
        rp, rs = oracle_mod.inverse_z_sampling(fr.origin, dirs_ds, inv_near, inv_bound, step, 16)
        assert np.array_equal(pts.cpu().numpy().reshape(-1, 3), rp)
        rres = np.zeros((npx * 16, 4), np.float32)
        oracle_mod.inference_mlp(16, bnet_h, bps_h, qpi_h, T_ref, fr.centers, rp, fr.rays_d[::2, ::2], fr.params,
                                 1.0 / fr.far, fr.trans_stop, rres, quant16=True)
        oracle_mod.integrate_points(rres, rs, np.float32(fr.trans_stop), T_ref, col_ref)
        # transparency decides which pixels the next chunk evaluates: compare before continuing
        Tg = T.cpu().numpy().ravel()
        assert np.abs(Tg - T_ref).max() <= 2e-2, (step, np.abs(Tg - T_ref).max())
    cg = col.cpu().numpy().reshape(-1, 3)
    err = np.abs(cg - col_ref)
    assert err.max() <= 2e-2 and err.mean() <= 1e-3, (err.max(), err.mean())


def test_inference_mlp_blocks_with_every_pixel_stopped(oracle_mod):
    """Pixels of the first half of the query order (whole blocks) are below trans_stop, so those
    blocks have no live column at all; the live blocks match the oracle and no stopped sample is
    written."""
    fr = S.reflection_frame(seed=77, H=64, W=96, n_groups=4)
    DH, DW = fr.net_idx.shape
    spr = 7
    rng = np.random.default_rng(77)
    qpi, bnet, bps = S.reflection_blocks(fr.net_idx)
    T = rng.uniform(0.5, 1.0, DH * DW).astype(np.float32)
    stopped = np.asarray(qpi[: len(qpi) // 2])
    T[stopped] = 0.0                                            # every pixel of these blocks stops
    samples = (fr.origin + rng.uniform(-4, 4, (DH * DW * spr, 3))).astype(np.float32)
    got = run_mlp(fr, qpi, bnet, bps, T, samples, spr, to_dev(fr.rays_d)[::2, ::2])
    ref = np.full_like(got, SENTINEL)
    oracle_mod.inference_mlp(spr, bnet, bps, qpi, T, fr.centers, samples, fr.rays_d[::2, ::2], fr.params,
                             1.0 / fr.far, fr.trans_stop, ref, quant16=True)
    live = live_mask(fr, bnet, bps, qpi, T, spr)
    assert live.any() and not live.all()
    check_close(got, ref, live)
    dead = (T.reshape(-1)[:, None] < fr.trans_stop).repeat(spr, 1).reshape(-1)
    assert np.all(got.reshape(-1, 4)[dead] == SENTINEL)


def test_packed_weights_follow_the_parameter_tensor():
    """Two same-sized parameter tensors used one after the other, the first freed before the second
    is made (so the caching allocator hands the second the first's memory): the second frame is
    rendered with its own weights, equal to a render after the packing cache is cleared."""
    H, W, spr = 40, 70, 5
    fr = S.reflection_frame(seed=21, H=H, W=W, n_groups=3)
    DH, DW = fr.net_idx.shape
    rng = np.random.default_rng(22)
    qpi, bnet, bps = S.reflection_blocks(fr.net_idx)
    T = np.ones(DH * DW, np.float32)
    samples = (fr.origin + rng.uniform(-4, 4, (DH * DW * spr, 3))).astype(np.float32)
    rays = to_dev(fr.rays_d)[::2, ::2]

    def render(params):
        res = torch.full((DH, DW, spr, 4), SENTINEL, device=DEV)
        FR.inference_mlp(spr, len(bnet), 256, 1.0 / fr.far, fr.trans_stop, to_dev(bnet), to_dev(T), to_dev(qpi),
                         to_dev(bps), to_dev(fr.centers), to_dev(samples), rays, params, res)
        torch.cuda.synchronize()
        return res

    a = to_dev(fr.params)
    first = render(a)
    addr = a.data_ptr()
    del a
    b = to_dev(fr.params) * torch.tensor(-0.5, dtype=torch.float16, device=DEV)   # different weights
    second = render(b)
    FR._packed_cache.clear()
    fresh = render(b)
    assert torch.equal(second, fresh)
    assert not torch.equal(second, first)
    print("second tensor reused the first's memory:", b.data_ptr() == addr)
