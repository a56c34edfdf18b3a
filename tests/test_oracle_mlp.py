"""Pins of the oracle's reflection-MLP restatement (inference_mlp, rendering/src/rayreflection_render.cu)
against library routines: numpy's float16 rounding, scipy's spherical harmonics, and torch.nn
layers assembled from the reference's parameter packing."""
import math

import numpy as np
import pytest
import torch

from snisr import scene as S


def test_round_half_matches_numpy_float16(oracle_mod):
    rng = np.random.default_rng(0)
    v = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-9, 6, 4000),
                        [0.0, -0.0, 65504.0, 65519.99, 65520.0, -70000.0, 6e-8, 2.98e-8, 2.99e-8, 1e-9]])
    with np.errstate(over="ignore"):
        ref = v.astype(np.float16).astype(np.float64)
    got = np.array([oracle_mod.round_half(float(x)) for x in v])
    assert np.array_equal(got, ref)


def _real_sh_scipy(dirs):
    """Real SH up to degree 2 as sqrt(2) Re / Im of scipy's complex SH (Condon-Shortley phase);
    m < 0 takes the imaginary part of Y_l^|m|.  Order (l, m): (0,0) (1,-1) (1,0) (1,1) (2,-2) ...
    (2,2)."""
    from scipy.special import sph_harm_y
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    theta = np.arccos(np.clip(z, -1.0, 1.0))
    phi = np.arctan2(y, x)
    cols = []
    for l in range(3):
        for m in range(-l, l + 1):
            Y = sph_harm_y(l, abs(m), theta, phi)
            cols.append(Y.real if m == 0 else math.sqrt(2.0) * (Y.imag if m < 0 else Y.real))
    return np.stack(cols, -1)


def test_sh2_basis_matches_scipy(oracle_mod):
    rng = np.random.default_rng(1)
    d = rng.standard_normal((5000, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    got = oracle_mod.sh2_basis(d)
    assert np.allclose(got, _real_sh_scipy(d), rtol=0, atol=1e-12)


def test_sh2_basis_orthonormal(oracle_mod):
    """Gauss-Legendre in cos(theta) x uniform phi integrates the degree <= 4 products exactly."""
    mu, w = np.polynomial.legendre.leggauss(12)
    phi = np.arange(24) * (2 * np.pi / 24)
    M, P = np.meshgrid(mu, phi, indexing="ij")
    s = np.sqrt(1 - M ** 2)
    d = np.stack([s * np.cos(P), s * np.sin(P), M], -1).reshape(-1, 3)
    wt = (w[:, None] * np.full(24, 2 * np.pi / 24)[None, :]).reshape(-1)
    B = oracle_mod.sh2_basis(d)
    G = (B * wt[:, None]).T @ B
    assert np.allclose(G, np.eye(9), atol=1e-12)


@torch.no_grad()
def _torch_reference(params_f16, centers, inverse_far, samples, rays_d, net_idx, pixel_starts, qpi, T, trans_stop,
                     spr):
    """The reflection MLP written with torch.nn layers in float64: the packed vector is unpacked per
    layer as bias (out) then W as [in][out] (nn.Linear.weight = W^T); Positional_Encoding as in
    training/scripts/networks.py:4-16; SH colour with coefficient r = 3 * basis + channel."""
    dims = S.REF_MLP_DIMS
    P = torch.from_numpy(params_f16.astype(np.float64))
    H, W = rays_d.shape[0], rays_d.shape[1]
    out = {}
    for b in range(len(net_idx)):
        g = int(net_idx[b])
        if g < 0:
            continue
        layers, off = [], g * S.REF_MLP_PARAMS
        for i, o in dims:
            lin = torch.nn.Linear(i, o, dtype=torch.float64)
            with torch.no_grad():
                lin.bias.copy_(P[off:off + o])
                lin.weight.copy_(P[off + o:off + o + i * o].reshape(i, o).T)
            off += o + i * o
            layers.append(lin)
        for p in range(pixel_starts[b], pixel_starts[b + 1]):
            ori = int(qpi[p])
            if T[ori] < trans_stop:
                continue
            for j in range(spr):
                task = ori * spr + j
                x32 = (samples[task].astype(np.float32) - centers[g]) * np.float32(inverse_far)
                x = torch.from_numpy(x32.astype(np.float64))
                enc = [x]
                for k in range(10):
                    enc += [torch.sin(2.0 ** k * x), torch.cos(2.0 ** k * x)]
                h = torch.cat(enc)
                for li, lin in enumerate(layers):
                    h = lin(h)
                    if li < len(layers) - 1:
                        h = torch.nn.functional.leaky_relu(h, 0.01)
                dv = torch.from_numpy(rays_d[ori // W, ori % W].astype(np.float64))
                dv = dv / torch.linalg.norm(dv)
                basis = torch.from_numpy(_real_sh_scipy(dv.numpy()[None])[0])
                coef = h[:27].reshape(9, 3)
                rgb = torch.sigmoid((basis[:, None] * coef).sum(0))
                out[task] = np.concatenate([rgb.detach().numpy(), [max(float(h[27]), 0.0)]])
    return out


def _small_case(seed, DH=4, DW=6, n_groups=3, spr=4):
    rng = np.random.default_rng(seed)
    params, centers = S.make_reflection_groups(seed, n_groups, (0, 0, 0), (4, 4, 1))
    net = rng.integers(-1, n_groups, DH * DW).astype(np.int16)
    qpi, bnet, bps = S.reflection_blocks(net, max_pixels_per_block=3)
    samples = rng.uniform(0.0, 4.0, (DH * DW * spr, 3)).astype(np.float32)
    rays = rng.standard_normal((DH, DW, 3)).astype(np.float32)
    T = rng.uniform(0.0, 1.0, DH * DW).astype(np.float32)
    return params, centers, samples, rays, bnet, bps, qpi, T, spr


def test_inference_mlp_oracle_vs_torch_layers(oracle_mod):
    params, centers, samples, rays, bnet, bps, qpi, T, spr = _small_case(3)
    inv_far = 1.0 / 5.744562646538029
    res = np.full((samples.shape[0], 4), -7.0, np.float32)
    oracle_mod.inference_mlp(spr, bnet, bps, qpi, T, centers, samples, rays, params, inv_far, 0.2, res)
    ref = _torch_reference(params, centers, inv_far, samples, rays, bnet, bps, qpi, T, 0.2, spr)
    assert len(ref) > 10
    for task in range(res.shape[0]):
        if task in ref:
            assert np.allclose(res[task], ref[task], rtol=1e-6, atol=1e-6), task
        else:
            assert (res[task] == -7.0).all(), task   # skipped samples stay untouched


def test_inference_mlp_quant16_close_to_exact(oracle_mod):
    """FP16 rounding of inputs and activations moves outputs by at most a few 1e-3."""
    params, centers, samples, rays, bnet, bps, qpi, T, spr = _small_case(4, DH=8, DW=8, spr=8)
    a = np.zeros((samples.shape[0], 4), np.float32)
    b = np.zeros_like(a)
    oracle_mod.inference_mlp(spr, bnet, bps, qpi, T, centers, samples, rays, params, 0.2, 0.0, a)
    oracle_mod.inference_mlp(spr, bnet, bps, qpi, T, centers, samples, rays, params, 0.2, 0.0, b, quant16=True)
    err = np.abs(a - b)
    assert err[:, :3].max() < 1e-2 and err[:, :3].max() > 0
