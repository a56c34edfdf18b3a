"""The reference's training-form compositing and specular head as autograd Functions over the HIP
library (forward and hand-written adjoint kernels, include/snisr.h):

- ``volume_rendering(rgb, sigma, dists, init_T=None)`` is utils.volume_rendering
  (training/scripts/utils.py:823-837): returns (out [B, 3], transparency [B, N, 1], alpha [B, N, 1])
  with transparency the exclusive cumprod of (1 - alpha + 1e-10) started at init_T.  Passing the
  in-tile pass's ``transparency[:, -1:]`` as init_T chains the background pass behind it
  (training/scripts/tile/tile.py:703-704); the gradient flows into both passes.
- ``PEMLP`` is networks.MLP (training/scripts/networks.py:4-41) with L = 10, width 64: the surface
  MLP (depth 8, 4 outputs, ReLU) and the reflection MLP (depth 6, 28 outputs, LeakyReLU 0.01), its
  parameters one flat tensor in nn.Linear order, forward and backward on the HIP kernels.
- ``flag_cells``, ``bake_voxels``, ``inference_diffuse`` and ``query_voxel`` are the diffuse branch
  around the surface MLP: Tile.compute_voxels (training/scripts/tile/tile.py:283-295, the MLP at the
  occupied cells' centres, sigmoid / ReLU head, scattered into the dilated grid), Tile.inference_diffuse
  (tile.py:175-194, background samples) and Tile.query_voxel (tile.py:225-236, trilinear grid_sample
  of the baked grid at the in-tile samples), each with its adjoint, so the gradient of a diffuse
  training step reaches the MLP through the bake as in the reference.  The grid is [D, D, D, 4]
  channel-last, the memory of the reference's permuted [1, 4, D, D, D] voxels.
- ``sh_sigmoid(coeffs, viewdirs, deg)`` is the specular head of Tile.inference_specular
  (training/scripts/tile/tile.py:212-219): sigmoid(eval_sh(deg, coeffs as [..., 3, (deg+1)^2],
  normalised viewdirs)) for coefficients laid out [basis][rgb] as the MLP emits them; ``viewdirs``
  is one direction per ray, shared by the ray's samples.
"""
from __future__ import annotations

import torch

from . import _native as N
from .device import _ptr, _stream


def _f32(t: torch.Tensor, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32:
        raise ValueError(f"{name} must be a float32 CUDA (HIP) tensor")
    return t.contiguous()


class _VolumeRendering(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rgb, sigma, dists, init_T):
        B, n = sigma.shape[0], sigma.shape[1]
        rgb_c, sig_c, dst_c = _f32(rgb, "rgb"), _f32(sigma, "sigma"), _f32(dists, "dists")
        if rgb_c.numel() != 3 * B * n or dst_c.numel() != B * n or sig_c.numel() != B * n:
            raise ValueError("rgb [B, N, 3], sigma and dists [B, N(, 1)] must agree")
        it = None if init_T is None else _f32(init_T, "init_T")
        if it is not None and it.numel() != B:
            raise ValueError("init_T must hold one value per ray")
        dev = rgb.device
        out = torch.empty((B, 3), dtype=torch.float32, device=dev)
        trans = torch.empty((B, n, 1), dtype=torch.float32, device=dev)
        alpha = torch.empty((B, n, 1), dtype=torch.float32, device=dev)
        N.check(N.lib().snisr_volume_render_forward(_ptr(rgb_c), _ptr(sig_c), _ptr(dst_c), _ptr(it), B, n, _ptr(out),
                                                    _ptr(trans), _ptr(alpha), _stream()), "volume_render_forward")
        ctx.save_for_backward(rgb_c, dst_c, trans, alpha)
        ctx.init_shape = None if init_T is None else init_T.shape
        ctx.sigma_shape = sigma.shape
        return out, trans, alpha

    @staticmethod
    def backward(ctx, g_out, g_trans, g_alpha):
        rgb, dists, trans, alpha = ctx.saved_tensors
        B, n = trans.shape[0], trans.shape[1]
        dev = rgb.device
        g_out = torch.zeros((B, 3), dtype=torch.float32, device=dev) if g_out is None else g_out.contiguous()
        gt = None if g_trans is None else g_trans.contiguous()
        ga = None if g_alpha is None else g_alpha.contiguous()
        g_rgb = torch.empty_like(rgb)
        g_sig = torch.empty((B, n), dtype=torch.float32, device=dev)
        g_init = torch.empty(B, dtype=torch.float32, device=dev) if ctx.init_shape is not None else None
        N.check(N.lib().snisr_volume_render_backward(_ptr(rgb), _ptr(dists), _ptr(trans), _ptr(alpha), B, n,
                                                     _ptr(g_out), _ptr(gt), _ptr(ga), _ptr(g_rgb), _ptr(g_sig),
                                                     _ptr(g_init), _stream()), "volume_render_backward")
        g_init_out = None if g_init is None else g_init.reshape(ctx.init_shape)
        return g_rgb, g_sig.reshape(ctx.sigma_shape), None, g_init_out


def volume_rendering(rgb: torch.Tensor, sigma: torch.Tensor, dists: torch.Tensor, init_T=None):
    """utils.volume_rendering on the HIP kernels: rgb [B, N, 3], sigma and dists [B, N, 1] (or
    [B, N]); init_T a float (1 by default) or a [B, 1] tensor.  Returns (out, transparency, alpha)."""
    if init_T is not None and not isinstance(init_T, torch.Tensor):
        if float(init_T) == 1.0:
            init_T = None
        else:
            init_T = torch.full((sigma.shape[0], 1), float(init_T), dtype=torch.float32, device=sigma.device)
    return _VolumeRendering.apply(rgb, sigma, dists, init_T)


class _SHSigmoid(torch.autograd.Function):
    @staticmethod
    def forward(ctx, coeffs, dirs, deg, spr):
        q = _f32(coeffs, "coeffs")
        d = _f32(dirs, "viewdirs")
        n = q.numel() // q.shape[-1]
        stride = q.shape[-1]
        if d.numel() * spr != 3 * n:
            raise ValueError("viewdirs must hold one direction per ray of samples_per_ray samples")
        spec = torch.empty((*q.shape[:-1], 3), dtype=torch.float32, device=q.device)
        N.check(N.lib().snisr_sh_sigmoid_forward(_ptr(q), stride, _ptr(d), n, spr, deg, _ptr(spec), _stream()),
                "sh_sigmoid_forward")
        ctx.save_for_backward(spec, d)
        ctx.meta = (n, spr, deg, stride, q.shape)
        return spec

    @staticmethod
    def backward(ctx, g_spec):
        spec, d = ctx.saved_tensors
        n, spr, deg, stride, shape = ctx.meta
        g = torch.zeros(shape, dtype=torch.float32, device=spec.device)
        N.check(N.lib().snisr_sh_sigmoid_backward(_ptr(spec), _ptr(d), n, spr, deg, _ptr(g_spec.contiguous()), _ptr(g),
                                                  stride, _stream()), "sh_sigmoid_backward")
        return g, None, None, None


def sh_sigmoid(coeffs: torch.Tensor, viewdirs: torch.Tensor, deg: int) -> torch.Tensor:
    """sigmoid(eval_sh(deg, ...)) of the MLP output rows ``coeffs`` [B, S, K] (K >= 27: [basis][rgb]
    then e.g. sigma) at the rays' directions ``viewdirs`` [B, 3] (normalised inside).  Returns [B, S, 3];
    the gradient reaches coeffs[..., :27]."""
    if coeffs.shape[-1] < 27:
        raise ValueError("coeffs rows need 27 entries ([basis][rgb] of degree 2)")
    spr = coeffs.numel() // coeffs.shape[-1] // max(viewdirs.numel() // 3, 1)
    return _SHSigmoid.apply(coeffs, viewdirs, int(deg), int(spr))


class _PEMLP(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, points, depth, n_out, slope):
        p = _f32(params, "params")
        x = _f32(points, "points")
        n = x.numel() // 3
        L = N.lib()
        if p.numel() != L.snisr_pemlp_param_count(10, depth, n_out):
            raise ValueError("params do not match the MLP shape")
        out = torch.empty((*points.shape[:-1], n_out), dtype=torch.float32, device=x.device)
        ws = torch.empty(L.snisr_pemlp_workspace_size(0, 10, depth, n_out), dtype=torch.uint8, device=x.device)
        N.check(L.snisr_pemlp_forward(_ptr(p), _ptr(x), n, 10, depth, n_out, float(slope), _ptr(out), _ptr(ws),
                                      ws.numel(), _stream()), "pemlp_forward")
        ctx.save_for_backward(p, x)
        ctx.meta = (n, depth, n_out, float(slope))
        return out

    @staticmethod
    def backward(ctx, g_out):
        p, x = ctx.saved_tensors
        n, depth, n_out, slope = ctx.meta
        L = N.lib()
        ws = torch.empty(L.snisr_pemlp_workspace_size(n, 10, depth, n_out), dtype=torch.uint8, device=p.device)
        g = torch.empty_like(p)
        N.check(L.snisr_pemlp_backward(_ptr(p), _ptr(x), n, 10, depth, n_out, slope, _ptr(g_out.contiguous()), _ptr(g),
                                       _ptr(ws), ws.numel(), _stream()), "pemlp_backward")
        return g, None, None, None, None


class PEMLP(torch.nn.Module):
    """networks.MLP(hiden_depth=depth, hiden_width=64, in_channel=3, L=10, num_out=n_out, activation)
    with activation LeakyReLU(slope) (slope 0: ReLU) on the HIP kernels.  ``params`` is the flat
    parameter vector in nn.Linear order; ``load_linear_state`` takes the reference module's state dict
    (keys mlp.mlp.{2k}.weight / .bias)."""

    def __init__(self, depth: int, n_out: int, slope: float, device="cuda"):
        super().__init__()
        self.depth, self.n_out, self.slope = int(depth), int(n_out), float(slope)
        count = N.lib().snisr_pemlp_param_count(10, self.depth, self.n_out)
        if count < 0:
            raise ValueError(N.lib().snisr_last_error().decode())
        self.params = torch.nn.Parameter(torch.zeros(count, dtype=torch.float32, device=device))

    def load_linear_state(self, state: dict):
        parts = []
        for k in range(self.depth):
            parts += [state[f"mlp.mlp.{2 * k}.weight"].reshape(-1), state[f"mlp.mlp.{2 * k}.bias"].reshape(-1)]
        with torch.no_grad():
            self.params.copy_(torch.cat([t.to(self.params.device, torch.float32) for t in parts]))
        return self

    def forward(self, points: torch.Tensor) -> torch.Tensor:
        return _PEMLP.apply(self.params, points, self.depth, self.n_out, self.slope)


def _vec3(v, name):
    t = [float(a) for a in (v.tolist() if hasattr(v, "tolist") else v)]
    if len(t) != 3:
        raise ValueError(f"{name} must have three components")
    return t


def flag_cells(nodes_flag: torch.Tensor) -> torch.Tensor:
    """Row-major int32 indices of the cells of a tile's dilated node grid (int16 [D, D, D]) whose flag
    is not -1, in order (the torch.where of Tile.compute_voxels).  Reads the count back once; cache
    the result while the flags do not change."""
    if not nodes_flag.is_cuda or nodes_flag.dtype != torch.int16:
        raise ValueError("nodes_flag must be an int16 CUDA (HIP) tensor")
    f = nodes_flag.contiguous()
    n = f.numel()
    L = N.lib()
    cells = torch.empty(max(n, 1), dtype=torch.int32, device=f.device)
    count = torch.zeros(1, dtype=torch.int32, device=f.device)
    ws = torch.empty(max(L.snisr_flag_cells_workspace_size(n), 16), dtype=torch.uint8, device=f.device)
    N.check(L.snisr_flag_cells(_ptr(f), n, _ptr(cells), _ptr(count), _ptr(ws), ws.numel(), _stream()), "flag_cells")
    return cells[:int(count.item())]


class _DiffuseHead(torch.autograd.Function):
    @staticmethod
    def forward(ctx, raw, cells, n_rows):
        r = _f32(raw, "raw").reshape(-1, 4)
        out = torch.empty((n_rows, 4), dtype=torch.float32, device=r.device)
        N.check(N.lib().snisr_diffuse_head_forward(_ptr(r), r.shape[0], _ptr(cells), n_rows, _ptr(out), _stream()),
                "diffuse_head_forward")
        ctx.save_for_backward(r, cells)
        ctx.raw_shape, ctx.n_rows = raw.shape, n_rows
        return out

    @staticmethod
    def backward(ctx, g_out):
        r, cells = ctx.saved_tensors
        g = torch.empty_like(r)
        N.check(N.lib().snisr_diffuse_head_backward(_ptr(r), r.shape[0], _ptr(cells), ctx.n_rows, _ptr(g_out.contiguous()),
                                                    _ptr(g), _stream()), "diffuse_head_backward")
        return g.reshape(ctx.raw_shape), None, None


def bake_voxels(mlp: "PEMLP", cells: torch.Tensor, D: int, min_corner, voxel_size: float, center, scale: float):
    """Tile.compute_voxels: the surface MLP at the centres of ``cells`` (from flag_cells) of the
    D^3 dilated grid whose first cell starts at ``min_corner``, inputs normalised as
    (x - center) / scale, through the sigmoid / ReLU head; returns the [D, D, D, 4] grid (zero at the
    other cells), differentiable with respect to ``mlp.params``."""
    if cells.dtype != torch.int32 or not cells.is_cuda:
        raise ValueError("cells must be an int32 CUDA (HIP) tensor")
    n = cells.numel()
    mc, c = _vec3(min_corner, "min_corner"), _vec3(center, "center")
    pts = torch.empty((n, 3), dtype=torch.float32, device=cells.device)
    N.check(N.lib().snisr_bake_points(_ptr(cells), n, int(D), *mc, float(voxel_size), *c, float(scale), _ptr(pts),
                                      _stream()), "bake_points")
    return _DiffuseHead.apply(mlp(pts), cells.contiguous(), int(D) ** 3).reshape(D, D, D, 4)


def normalize_points(x: torch.Tensor, center, scale: float) -> torch.Tensor:
    """(x - center) / scale of world points [..., 3]: norm_diffuse_point with the group centre and
    size, norm_specular_point with the group centre and cfg.far (training/scripts/tile/tileGroup.py:74-75)."""
    xc = _f32(x, "x")
    out = torch.empty_like(xc)
    N.check(N.lib().snisr_normalize_points(_ptr(xc), xc.numel() // 3, *_vec3(center, "center"), float(scale),
                                           _ptr(out), _stream()), "normalize_points")
    return out


def inference_diffuse(mlp: "PEMLP", x: torch.Tensor, center, scale: float):
    """Tile.inference_diffuse: (sigmoid rgb [..., 3], relu sigma [..., 1]) of the surface MLP at world
    points x [..., 3] normalised as (x - center) / scale."""
    n = x.numel() // 3
    out = _DiffuseHead.apply(mlp(normalize_points(x, center, scale).reshape(n, 3)), None, n).reshape(*x.shape[:-1], 4)
    return out[..., :3], out[..., 3:]


class _VoxelQuery(torch.autograd.Function):
    @staticmethod
    def forward(ctx, voxels, points, mc, vs):
        v, p = _f32(voxels, "voxels"), _f32(points, "points")
        D = v.shape[0]
        if v.shape != (D, D, D, 4):
            raise ValueError("voxels must be [D, D, D, 4]")
        n = p.numel() // 3
        rgb = torch.empty((*points.shape[:-1], 3), dtype=torch.float32, device=p.device)
        sigma = torch.empty((*points.shape[:-1], 1), dtype=torch.float32, device=p.device)
        N.check(N.lib().snisr_voxel_query_forward(_ptr(v), D, _ptr(p), n, *mc, vs, _ptr(rgb), _ptr(sigma), _stream()),
                "voxel_query_forward")
        ctx.save_for_backward(p)
        ctx.meta = (D, mc, vs)
        return rgb, sigma

    @staticmethod
    def backward(ctx, g_rgb, g_sigma):
        (p,) = ctx.saved_tensors
        D, mc, vs = ctx.meta
        g = torch.empty((D, D, D, 4), dtype=torch.float32, device=p.device)
        gr = None if g_rgb is None else g_rgb.contiguous()
        gs = None if g_sigma is None else g_sigma.contiguous()
        N.check(N.lib().snisr_voxel_query_backward(D, _ptr(p), p.numel() // 3, *mc, vs, _ptr(gr), _ptr(gs), _ptr(g),
                                                   _stream()), "voxel_query_backward")
        return g, None, None, None


def query_voxel(voxels: torch.Tensor, samples: torch.Tensor, min_corner, voxel_size: float):
    """Tile.query_voxel: trilinear samples of the [D, D, D, 4] grid at world points [..., 3] (zero
    outside); returns (rgb [..., 3], sigma [..., 1]), differentiable with respect to the grid."""
    return _VoxelQuery.apply(voxels, samples, tuple(_vec3(min_corner, "min_corner")), float(voxel_size))
