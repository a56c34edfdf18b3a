"""Differentiable render: a torch.autograd.Function over the hand-written HIP forward and adjoint
(snisr_train_forward / snisr_train_backward).  Parameters are the scene's density sigma
[NB,512], features [NB,512,F] (FP32) and per-tile MLP parameters [T,P]; the forward recomputes
nothing on the host and the backward recomputes the per-sample MLP forward on the device."""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N
from .device import DeviceScene, _attach_events, _ptr, _rays_desc, _render_outputs, _require_cuda, _stream, mfma_stats


class TrainWorkspace:
    """Workspace of one forward/backward pair (grown on demand).

    The first forward on a ray count checks the sample capacity (one host synchronisation) and grows
    the workspace to the batch's sample count with 25% headroom; later forwards with the same ray
    count run without synchronising, so a training step can be captured in a HIP graph.  A forward
    that outgrows the workspace renders NaN and is reported by ``overflowed()`` (synchronises).
    ``generation`` counts the forwards: a backward whose forward is no longer the last one on this
    workspace (its records were overwritten) raises."""

    def __init__(self, capacity_per_ray: int = 48):
        self.buf = None
        self.capacity_per_ray = capacity_per_ray
        self.sized_for = None
        self.generation = 0

    def get(self, ds: DeviceScene, n: int):
        need = N.lib().snisr_train_workspace_size(C.byref(ds.desc), n, self.capacity_per_ray * n)
        if need < 0:
            raise RuntimeError(N.lib().snisr_last_error().decode())
        if self.buf is None or self.buf.numel() < need:
            self.buf = torch.empty(need, dtype=torch.uint8, device=ds.sigma.device)
        return self.buf, need

    def overflowed(self) -> bool:
        """True if the last forward had more samples than the workspace holds (synchronises)."""
        return self.buf is not None and mfma_stats(self.buf)[1] != 0

    def samples_last(self) -> int:
        """Samples of the last forward (synchronises)."""
        return mfma_stats(self.buf)[0] if self.buf is not None else 0


def train_forward(ds: DeviceScene, origins, dirs, t_hit=None, workspace: TrainWorkspace | None = None,
                  events=None, check: bool | None = None, **march_kw):
    """snisr_train_forward.  ``check``: verify the sample capacity after the call and redo it with a
    grown workspace on overflow (synchronises); None = only the first time this workspace sees this
    ray count."""
    r = _rays_desc(origins, dirs, t_hit)
    p = ds.march_params(**march_kw)
    out, o = _render_outputs(r.n, dirs.device, True, 0)
    keep = _attach_events(o, events)  # noqa: F841
    ws = workspace or TrainWorkspace()
    if check is None:
        check = ws.sized_for != r.n
    buf, need = ws.get(ds, r.n)
    ws.generation += 1
    N.check(N.lib().snisr_train_forward(C.byref(ds.desc), C.byref(p), C.byref(r), C.byref(o), _ptr(buf), need,
                                        _stream()), "train_forward")
    if check:
        total, over = mfma_stats(buf)
        if over:  # grow to the exact sample count and redo
            ws.capacity_per_ray = int(total * 1.25 / max(r.n, 1)) + 1
            buf, need = ws.get(ds, r.n)
            N.check(N.lib().snisr_train_forward(C.byref(ds.desc), C.byref(p), C.byref(r), C.byref(o), _ptr(buf),
                                                need, _stream()), "train_forward")
            total, over = mfma_stats(buf)
            if over:
                raise RuntimeError("train_forward: sample capacity still exceeded after growing the workspace")
        ws.sized_for = r.n
        out["total_samples"] = total
    return out, ws


def train_backward(ds: DeviceScene, origins, dirs, grad_rgb, workspace: TrainWorkspace, t_hit=None, events=None,
                   **march_kw):
    r = _rays_desc(origins, dirs, t_hit)
    p = ds.march_params(**march_kw)
    g = grad_rgb.contiguous()
    _require_cuda(g, "grad_rgb", torch.float32)
    dsig = torch.empty_like(ds.sigma, dtype=torch.float32)
    dfeat = torch.empty(ds.feats.shape, dtype=torch.float32, device=ds.feats.device)
    dmlp = torch.empty_like(ds.mlp_params)
    arr, nev = None, 0
    if events:
        arr = (C.c_void_p * len(events))(*[C.c_void_p(e.cuda_event) for e in events])
        nev = len(events)
    buf, need = workspace.get(ds, r.n)
    N.check(N.lib().snisr_train_backward(C.byref(ds.desc), C.byref(p), C.byref(r), _ptr(g), _ptr(dsig), _ptr(dfeat),
                                         _ptr(dmlp), _ptr(buf), need, arr, nev, _stream()), "train_backward")
    return dsig, dfeat, dmlp


class VoxelRender(torch.autograd.Function):
    """rgb = VoxelRender.apply(sigma, feats, mlp_params, origins, dirs, t_hit, ds, workspace, march_kw)"""

    @staticmethod
    def forward(ctx, sigma, feats, mlp_params, origins, dirs, t_hit, ds: DeviceScene, workspace, march_kw):
        ds.set_params(mlp_params=mlp_params.detach().contiguous(), feats=feats.detach().contiguous(),
                      sigma=sigma.detach().contiguous())
        out, ws = train_forward(ds, origins, dirs, t_hit, workspace, **(march_kw or {}))
        ctx.ds, ctx.ws, ctx.march_kw = ds, ws, march_kw or {}
        ctx.generation = ws.generation
        ctx.save_for_backward(origins, dirs, t_hit if t_hit is not None else torch.empty(0, device=dirs.device))
        ctx.has_t_hit = t_hit is not None
        return out["rgb"]

    @staticmethod
    def backward(ctx, grad_rgb):
        origins, dirs, t_hit = ctx.saved_tensors
        if ctx.ws.generation != ctx.generation:
            raise RuntimeError("VoxelRender.backward: the workspace was used by another forward after this one; "
                               "give each pending forward its own TrainWorkspace")
        dsig, dfeat, dmlp = train_backward(ctx.ds, origins, dirs, grad_rgb, ctx.ws,
                                           t_hit if ctx.has_t_hit else None, **ctx.march_kw)
        return dsig, dfeat, dmlp, None, None, None, None, None, None


def render(sigma, feats, mlp_params, origins, dirs, ds: DeviceScene, t_hit=None, workspace=None, **march_kw):
    """Differentiable rgb [N,3] of the rays through the scene held by ``ds`` with the given parameters."""
    return VoxelRender.apply(sigma, feats, mlp_params, origins, dirs, t_hit, ds, workspace, march_kw)


class Adam(torch.optim.Optimizer):
    """torch.optim.Adam (no amsgrad) on the library's one-pass kernel (snisr_adam_step): the
    reference's optimizer, torch.optim.Adam(params, weight_decay=1e-5) (training/scripts/tile/
    tileGroup.py:342).  Parameters must be contiguous FP32 CUDA tensors; the state holds ``step``,
    ``exp_avg`` and ``exp_avg_sq`` like torch's."""

    def __init__(self, params, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, weight_decay=0.0):
        if not 0.0 <= betas[0] < 1.0 or not 0.0 <= betas[1] < 1.0:
            raise ValueError(f"betas must lie in [0, 1), got {betas}")
        super().__init__(params, dict(lr=lr, betas=betas, eps=eps, weight_decay=weight_decay))

    @torch.no_grad()
    def step(self, closure=None):
        loss = None
        if closure is not None:
            with torch.enable_grad():
                loss = closure()
        for group in self.param_groups:
            b1, b2 = group["betas"]
            for p in group["params"]:
                if p.grad is None:
                    continue
                _require_cuda(p, "param", torch.float32)
                _require_cuda(p.grad, "grad", torch.float32)
                st = self.state[p]
                if not st:
                    st["step"] = 0
                    st["exp_avg"] = torch.zeros_like(p)
                    st["exp_avg_sq"] = torch.zeros_like(p)
                st["step"] += 1
                N.check(N.lib().snisr_adam_step(_ptr(p), _ptr(p.grad), _ptr(st["exp_avg"]), _ptr(st["exp_avg_sq"]),
                                                p.numel(), float(group["lr"]), float(b1), float(b2),
                                                float(group["eps"]), float(group["weight_decay"]), st["step"],
                                                _stream()), "adam_step")
        return loss
