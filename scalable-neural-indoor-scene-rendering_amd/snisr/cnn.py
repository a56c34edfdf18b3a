"""The reference's screen-space refinement network (rendering/scripts/networks.py:5-81): a 5-channel
U-Net (rgb, inverse depth of the first surface, group-boundary flag) with 3x3 convolutions, average-
pool downsampling and bilinear (align_corners) upsampling, applied to the composited frame when
enabled (rendering/src/__init__.py:336-339).  The module and parameter names are the reference's,
so its trained state dict (loaded with ``torch.load(..., weights_only=True)``) loads as is.

``HipUnet`` runs the network on the library's FP16 MFMA convolutions (``snisr_unet_*``), which
replace the TensorRT engine the reference compiles from the same module in FP16
(__init__.py:85-88); ``refine`` with the module itself runs it through torch (MIOpen) and serves as
the FP32 reference."""
from __future__ import annotations

import torch
import torch.nn as nn


class Unet(nn.Module):
    def __init__(self, in_channel: int, mode: str = "bilinear", align_corners: bool = True):
        super().__init__()
        conv = lambda i, o: nn.Conv2d(in_channels=i, out_channels=o, kernel_size=3, stride=1, padding=1)
        self.conv1, self.conv2 = conv(in_channel, 64), conv(64, 32)
        self.conv3, self.conv4 = conv(32, 64), conv(64, 64)
        self.conv5, self.conv6 = conv(64, 128), conv(128, 128)
        self.conv7, self.conv8 = conv(192, 64), conv(64, 64)
        self.conv9, self.conv10 = conv(96, 32), conv(32, 3)
        self.relu = nn.ReLU()
        self.down = nn.AvgPool2d(kernel_size=2, stride=2)
        self.up = nn.Upsample(scale_factor=2, mode=mode, align_corners=align_corners)

    def forward(self, x):
        r = self.relu
        o2 = r(self.conv2(r(self.conv1(x))))
        o5 = r(self.conv4(r(self.conv3(self.down(o2)))))
        e7 = r(self.conv6(r(self.conv5(self.down(o5)))))
        e4 = r(self.conv8(r(self.conv7(torch.cat([self.up(e7), o5], dim=1)))))
        return r(self.conv10(r(self.conv9(torch.cat([self.up(e4), o2], dim=1)))))


class CNN(nn.Module):
    def __init__(self):
        super().__init__()
        self.cur_inchannel = 5
        self.RCONS = Unet(self.cur_inchannel)

    def forward(self, x):
        return self.RCONS(x)


def refine(cnn: nn.Module, frame: torch.Tensor, inverse_near: torch.Tensor, boundary: torch.Tensor) -> torch.Tensor:
    """[H, W, 3] refined frame, clamped to [0, 1], from the composited frame [H, W, 3], the
    full-resolution inverse depth [H, W] and the boundary flags [H, W(, 1)] (__init__.py:336-339).
    H and W must be multiples of 4 (two 2x poolings)."""
    H, W = frame.shape[0], frame.shape[1]
    if H % 4 or W % 4:
        raise ValueError("the refinement network needs H and W divisible by 4")
    dt = next(cnn.parameters()).dtype
    x = torch.cat([frame.permute(2, 0, 1), inverse_near.reshape(1, H, W), boundary.reshape(1, H, W).to(frame.dtype)],
                  0)[None].to(dt)
    with torch.no_grad():
        out = cnn(x)
    return out[0].permute(1, 2, 0).float().clamp(0, 1)


# (in, out) channels of conv1 .. conv10 (networks.py:8-15), the order snisr_unet_pack expects
UNET_CONVS = [(5, 64), (64, 32), (32, 64), (64, 64), (64, 128), (128, 128), (192, 64), (64, 64), (96, 32), (32, 3)]


class HipUnet:
    """The refinement network on FP16 MFMA convolutions (snisr_unet_pack / snisr_unet_refine).

    Built from a ``CNN`` / ``Unet`` module or its state dict (keys ``RCONS.convN.*`` or ``convN.*``);
    the weights are packed once on the device.  ``__call__(frame, inverse_near, boundary)`` returns
    the refined frame f32[H, W, 3] in [0, 1], like ``refine``."""

    def __init__(self, net, device=None):
        import ctypes as C
        from . import _native as N
        from .device import _ptr, _stream
        sd = net.state_dict() if hasattr(net, "state_dict") else net
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        ws, bs = [], []
        for i, (cin, cout) in enumerate(UNET_CONVS, start=1):
            wk = [k for k in sd if k.endswith(f"conv{i}.weight")]
            bk = [k for k in sd if k.endswith(f"conv{i}.bias")]
            if len(wk) != 1 or len(bk) != 1:
                raise ValueError(f"state dict has no unique conv{i}.weight / conv{i}.bias")
            w = sd[wk[0]].detach().to(dev, torch.float32).contiguous()
            b = sd[bk[0]].detach().to(dev, torch.float32).contiguous()
            if tuple(w.shape) != (cout, cin, 3, 3) or tuple(b.shape) != (cout,):
                raise ValueError(f"conv{i}: expected weight ({cout}, {cin}, 3, 3) and bias ({cout},), "
                                 f"got {tuple(w.shape)} and {tuple(b.shape)}")
            ws.append(w)
            bs.append(b)
        self.device = dev
        self.packed = torch.empty(int(N.lib().snisr_unet_packed_size()), dtype=torch.uint8, device=dev)
        wa = (C.c_void_p * 10)(*[t.data_ptr() for t in ws])
        ba = (C.c_void_p * 10)(*[t.data_ptr() for t in bs])
        N.check(N.lib().snisr_unet_pack(wa, ba, _ptr(self.packed), _stream()), "unet_pack")
        torch.cuda.current_stream().synchronize()  # the fp32 copies above may be freed after this
        self._ws = None

    def __call__(self, frame: torch.Tensor, inverse_near: torch.Tensor, boundary: torch.Tensor) -> torch.Tensor:
        from . import _native as N
        from .device import _ptr, _require_cuda, _stream
        H, W = int(frame.shape[0]), int(frame.shape[1])
        _require_cuda(frame, "frame", torch.float32)
        _require_cuda(inverse_near, "inverse_near", torch.float32)
        if frame.shape[2:] != (3,) or inverse_near.numel() != H * W or boundary.numel() != H * W:
            raise ValueError("frame must be [H, W, 3], inverse_near and boundary H*W elements")
        if boundary.dtype not in (torch.bool, torch.uint8):
            raise ValueError("boundary must be bool or uint8")
        _require_cuda(boundary, "boundary")
        need = int(N.lib().snisr_unet_workspace_size(H, W))
        if need < 0:
            raise ValueError(N.lib().snisr_last_error().decode())
        if self._ws is None or self._ws.numel() < need:
            self._ws = torch.empty(need, dtype=torch.uint8, device=frame.device)
        out = torch.empty((H, W, 3), dtype=torch.float32, device=frame.device)
        N.check(N.lib().snisr_unet_refine(_ptr(self.packed), _ptr(frame), _ptr(inverse_near), _ptr(boundary), H, W,
                                          _ptr(self._ws), self._ws.numel(), _ptr(out), _stream()), "unet_refine")
        return out
