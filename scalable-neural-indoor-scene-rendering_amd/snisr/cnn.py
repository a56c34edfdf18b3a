"""The reference's screen-space refinement network (rendering/scripts/networks.py:5-81): a 5-channel
U-Net (rgb, inverse depth of the first surface, group-boundary flag) with 3x3 convolutions, average-
pool downsampling and bilinear (align_corners) upsampling, applied to the composited frame when
enabled (rendering/src/__init__.py:336-339).  The module and parameter names are the reference's,
so its trained state dict (loaded with ``torch.load(..., weights_only=True)``) loads as is.  The
convolutions run through torch on the device (MIOpen on ROCm) in FP16 for the frame and FP32 for
checking; the reference compiles the same network with TensorRT in FP16 (__init__.py:85-88)."""
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
