"""The refinement U-Net on the library's FP16 MFMA convolutions (snisr_unet_*) against the
reference's module (rendering/scripts/networks.py:5-73) evaluated in FP32 by torch on the same
input: the FP16 activations bound the difference (max 2e-2, mean 2e-3 on the [0, 1] output)."""
import pytest
import torch

from snisr import cnn as CN

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _inputs(H, W, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    frame = torch.rand(H, W, 3, generator=g).to(DEV)
    inv = (torch.rand(H, W, generator=g) * 0.8 + 0.05).to(DEV)
    bnd = (torch.rand(H, W, generator=g) < 0.1).to(DEV)
    return frame, inv, bnd


def _fp32_reference(net, frame, inv, bnd):
    H, W = frame.shape[:2]
    x = torch.cat([frame.permute(2, 0, 1), inv.reshape(1, H, W), bnd.reshape(1, H, W).float()], 0)[None]
    with torch.no_grad():
        return net.to(DEV).float()(x)[0].permute(1, 2, 0).clamp(0, 1)


@pytest.mark.parametrize("H,W,seed", [(64, 96, 0), (36, 52, 1), (8, 4, 2), (720, 1280, 3)])
def test_hip_unet_matches_fp32_module(H, W, seed):
    print(f"hip unet {H}x{W} seed {seed}", flush=True)
    torch.manual_seed(seed)
    net = CN.CNN().eval()
    hip = CN.HipUnet(net, device=DEV)
    frame, inv, bnd = _inputs(H, W, seed)
    out = hip(frame, inv, bnd)
    torch.cuda.synchronize()
    ref = _fp32_reference(net, frame, inv, bnd)
    err = (out - ref).abs()
    assert out.shape == (H, W, 3) and bool(torch.isfinite(out).all())
    assert float(err.max()) < 2e-2 and float(err.mean()) < 2e-3, (float(err.max()), float(err.mean()))
    assert float(ref.std()) > 1e-3  # a non-trivial output


def test_hip_unet_repeatable_and_uint8_boundary():
    torch.manual_seed(5)
    net = CN.CNN().eval()
    hip = CN.HipUnet(net.state_dict(), device=DEV)  # from the state dict (RCONS.convN.* keys)
    frame, inv, bnd = _inputs(48, 64, 7)
    a = hip(frame, inv, bnd)
    b = hip(frame, inv, bnd.to(torch.uint8))
    assert torch.equal(a, b)


def test_hip_unet_rejects_bad_shapes():
    net = CN.CNN().eval()
    hip = CN.HipUnet(net, device=DEV)
    frame, inv, bnd = _inputs(6, 8, 0)  # H not a multiple of 4
    with pytest.raises((ValueError, RuntimeError)):
        hip(frame, inv, bnd)
    with pytest.raises(ValueError):
        CN.HipUnet({"conv1.weight": torch.zeros(64, 5, 3, 3)}, device=DEV)
