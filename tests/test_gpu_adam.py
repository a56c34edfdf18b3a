"""The one-pass Adam step (snisr_adam_step, snisr.train.Adam) against torch.optim.Adam, the
reference's optimizer (training/scripts/tile/tileGroup.py:342), on the same parameters and
gradients over several steps, with and without weight decay, on aligned and ragged sizes."""
import pytest
import torch

from snisr import train as TR

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


@pytest.mark.parametrize("n,wd", [(1 << 20, 0.0), (1 << 20, 1e-5), (1001, 1e-5), (3, 0.0)])
def test_adam_matches_torch(n, wd):
    print(f"adam n={n} wd={wd}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(n)
    p0 = torch.randn(n, generator=g).to(DEV)
    a = p0.clone().requires_grad_(True)
    b = p0.clone().requires_grad_(True)
    oa = TR.Adam([a], lr=1e-3, weight_decay=wd)
    ob = torch.optim.Adam([b], lr=1e-3, weight_decay=wd)
    for k in range(5):
        grad = (torch.randn(n, generator=g) * (0.1 + k)).to(DEV)
        a.grad = grad.clone()
        b.grad = grad.clone()
        oa.step()
        ob.step()
    torch.cuda.synchronize()
    # torch's kernels are built with FMA contraction and this one without, so single roundings differ;
    # the bar is FP32 rounding relative to each tensor's scale (a wrong beta, bias correction or
    # weight-decay term is off by orders of magnitude more)
    for x, y in ((a.detach(), b.detach()), (oa.state[a]["exp_avg"], ob.state[b]["exp_avg"]),
                 (oa.state[a]["exp_avg_sq"], ob.state[b]["exp_avg_sq"])):
        err, scale = float((x - y).abs().max()), float(y.abs().max())
        assert err <= 2e-6 * scale, (err, scale)
    assert float((a - b).abs().max().detach()) < 1e-3 * float((b - p0).abs().max().detach())  # tiny next to the update itself
    assert not torch.equal(a, p0)


def test_adam_rejects_bad_arguments():
    a = torch.zeros(8, device=DEV, requires_grad=True)
    with pytest.raises(ValueError):
        TR.Adam([a], betas=(1.0, 0.999))
    c = torch.zeros(8, device=DEV, dtype=torch.float64, requires_grad=True)
    c.grad = torch.zeros_like(c)
    with pytest.raises(ValueError):
        TR.Adam([c]).step()
