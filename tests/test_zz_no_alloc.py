"""The allocation-free regrouping (snisr_regroup_pixels_static, workspace in a fixed device array of
the library) against snisr_regroup_pixels with a caller workspace, through the C ABI: identical
query, block groups, block starts and count on two calls in a row, the second with a smaller input
(what the first call left in the shared array must not leak into the second), and the refusal of
inputs beyond the static maximum before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from snisr import _native as N

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _run(net, P, static):
    L = N.lib()
    n = net.numel()
    mb = int(L.snisr_regroup_max_blocks(n, P))
    query = torch.full((max(n, 1),), -3, dtype=torch.int32, device=DEV)
    bnet = torch.full((max(mb, 1),), 7, dtype=torch.int16, device=DEV)
    bps = torch.full((mb + 1,), 7, dtype=torch.int32, device=DEV)
    nbl = torch.full((1,), -9, dtype=torch.int32, device=DEV)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    if static:
        rc = L.snisr_regroup_pixels_static(_ptr(net), n, P, mb, _ptr(query), _ptr(bnet), _ptr(bps), _ptr(nbl), stream)
    else:
        ws = torch.empty(int(L.snisr_regroup_workspace_size(n)), dtype=torch.uint8, device=DEV)
        rc = L.snisr_regroup_pixels(_ptr(net), n, P, mb, _ptr(query), _ptr(bnet), _ptr(bps), _ptr(nbl), _ptr(ws),
                                    ws.numel(), stream)
    assert rc == 0, L.snisr_last_error()
    torch.cuda.synchronize()
    k = int(nbl)
    return query[:n].cpu().numpy(), bnet[:k].cpu().numpy(), bps[:k + 1].cpu().numpy(), k


def test_static_regroup_equals_workspace_regroup_on_consecutive_calls():
    L = N.lib()
    assert L.snisr_regroup_pixels_static_max() == 1 << 20
    big = torch.from_numpy(np.random.default_rng(0).integers(-1, 300, 921_600).astype(np.int16)).to(DEV)
    small = torch.from_numpy(np.random.default_rng(1).integers(-1, 7, 4097).astype(np.int16)).to(DEV)
    for net in (big, small):
        a = _run(net, 64, static=True)
        b = _run(net, 64, static=False)
        assert a[3] == b[3]
        for x, y in zip(a[:3], b[:3]):
            assert np.array_equal(x, y)


def test_static_regroup_refuses_inputs_beyond_its_maximum():
    L = N.lib()
    n = (1 << 20) + 1
    fake = C.c_void_p(256)
    rc = L.snisr_regroup_pixels_static(fake, n, 64, int(L.snisr_regroup_max_blocks(n, 64)), fake, fake, fake, fake,
                                       None)
    assert rc == -5 and b"maximum" in L.snisr_last_error()
