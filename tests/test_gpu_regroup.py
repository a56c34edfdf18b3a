"""The library's own stable radix sort and device-side regrouping (A8; rendering/src/reorder.cu:11-94,
rendering/src/__init__.py:272-291) against numpy's stable argsort and the reference's block rule:
the whole signed int16 key range, every distinct key at once, sizes that take more than 4 tiles per
chunk, graph capture of the regrouping replayed on new inputs, and the refusals (sort_by_key on a
capturing stream, too few output blocks)."""
import numpy as np
import pytest
import torch

from snisr import _native as N
from snisr import fastrendering as FR
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _keys(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "full_range":
        return rng.integers(-32768, 32767, n, endpoint=True).astype(np.int16)
    if kind == "all_distinct":   # every int16 value at least once, shuffled
        k = np.concatenate([np.arange(-32768, 32768), rng.integers(-32768, 32767, n - 65536, endpoint=True)])
        return rng.permutation(k).astype(np.int16)
    if kind == "constant":
        return np.full(n, -1, np.int16)
    if kind == "sorted_desc":
        return np.sort(rng.integers(-5, 300, n))[::-1].astype(np.int16)
    return rng.integers(-1, 12, n).astype(np.int16)


CASES = [("groups", 1), ("groups", 255), ("groups", 256), ("groups", 257), ("full_range", 100_000),
         ("all_distinct", 70_000), ("constant", 230_400), ("sorted_desc", 50_000), ("groups", 3_000_001)]


@pytest.mark.parametrize("kind,n", CASES)
def test_sort_by_key_is_stable_sort(kind, n):
    print(f"sort_by_key {kind} n={n}", flush=True)
    keys = _keys(kind, n, n)
    vals = np.random.default_rng(1).integers(0, 1 << 30, n).astype(np.int32)
    starts = np.arange(n, dtype=np.int32)
    kg, vg, sg = (torch.from_numpy(a).to(DEV) for a in (keys, vals, starts))
    m = FR.sort_by_key(kg, vg, sg)
    order = np.argsort(keys, kind="stable")
    ks = keys[order]
    heads = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]]))
    assert m == heads.size
    assert np.array_equal(vg.cpu().numpy(), vals[order])
    kk = kg.cpu().numpy()
    assert np.array_equal(kk[:m], ks[heads])
    assert np.array_equal(kk[m:], ks[m:])
    assert np.array_equal(sg.cpu().numpy()[:m], heads.astype(np.int32))


def _regroup_buffers(n, P=64):
    mb = FR.regroup_max_blocks(n, P)
    return (mb, torch.empty(n, dtype=torch.int32, device=DEV), torch.full((mb,), 7, dtype=torch.int16, device=DEV),
            torch.full((mb + 1,), 7, dtype=torch.int32, device=DEV), torch.zeros(1, dtype=torch.int32, device=DEV),
            torch.empty(int(N.lib().snisr_regroup_workspace_size(n)), dtype=torch.uint8, device=DEV))


def _check_regroup(net, query, bnet, bps, nbl, P=64):
    n = net.size
    q_ref, bnet_ref, bps_ref = S.reflection_blocks(net, P)
    k = int(nbl)
    assert k == bnet_ref.size
    assert np.array_equal(query.cpu().numpy(), q_ref)
    assert np.array_equal(bnet[:k].cpu().numpy(), bnet_ref)
    assert np.array_equal(bps[:k + 1].cpu().numpy(), bps_ref)
    assert bool((bnet[k:] == -1).all()) and bool((bps[k:] == n).all())


@pytest.mark.parametrize("kind,n,P", [("full_range", 100_000, 64), ("all_distinct", 70_000, 64),
                                      ("constant", 230_400, 64), ("groups", 3_000_001, 64), ("groups", 999, 7),
                                      ("groups", 1, 64)])
def test_regroup_pixels_key_ranges(kind, n, P):
    print(f"regroup {kind} n={n} P={P}", flush=True)
    net = _keys(kind, n, n + 1)
    mb, query, bnet, bps, nbl, ws = _regroup_buffers(n, P)
    FR.regroup_pixels(torch.from_numpy(net).to(DEV), P, query, bnet, bps, nbl, ws)
    _check_regroup(net, query, bnet, bps, nbl, P)


def test_regroup_pixels_empty():
    mb, query, bnet, bps, nbl, ws = _regroup_buffers(0)
    nbl.fill_(5)
    FR.regroup_pixels(torch.empty(0, dtype=torch.int16, device=DEV), 64, query, bnet, bps, nbl, ws)
    assert int(nbl) == 0 and int(bps[0]) == 0


def test_regroup_pixels_graph_replays_on_new_inputs():
    """The regrouping captured once (230,400 pixels, the 720p reflection grid) and replayed on three
    new group maps copied into its input buffer, with eager work between the replays."""
    print("regroup graph 230400", flush=True)
    n = 230_400
    net_dev = torch.zeros(n, dtype=torch.int16, device=DEV)
    mb, query, bnet, bps, nbl, ws = _regroup_buffers(n)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        FR.regroup_pixels(net_dev, 64, query, bnet, bps, nbl, ws)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        FR.regroup_pixels(net_dev, 64, query, bnet, bps, nbl, ws)
    for seed in (3, 4, 5):
        net = np.random.default_rng(seed).integers(-1, 9 + 40 * seed, n).astype(np.int16)
        net_dev.copy_(torch.from_numpy(net))
        g.replay()
        torch.cuda.synchronize()
        _check_regroup(net, query, bnet, bps, nbl)
        _ = torch.randn(1000, device=DEV).sum().item()


def test_sort_by_key_refuses_capturing_stream():
    """sort_by_key returns its group count to the host (it synchronises), so under stream capture it
    is refused with SNISR_ERR_UNSUPPORTED before any launch; the inputs stay untouched."""
    n = 1000
    keys = torch.from_numpy(np.random.default_rng(0).integers(-1, 5, n).astype(np.int16)).to(DEV)
    before = keys.clone()
    vals = torch.arange(n, dtype=torch.int32, device=DEV)
    starts = torch.arange(n, dtype=torch.int32, device=DEV)
    FR.sort_by_key(keys.clone(), vals.clone(), starts.clone())   # sizes the cached workspace outside capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    err = None
    with torch.cuda.graph(g):
        try:
            FR.sort_by_key(keys, vals, starts)
        except N.SnisrError as e:
            err = e
        twice = starts * 2   # keeps the captured graph non-empty
    assert err is not None and err.status == -6, err
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(keys, before)
    assert torch.equal(twice, 2 * torch.arange(n, dtype=torch.int32, device=DEV))
