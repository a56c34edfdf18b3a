"""GPU parity of the reference-format operators (the FASTRENDERING stand-in module) against the
oracle: bit-exact for rays, tile lists, inverse-depth samples, compositing, sorting and pixel
assignment."""
import numpy as np
import pytest
import torch

from snisr import fastrendering as FR
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def to_dev(a, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a)).to(DEV)
    return t if dtype is None else t.to(dtype)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_compute_rays_and_tracing_tiles(oracle_mod):
    """The first two ops of FRender.render (rendering/src/__init__.py:258-261), reference call form."""
    sc = S.make_scene(21, tile_shape=(6, 5, 3), res=8, occupancy=0.0)
    sc.tile_index_map = np.where(np.random.default_rng(0).random(90) < 0.5, np.arange(90), -1).astype(np.int32)
    osc = oracle_mod.OracleScene(sc)
    cam = S.look_at_camera((-1.0, 7.5, 2.0), (3.0, 2.5, 1.5), 64, 96, 80.0)
    rays_d = torch.full((cam.H, cam.W, 3), -1.0, device=DEV)
    FR.compute_rays(to_dev(cam.intrinsic), to_dev(cam.c2w), 256, rays_d)
    ref_d = oracle_mod.compute_rays(cam.H, cam.W, cam.intrinsic, cam.c2w)
    assert np.array_equal(bits(rays_d.cpu().numpy().reshape(-1, 3)), bits(ref_d))
    scene_size = torch.tensor(np.asarray(sc.tile_shape, np.float32) * np.float32(sc.tile_size), device=DEV)
    for K in (1, 7, 30):
        vis = torch.full((cam.H, cam.W, K), -1, device=DEV, dtype=torch.int32)
        FR.tracing_tiles(rays_d, to_dev(sc.tile_index_map), to_dev(cam.origin),
                         to_dev(np.asarray(sc.tile_shape, np.int32)), to_dev(sc.scene_min), scene_size,
                         float(sc.tile_size), 256, vis)
        ref = oracle_mod.trace_tiles(osc, cam.origin, ref_d, K)
        assert np.array_equal(vis.cpu().numpy().reshape(-1, K), ref), K
    with pytest.raises(ValueError):
        FR.tracing_tiles(rays_d, to_dev(sc.tile_index_map), to_dev(cam.origin),
                         to_dev(np.asarray(sc.tile_shape, np.int32)), to_dev(sc.scene_min), scene_size * 2,
                         float(sc.tile_size), 256, vis)


@pytest.mark.parametrize("H,W", [(36, 64), (1, 1), (7, 13)])
def test_inverse_z_sampling_strided(oracle_mod, H, W):
    """Four 16-sample chunks over a rays_d[::2, ::2] view, as FRender.render calls it
    (rendering/src/__init__.py:304-321), with the reference's inverse_bound host math."""
    rng = np.random.default_rng(H * 100 + W)
    full = rng.standard_normal((2 * H, 2 * W, 3)).astype(np.float32)
    rays_full = to_dev(full)
    view = rays_full[::2, ::2]
    dirs = np.ascontiguousarray(full[::2, ::2]).reshape(-1, 3)
    origin = np.array([0.3, -1.2, 2.5], np.float32)
    near = rng.uniform(0.05, 8.0, (H, W)).astype(np.float32)
    inv_near = (np.float32(1.0) / near).astype(np.float32)
    inv_near[0, 0] = np.float32(1e-8)   # a pixel no tile hit keeps the reference's init value
    far = np.float32(3.0)
    sample_far = (np.float32(1.0) / inv_near + far).astype(np.float32)
    inv_bound = ((np.float32(1.0) / sample_far - inv_near) / np.float32(63)).astype(np.float32)
    for step in range(0, 64, 16):
        pts = torch.full((H, W, 16, 3), -1.0, device=DEV)
        st = torch.zeros((H, W, 16), device=DEV)
        FR.inverse_z_sampling(to_dev(origin), view, to_dev(inv_near), to_dev(inv_bound), 256, step, 16, pts, st)
        rp, rs = oracle_mod.inverse_z_sampling(origin, dirs, inv_near, inv_bound, step, 16)
        assert np.array_equal(bits(pts.cpu().numpy().reshape(-1, 3)), bits(rp)), step
        assert np.array_equal(bits(st.cpu().numpy().reshape(-1)), bits(rs)), step
        assert (st.cpu().numpy() >= 0).all()


def test_integrate_points_in_place(oracle_mod):
    """Four chained 16-sample chunks with early-stopped pixels left untouched
    (rendering/src/volume_rendering.cu:5-58)."""
    rng = np.random.default_rng(5)
    H, W = 45, 80
    T = np.ones((H, W, 1), np.float32)
    T[0, :10] = 0.001   # already below early_stop: skipped
    Cc = rng.random((H, W, 3)).astype(np.float32) * np.float32(0.1)
    Tg, Cg = to_dev(T), to_dev(Cc)
    for chunk in range(4):
        res = rng.random((H, W, 16, 4)).astype(np.float32)
        res[..., 3] *= np.float32(30.0)
        res[1, :, :, 3] = 0.0           # empty space
        res[2, :, :, 3] = 1e6           # opaque
        st = rng.uniform(0.0, 0.2, (H, W, 16)).astype(np.float32)
        FR.integrate_points(to_dev(res), to_dev(st), 512, 0.01, Tg, Cg)
        oracle_mod.integrate_points(res, st, np.float32(0.01), T.reshape(-1), Cc.reshape(-1, 3))
        assert np.array_equal(bits(Tg.cpu().numpy()), bits(T)), chunk
        assert np.array_equal(bits(Cg.cpu().numpy()), bits(Cc)), chunk
    # a single pixel (one partial 64-pixel block)
    T1, C1 = np.ones((1, 1, 1), np.float32), np.zeros((1, 1, 3), np.float32)
    res = rng.random((1, 1, 16, 4)).astype(np.float32)
    st = rng.uniform(0.0, 0.2, (1, 1, 16)).astype(np.float32)
    T1g, C1g = to_dev(T1), to_dev(C1)
    FR.integrate_points(to_dev(res), to_dev(st), 512, 0.01, T1g, C1g)
    oracle_mod.integrate_points(res, st, np.float32(0.01), T1.reshape(-1), C1.reshape(-1, 3))
    assert np.array_equal(bits(T1g.cpu().numpy()), bits(T1))
    assert np.array_equal(bits(C1g.cpu().numpy()), bits(C1))


@pytest.mark.parametrize("n,nkeys", [(1, 1), (230400, 40), (230400, 1), (1000, 700), (1 << 20, 3000)])
def test_sort_by_key_matches_stable_sort(n, nkeys):
    """thrust::sort_by_key + unique_by_key (rendering/src/reorder.cu:11-23) against numpy's stable
    argsort and the run heads of the sorted keys; keys include -1 (pixels with no group)."""
    rng = np.random.default_rng(n + nkeys)
    keys = rng.integers(-1, nkeys, n).astype(np.int16)
    vals = np.arange(n, dtype=np.int32)
    starts = np.arange(n, dtype=np.int32)
    kg, vg, sg = to_dev(keys), to_dev(vals), to_dev(starts)
    m = FR.sort_by_key(kg, vg, sg)
    order = np.argsort(keys, kind="stable")
    ks = keys[order]
    heads = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]]))
    assert m == heads.size
    assert np.array_equal(vg.cpu().numpy(), vals[order])
    assert np.array_equal(kg.cpu().numpy()[:m], ks[heads])
    assert np.array_equal(sg.cpu().numpy()[:m], starts[heads])


def test_sort_by_key_empty():
    e16 = torch.empty(0, dtype=torch.int16, device=DEV)
    e32 = torch.empty(0, dtype=torch.int32, device=DEV)
    assert FR.sort_by_key(e16, e32, e32.clone()) == 0


def test_reorder_pipeline_matches_oracle(oracle_mod):
    """sort_by_key -> block_starts -> assign_pixels exactly as FRender.render builds them
    (rendering/src/__init__.py:277-291), max_pixels_per_block 64."""
    rng = np.random.default_rng(9)
    DH, DW = 360, 640
    npx = DH * DW
    net = rng.integers(-1, 120, npx).astype(np.int16)
    net[:5000] = 7                      # one large group
    kg = to_dev(net)
    q = torch.arange(npx, dtype=torch.int32, device=DEV)
    ps = torch.arange(npx, dtype=torch.int32, device=DEV)
    m = FR.sort_by_key(kg, q, ps)
    tail = torch.full((1,), npx, dtype=torch.int32, device=DEV)
    per = torch.cat([ps[1:m], tail]) - ps[:m]
    pstarts = torch.cat([ps[:m], tail])
    bstarts = torch.cat([torch.zeros(1, dtype=torch.int32, device=DEV),
                         torch.cumsum(torch.ceil(per / 64), dim=0).int()])
    nb = int(bstarts[-1])
    new_net = torch.empty(nb, dtype=torch.int16, device=DEV)
    new_ps = torch.empty(nb, dtype=torch.int32, device=DEV)
    FR.assign_pixels(kg[:m], pstarts, bstarts, 256, new_net, new_ps)
    ref_net, ref_ps = oracle_mod.assign_pixels(kg[:m].cpu().numpy(), pstarts.cpu().numpy(), bstarts.cpu().numpy())
    assert np.array_equal(new_net.cpu().numpy(), ref_net)
    assert np.array_equal(new_ps.cpu().numpy(), ref_ps)
    # every chunk holds at most 64 pixels and the chunks tile [0, npx)
    edges = np.concatenate([new_ps.cpu().numpy(), [npx]])
    sizes = np.diff(edges)
    assert sizes.min() >= 1 and sizes.max() <= 64 and edges[0] == 0


@pytest.mark.parametrize("H,W,ksize", [(720, 1280, 3), (45, 77, 5), (17, 33, 9), (8, 8, 1)])
def test_compute_group_boundary_bit_exact(oracle_mod, H, W, ksize):
    rng = np.random.default_rng(H + W)
    blocks = rng.integers(-1, 7, ((H + 5) // 6, (W + 6) // 7))
    net = np.repeat(np.repeat(blocks, 6, 0), 7, 1)[:H, :W].astype(np.int16)
    pre = np.zeros((H, W, 1), bool)
    pre[H // 2, W // 3] = True
    b = to_dev(pre)
    FR.compute_group_boundary(to_dev(net), b, ksize, 512)
    ref = oracle_mod.group_boundary(net, ksize, pre.reshape(H, W).astype(np.uint8).copy()).astype(bool)
    assert np.array_equal(b.cpu().numpy().reshape(H, W), ref)
    assert ref.any() and not ref.all()
