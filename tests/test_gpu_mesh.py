"""GPU parity of the BVH nearest hit with the oracle's brute force over all triangles:
triangle ids and t bit-exact."""
import numpy as np
import pytest
import torch

from snisr import device as D
from snisr import scene as S

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def check(oracle_mod, mesh, origins, dirs):
    dm = D.DeviceMesh(mesh.vertices, mesh.faces)
    t, ids = dm.first_hit(to_dev(origins), to_dev(dirs))
    dn = oracle_mod.normalize(dirs)
    t_ref, id_ref = oracle_mod.mesh_hit(mesh.vertices, mesh.faces, origins, dn)
    ids = ids.cpu().numpy()
    t = t.cpu().numpy()
    assert np.array_equal(ids, id_ref), np.flatnonzero(ids != id_ref)[:10]
    assert np.array_equal(t.view(np.uint32), t_ref.view(np.uint32))
    return ids


def test_mesh_hit_c2_mesh_camera_rays(oracle_mod):
    sc, cam, mesh = S.config_c2()
    dirs = S.camera_rays_np(cam)
    idx = np.random.default_rng(0).choice(dirs.shape[0], 1500, replace=False)
    ids = check(oracle_mod, mesh, cam.origin, dirs[idx])
    assert np.mean(ids >= 0) > 0.9


def test_mesh_hit_random_rays_and_edge_cases(oracle_mod):
    mesh = S.room_mesh(7, (0, 0, 0), (2, 2, 1), 24000)
    rng = np.random.default_rng(1)
    n = 6000
    origins = rng.uniform([-0.5, -0.5, -0.5], [2.5, 2.5, 1.5], (n, 3)).astype(np.float32)
    dirs = rng.standard_normal((n, 3)).astype(np.float32)
    dirs[:300, 1:] = 0.0          # axis-parallel
    dirs[300:600, 2] = 0.0
    dirs[600:700, 0] = -0.0
    origins[700:800] = np.array([5.0, 5.0, 5.0], np.float32)  # outside everything, pointing away
    dirs[700:800] = np.abs(dirs[700:800])
    check(oracle_mod, mesh, origins, dirs)


def test_mesh_hit_tiny_and_empty_meshes(oracle_mod):
    one = S.box_mesh(np.array([[0.2, 0.2, 0.2]]), np.array([[0.4, 0.5, 0.6]]))
    rng = np.random.default_rng(2)
    origins = rng.uniform(-1, 1.5, (500, 3)).astype(np.float32)
    dirs = (np.array([0.3, 0.35, 0.4]) - origins + rng.normal(0, 0.1, (500, 3))).astype(np.float32)
    check(oracle_mod, one, origins, dirs)
    tri = S.Mesh(vertices=one.vertices[one.faces[0]].copy(), faces=np.array([[0, 1, 2]], np.int32))
    check(oracle_mod, tri, origins, dirs)
    empty = S.Mesh(vertices=np.zeros((3, 3), np.float32), faces=np.zeros((0, 3), np.int32))
    dm = D.DeviceMesh(empty.vertices, empty.faces)
    t, ids = dm.first_hit(to_dev(origins), to_dev(dirs))
    assert torch.all(ids == -1) and torch.all(torch.isinf(t))


def test_packet_traversal_matches_per_lane_and_brute_force(oracle_mod):
    """Camera 8x8 packets take the wavefront-shared traversal: the full C2 frame equals the per-lane
    traversal bit-exactly, a subset equals the brute force, and incoherent rays laid out as
    packets (the worst case for a shared traversal) equal the brute force."""
    sc, cam, mesh = S.config_c2()
    dm = D.DeviceMesh(mesh.vertices, mesh.faces)
    dirs = S.camera_rays_np(cam)
    o, d = to_dev(cam.origin), to_dev(dirs)
    t_pk, id_pk = dm.first_hit(o, d, image_width=cam.W)
    t_ln, id_ln = dm.first_hit(o, d)
    assert torch.equal(id_pk, id_ln) and torch.equal(t_pk.view(torch.int32), t_ln.view(torch.int32))
    idx = np.random.default_rng(3).choice(dirs.shape[0], 1500, replace=False)
    t_ref, id_ref = oracle_mod.mesh_hit(mesh.vertices, mesh.faces, cam.origin, oracle_mod.normalize(dirs[idx]))
    assert np.array_equal(id_pk.cpu().numpy()[idx], id_ref)
    assert np.array_equal(t_pk.cpu().numpy()[idx].view(np.uint32), t_ref.view(np.uint32))
    small = S.room_mesh(8, (0, 0, 0), (2, 2, 1), 12000)
    dsm = D.DeviceMesh(small.vertices, small.faces)
    rng = np.random.default_rng(4)
    n = 96 * 64
    origins = rng.uniform([-0.5, -0.5, -0.5], [2.5, 2.5, 1.5], (n, 3)).astype(np.float32)
    rd = rng.standard_normal((n, 3)).astype(np.float32)
    t, ids = dsm.first_hit(to_dev(origins), to_dev(rd), image_width=96)
    t_ref, id_ref = oracle_mod.mesh_hit(small.vertices, small.faces, origins, oracle_mod.normalize(rd))
    assert np.array_equal(ids.cpu().numpy(), id_ref)
    assert np.array_equal(t.cpu().numpy().view(np.uint32), t_ref.view(np.uint32))


@pytest.mark.parametrize("nb", [8, 32])
def test_mesh_hit_tiles_first_and_double_match_oracle(oracle_mod, nb):
    """Per-tile first and double hits (ray_firstHit_inTile_kernel / ray_doubleHit_inTile_kernel,
    training/src/cuda_func.cu:130-444) through the host-built face lists against the oracle's walk
    over every nearby face: t, faces and both double-hit t bit-exact."""
    mesh = S.room_mesh(11, (0, 0, 0), (2, 2, 1), 5000, size_range=(0.02, 0.15))
    grid = S.make_scene(0, tile_shape=(2, 2, 1), res=8, occupancy=0.0)
    tm = D.TileFaceMesh(mesh.vertices, mesh.faces, grid.tile_centers, 1.0, num_block=nb)
    rng = np.random.default_rng(nb)
    eye = np.array([0.3, 0.4, 0.5], np.float32)
    n = 800
    d = rng.normal(0, 1, (n, 3))
    d[:50, 1:] = 0.0                                  # axis-parallel rays
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    vis = oracle_mod.trace_tiles(oracle_mod.OracleScene(grid), eye, d, 4)
    o = np.repeat(eye[None], n, 0)
    ref_t, ref_f = oracle_mod.mesh_hit_tiles(mesh.vertices, mesh.faces, grid.tile_centers, 1.0, nb, o, d, vis, 0)
    ref_d = oracle_mod.mesh_hit_tiles(mesh.vertices, mesh.faces, grid.tile_centers, 1.0, nb, o, d, vis, 1)
    hits = torch.full((n, 4), float("nan"), device=DEV)
    t, f = tm.hits(to_dev(o), to_dev(d), to_dev(vis), mode=0, hits=hits)
    dd = tm.hits(to_dev(o), to_dev(d), to_dev(vis), mode=1)
    t, f, dd = t.cpu().numpy(), f.cpu().numpy(), dd.cpu().numpy()
    assert np.array_equal(f, ref_f)
    assert np.array_equal(t.view(np.uint32), ref_t.view(np.uint32))    # NaN where no hit, in both
    assert np.array_equal(dd.view(np.uint32), ref_d.view(np.uint32))
    assert (f >= 0).sum() > 300 and (dd[..., 1] > 0).sum() > 200
