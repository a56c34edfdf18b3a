"""Seeded synthetic scenes in the renderer's tiled sparse-brick format.

The scene is a grid of cubic tiles (the reference's ``tile_IndexMap`` / ``centers``,
rendering/src/__init__.py:111-117).  Each tile holds an n^3 voxel grid cut into 8^3-voxel
bricks; a brick is either empty or fully occupied (coherent occupancy).  Per occupied
brick the data is stored as

* ``sigma``  float32 [n_bricks, 512]         density, voxel-major ((x*8+y)*8+z)
* ``feats``  f16/f32 [n_bricks, 512, F]       features, F contiguous per voxel

and ``brick_map`` maps (tile, brick cell) -> brick id or -1 (the role of the reference's
``block_IndexMap``, __init__.py:118).  Empty voxels read as zero.

Stated decisions for the synthetic inputs (the reference ships no data):

* random streams: numpy ``Generator(PCG64)``, one child ``SeedSequence`` per component,
  so the draws of one component never depend on another's size;
* features ~ N(0, 0.5^2) (0.5 is the standard deviation);
* density: ``LogNormal(0, 1)`` read as optical depth per voxel length, i.e. the stored
  world-space sigma is ``exp(N(0,1)) / voxel_size``; with this unit the T<1e-3 early
  termination is reached inside a few voxels, as on the reference's opaque scenes;
* MLP weights Xavier-uniform, biases 0; FP16 scenes round the weights to FP16 values so
  that FP16 storage is lossless and every precision sees the same network;
* the view encoding is ``Positional_Encoding(L=2)`` of the unit direction (15 values).
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

BRICK = 8
BRICK_VOX = BRICK ** 3


@dataclasses.dataclass
class Camera:
    H: int
    W: int
    origin: np.ndarray      # f32[3]
    c2w: np.ndarray         # f32[9] row-major 3x3, columns = right, down, forward
    intrinsic: np.ndarray   # f32[3] = (cx, cy, 1/focal)  (rendering/src/__init__.py:51)


@dataclasses.dataclass
class Mesh:
    vertices: np.ndarray    # f32[V,3]
    faces: np.ndarray       # i32[F,3]


@dataclasses.dataclass
class Scene:
    tile_shape: np.ndarray      # i32[3]
    scene_min: np.ndarray       # f32[3]
    tile_size: float
    tile_index_map: np.ndarray  # i32[prod(tile_shape)]
    tile_centers: np.ndarray    # f32[T,3]
    tile_res: np.ndarray        # i32[T]
    tile_brick_base: np.ndarray # i32[T]
    brick_map: np.ndarray       # i32[sum nb^3]
    sigma: np.ndarray           # f32[NB,512]
    feats: np.ndarray           # f16 or f32 [NB,512,F]
    mlp_in: int
    mlp_width: int
    mlp_layers: int
    mlp_params: np.ndarray      # f32[T,P]
    sample_step: float
    early_stop: float = 1e-3
    trans_th: float = 0.8
    max_tiles: int = 30

    @property
    def n_tiles(self) -> int:
        return int(self.tile_centers.shape[0])

    @property
    def feat_dim(self) -> int:
        return int(self.feats.shape[2])

    @property
    def feat_half(self) -> bool:
        return self.feats.dtype == np.float16

    @property
    def n_bricks(self) -> int:
        return int(self.sigma.shape[0])


def mlp_layer_dims(mlp_in: int, width: int, layers: int):
    """(in, out) of every Linear layer: in->width, (layers-2) x width->width, width->3."""
    dims = []
    i = mlp_in
    for l in range(layers):
        o = 3 if l == layers - 1 else width
        dims.append((i, o))
        i = o
    return dims


def mlp_param_count(mlp_in: int, width: int, layers: int) -> int:
    return sum(o * i + o for i, o in mlp_layer_dims(mlp_in, width, layers))


def mlp_unpack(params: np.ndarray, mlp_in: int, width: int, layers: int):
    """Split one tile's packed vector into [(W[out,in], b[out]), ...]."""
    out, off = [], 0
    for i, o in mlp_layer_dims(mlp_in, width, layers):
        W = params[off:off + o * i].reshape(o, i)
        off += o * i
        b = params[off:off + o]
        off += o
        out.append((W, b))
    return out


def _child_rngs(seed: int, n: int):
    ss = np.random.SeedSequence(seed)
    return [np.random.Generator(np.random.PCG64(s)) for s in ss.spawn(n)]


def make_scene(seed: int, tile_shape=(1, 1, 1), tile_size=1.0, scene_min=(0.0, 0.0, 0.0),
               res=64, occupancy=0.1, feat_dim=16, feat_half=False, mlp_width=64, mlp_layers=3,
               sample_step=1.0 / 256, early_stop=1e-3, max_tiles=30, res_choices=None,
               occupancy_range=None) -> Scene:
    """Seeded tiled scene.  ``res_choices``/``occupancy_range`` draw per-tile values."""
    r_occ, r_sig, r_feat, r_mlp, r_tile = _child_rngs(seed, 5)
    tile_shape = np.asarray(tile_shape, np.int32)
    Tx, Ty, Tz = (int(v) for v in tile_shape)
    T = Tx * Ty * Tz
    tile_index_map = np.arange(T, dtype=np.int32)
    cells = np.stack(np.meshgrid(np.arange(Tx), np.arange(Ty), np.arange(Tz), indexing="ij"), -1).reshape(-1, 3)
    smin = np.asarray(scene_min, np.float32)
    ts = np.float32(tile_size)
    tile_centers = (smin[None, :] + (cells.astype(np.float32) + np.float32(0.5)) * ts).astype(np.float32)
    if res_choices is not None:
        tile_res = r_tile.choice(np.asarray(res_choices, np.int32), size=T).astype(np.int32)
    else:
        tile_res = np.full(T, res, np.int32)
    if occupancy_range is not None:
        occ = r_tile.uniform(occupancy_range[0], occupancy_range[1], size=T)
    else:
        occ = np.full(T, occupancy)
    assert np.all(tile_res % BRICK == 0)
    nb = tile_res // BRICK
    counts = nb.astype(np.int64) ** 3
    base = np.concatenate([[0], np.cumsum(counts)[:-1]]).astype(np.int32)
    occupied = np.concatenate([r_occ.random(int(c)) < p for c, p in zip(counts, occ)])
    brick_map = np.full(int(counts.sum()), -1, np.int32)
    nbrk = int(occupied.sum())
    brick_map[occupied] = np.arange(nbrk, dtype=np.int32)
    vs = (ts / tile_res.astype(np.float32)).astype(np.float32)
    # voxel size of every occupied brick's tile
    brick_tile = np.repeat(np.arange(T), counts)[occupied]
    sigma = np.empty((nbrk, BRICK_VOX), np.float32)
    feats = np.empty((nbrk, BRICK_VOX, feat_dim), np.float16 if feat_half else np.float32)
    chunk = 4096
    for s in range(0, nbrk, chunk):
        e = min(nbrk, s + chunk)
        z = r_sig.standard_normal((e - s, BRICK_VOX), dtype=np.float32)
        sigma[s:e] = (np.exp(z) / vs[brick_tile[s:e]][:, None]).astype(np.float32)
        f = r_feat.standard_normal((e - s, BRICK_VOX, feat_dim), dtype=np.float32) * np.float32(0.5)
        feats[s:e] = f.astype(feats.dtype)
    mlp_in = feat_dim + 15
    P = mlp_param_count(mlp_in, mlp_width, mlp_layers)
    params = np.empty((T, P), np.float32)
    for t in range(T):
        chunks = []
        for i, o in mlp_layer_dims(mlp_in, mlp_width, mlp_layers):
            a = math.sqrt(6.0 / (i + o))
            W = r_mlp.uniform(-a, a, size=(o, i)).astype(np.float32)
            chunks += [W.ravel(), np.zeros(o, np.float32)]
        params[t] = np.concatenate(chunks)
    if feat_half:
        params = params.astype(np.float16).astype(np.float32)
    return Scene(tile_shape=tile_shape, scene_min=smin, tile_size=float(ts), tile_index_map=tile_index_map,
                 tile_centers=tile_centers, tile_res=tile_res, tile_brick_base=base, brick_map=brick_map,
                 sigma=sigma, feats=feats, mlp_in=mlp_in, mlp_width=mlp_width, mlp_layers=mlp_layers,
                 mlp_params=params, sample_step=float(np.float32(sample_step)), early_stop=early_stop,
                 max_tiles=max_tiles)


def look_at_camera(eye, target, H, W, focal, world_up=(0.0, 0.0, 1.0)) -> Camera:
    """Pinhole camera; rays are d = R [ (px-cx)/f, (py-cy)/f, 1 ] (compute_rays.cu:7-16),
    so R's columns are right, down (image y grows downwards) and forward."""
    eye = np.asarray(eye, np.float64)
    fwd = np.asarray(target, np.float64) - eye
    fwd /= np.linalg.norm(fwd)
    up = np.asarray(world_up, np.float64)
    if abs(np.dot(fwd, up)) > 0.999:
        up = np.array([0.0, 1.0, 0.0])
    right = np.cross(fwd, up)
    right /= np.linalg.norm(right)
    down = np.cross(fwd, right)
    R = np.stack([right, down, fwd], axis=1)
    return Camera(H=H, W=W, origin=eye.astype(np.float32), c2w=R.astype(np.float32).ravel(),
                  intrinsic=np.array([W / 2.0, H / 2.0, 1.0 / focal], np.float32))


def box_mesh(lo: np.ndarray, hi: np.ndarray) -> Mesh:
    """12 triangles per axis-aligned box (lo, hi: [B,3])."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    B = lo.shape[0]
    corners = np.empty((B, 8, 3), np.float32)
    for c in range(8):
        corners[:, c, 0] = np.where(c & 1, hi[:, 0], lo[:, 0])
        corners[:, c, 1] = np.where(c & 2, hi[:, 1], lo[:, 1])
        corners[:, c, 2] = np.where(c & 4, hi[:, 2], lo[:, 2])
    quads = [(0, 2, 6, 4), (1, 5, 7, 3), (0, 4, 5, 1), (2, 3, 7, 6), (0, 1, 3, 2), (4, 6, 7, 5)]
    tri = []
    for a, b, c, d in quads:
        tri += [(a, b, c), (a, c, d)]
    tri = np.asarray(tri, np.int32)
    faces = (tri[None, :, :] + 8 * np.arange(B, dtype=np.int32)[:, None, None]).reshape(-1, 3)
    return Mesh(vertices=corners.reshape(-1, 3), faces=faces)


def room_mesh(seed: int, room_lo, room_hi, n_triangles: int, avoid=None, size_range=(0.01, 0.05)) -> Mesh:
    """Room walls (6 thin boxes enclosing the room) plus random furniture boxes up to
    ``n_triangles`` (12 per box).  Boxes containing the point ``avoid`` are redrawn."""
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed).spawn(6)[5]))
    lo = np.asarray(room_lo, np.float32)
    hi = np.asarray(room_hi, np.float32)
    th = np.float32(0.05)
    walls_lo, walls_hi = [], []
    for a in range(3):
        for side in range(2):
            wl = lo - th
            wh = hi + th
            if side == 0:
                wh = wh.copy(); wh[a] = lo[a]
            else:
                wl = wl.copy(); wl[a] = hi[a]
            walls_lo.append(wl); walls_hi.append(wh)
    nbox = max(0, -(-n_triangles // 12) - 6)
    size = rng.uniform(size_range[0], size_range[1], size=(nbox, 3)).astype(np.float32)
    pos = rng.uniform(lo, hi, size=(nbox, 3)).astype(np.float32)
    blo = np.minimum(np.maximum(pos - size / 2, lo), hi)
    bhi = np.minimum(np.maximum(pos + size / 2, lo), hi)
    if avoid is not None:
        a = np.asarray(avoid, np.float32)
        bad = np.all((blo - 0.05 <= a) & (a <= bhi + 0.05), axis=1)
        blo, bhi = blo[~bad], bhi[~bad]
    m = box_mesh(np.concatenate([np.asarray(walls_lo), blo]), np.concatenate([np.asarray(walls_hi), bhi]))
    return m


# --------------------------------------------------------------------------- configs
def config_c1(seed: int = 1):
    """Plain FP32 single tile: 64^3 over the unit cube, 8^3 bricks ~ Bernoulli(0.1), F=16 FP32,
    MLP 31->64->64->3, 256 samples per unit length, T<1e-3; 320x180 pinhole, 60 deg hFOV,
    on a seeded point of the radius-2 sphere around the cube centre."""
    sc = make_scene(seed, tile_shape=(1, 1, 1), tile_size=1.0, res=64, occupancy=0.1, feat_dim=16,
                    feat_half=False, mlp_width=64, mlp_layers=3, sample_step=1.0 / 256, early_stop=1e-3)
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed).spawn(7)[6]))
    v = rng.standard_normal(3)
    v /= np.linalg.norm(v)
    centre = np.array([0.5, 0.5, 0.5])
    W, H = 320, 180
    focal = (W / 2) / math.tan(math.radians(60) / 2)
    cam = look_at_camera(centre + 2.0 * v, centre, H, W, focal)
    return sc, cam, None


def room_camera(seed: int, room_lo, room_hi, H, W, focal=750.0) -> Camera:
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed).spawn(8)[7]))
    lo = np.asarray(room_lo, np.float64)
    hi = np.asarray(room_hi, np.float64)
    eye = lo + (hi - lo) * np.array([rng.uniform(0.2, 0.8), rng.uniform(0.2, 0.8), rng.uniform(0.35, 0.65)])
    yaw = rng.uniform(0, 2 * math.pi)
    pitch = rng.uniform(-0.15, 0.15)
    fwd = np.array([math.cos(yaw) * math.cos(pitch), math.sin(yaw) * math.cos(pitch), math.sin(pitch)])
    return look_at_camera(eye, eye + fwd, H, W, focal)


def config_c2(seed: int = 2, H: int = 720, W: int = 1280, n_triangles: int = 200_000, feat_half=True):
    """16-tile (4x4x1) room of 128^3 tiles at 5% brick occupancy, F=32 FP16 features, per-tile
    MLP 47->128->128->128->3; sample step 0.2 voxel (the reference's sample_step_scale,
    rendering/src/__init__.py:41); proxy mesh of room walls + furniture boxes (~200k triangles);
    camera at a seeded pose inside the room, focal 750 (rendering/config/base.yaml:7)."""
    sc = make_scene(seed, tile_shape=(4, 4, 1), tile_size=1.0, res=128, occupancy=0.05, feat_dim=32,
                    feat_half=feat_half, mlp_width=128, mlp_layers=4, sample_step=0.2 / 128, early_stop=1e-3)
    lo, hi = (0.0, 0.0, 0.0), (4.0, 4.0, 1.0)
    cam = room_camera(seed, lo, hi, H, W)
    mesh = room_mesh(seed, lo, hi, n_triangles, avoid=cam.origin) if n_triangles else None
    return sc, cam, mesh


def config_c3_rays(sc: Scene, seed: int = 3, rays_per_tile: int = 8192, room_lo=(0.0, 0.0, 0.0),
                   room_hi=(4.0, 4.0, 1.0)):
    """Training batch of the 16-tile room: for every tile, ``rays_per_tile`` rays from seeded points
    in the room towards seeded points inside that tile (origins f32[N,3], dirs f32[N,3])."""
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed).spawn(9)[8]))
    lo = np.asarray(room_lo, np.float64)
    hi = np.asarray(room_hi, np.float64)
    o, d = [], []
    for t in range(sc.n_tiles):
        c = sc.tile_centers[t].astype(np.float64)
        org = lo + (hi - lo) * rng.uniform(0.02, 0.98, (rays_per_tile, 3))
        tgt = c + rng.uniform(-0.5, 0.5, (rays_per_tile, 3)) * sc.tile_size
        o.append(org)
        d.append(tgt - org)
    return np.concatenate(o).astype(np.float32), np.concatenate(d).astype(np.float32)


def perturbed_copy(sc: Scene, seed: int, rel: float = 0.01) -> Scene:
    """Target scene of the training config: features and MLP weights perturbed by
    N(0, rel^2) * |value| (density unchanged)."""
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed).spawn(10)[9]))
    c = dataclasses.replace(sc)
    f = sc.feats.astype(np.float32)
    c.feats = (f + rng.standard_normal(f.shape, dtype=np.float32) * np.float32(rel) * np.abs(f)).astype(sc.feats.dtype)
    m = sc.mlp_params
    c.mlp_params = (m + rng.standard_normal(m.shape, dtype=np.float32) * np.float32(rel) * np.abs(m)).astype(np.float32)
    return c


def config_c4(seed: int = 4, n_rays: int = 4_000_000, zipf_s: float = 1.2, tile_shape=(4, 4, 4)):
    """Batched uneven tiles: 64 tiles (4x4x4 of unit size) with grids drawn from {64, 96, 128}^3 and
    brick occupancy from U(0.02, 0.15), F=32 FP16 features, per-tile width-128 x4 MLP (run in BF16
    on MFMA).  Per-tile ray counts follow Zipf(zipf_s) over a seeded ranking of the tiles and sum to
    n_rays; every ray starts at a seeded point of the scene volume and points at a seeded point inside
    its tile.  Returns (scene, origins f32[N,3], dirs f32[N,3], rays per tile i64[T])."""
    sc = make_scene(seed, tile_shape=tile_shape, tile_size=1.0, res_choices=(64, 96, 128),
                    occupancy_range=(0.02, 0.15), feat_dim=32, feat_half=True, mlp_width=128, mlp_layers=4,
                    sample_step=0.2 / 128, early_stop=1e-3)
    rng = np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed).spawn(11)[10]))
    T = sc.n_tiles
    rank = rng.permutation(T) + 1
    w = rank.astype(np.float64) ** (-zipf_s)
    counts = np.floor(n_rays * w / w.sum()).astype(np.int64)
    counts[np.argsort(rank)[: n_rays - int(counts.sum())]] += 1     # remainder to the top ranks
    hi = np.asarray(tile_shape, np.float64) * sc.tile_size + sc.scene_min
    lo = sc.scene_min.astype(np.float64)
    tile_of_ray = np.repeat(np.arange(T), counts)
    org = lo + (hi - lo) * rng.uniform(0.0, 1.0, (n_rays, 3))
    tgt = sc.tile_centers[tile_of_ray].astype(np.float64) + rng.uniform(-0.5, 0.5, (n_rays, 3)) * sc.tile_size
    perm = rng.permutation(n_rays)           # rays of all tiles interleaved, as a batch would arrive
    return sc, org[perm].astype(np.float32), (tgt - org)[perm].astype(np.float32), counts


def config_c5(seed: int = 5, H: int = 1080, W: int = 1920, n_views: int = 4, n_triangles: int = 2_000_000):
    """256 tiles (16x16 floor plan) of 128^3 at 4%, F=32 FP16, width-128x4 MLP, 2M-triangle mesh,
    n_views seeded 1920x1080 views."""
    sc = make_scene(seed, tile_shape=(16, 16, 1), tile_size=1.0, res=128, occupancy=0.04, feat_dim=32,
                    feat_half=True, mlp_width=128, mlp_layers=4, sample_step=0.2 / 128, early_stop=1e-3)
    lo, hi = (0.0, 0.0, 0.0), (16.0, 16.0, 1.0)
    cams = [room_camera(seed * 100 + v, lo, hi, H, W) for v in range(n_views)]
    mesh = room_mesh(seed, lo, hi, n_triangles, avoid=cams[0].origin)
    return sc, cams, mesh


def camera_rays_np(cam: Camera) -> np.ndarray:
    """float32 restatement of compute_rays (compute_rays.cu:7-44) for host-side use."""
    py, px = np.meshgrid(np.arange(cam.H, dtype=np.float32), np.arange(cam.W, dtype=np.float32), indexing="ij")
    x = (px - cam.intrinsic[0]) * cam.intrinsic[2]
    y = (py - cam.intrinsic[1]) * cam.intrinsic[2]
    R = cam.c2w.reshape(3, 3)
    d = np.stack([R[r, 0] * x + R[r, 1] * y + R[r, 2] for r in range(3)], -1)
    return d.reshape(-1, 3).astype(np.float32)


# ---------------------------------------------------------------------------- reflection pass
# The reference's second render pass (rendering/src/__init__.py:272-335): per downscaled pixel, 64
# inverse-depth samples behind the first surface, a per-group MLP (inference_mlp) and compositing.

REF_MLP_DIMS = ((63, 64), (64, 64), (64, 64), (64, 64), (64, 64), (64, 28))
REF_MLP_PARAMS = sum(o + i * o for i, o in REF_MLP_DIMS)  # 22,556 halves per group


@dataclasses.dataclass
class ReflectionFrame:
    origin: np.ndarray          # f32[3]
    rays_d: np.ndarray          # f32[H, W, 3] full-resolution directions (compute_rays, not normalised)
    inverse_near: np.ndarray    # f32[DH, DW] 1 / depth of the first surface (1e-8: no surface)
    net_idx: np.ndarray         # i16[DH, DW] group of each downscaled pixel, -1 = none
    params: np.ndarray          # f16[G * 22556] reference packing: per layer bias[out], then W[in][out]
    centers: np.ndarray         # f32[G, 3]
    far: float                  # scene diagonal (rendering/src/__init__.py:132)
    trans_stop: float           # early_stop_R (rendering/src/__init__.py:39)
    scale: int                  # reflection_scale


def make_reflection_groups(seed: int, n_groups: int, room_lo, room_hi):
    """Per group a reflection MLP in the reference packing (Kaiming-uniform hidden weights, Xavier
    output layer, biases ~ N(0, 0.05^2), all rounded to FP16) and a centre inside the room."""
    rs = np.random.SeedSequence(seed).spawn(2)
    rng = np.random.Generator(np.random.PCG64(rs[0]))
    out = np.empty(n_groups * REF_MLP_PARAMS, np.float16)
    off = 0
    for _ in range(n_groups):
        for li, (i, o) in enumerate(REF_MLP_DIMS):
            a = math.sqrt(6.0 / i) if li < len(REF_MLP_DIMS) - 1 else math.sqrt(6.0 / (i + o))
            out[off:off + o] = rng.normal(0.0, 0.05, o)
            off += o
            out[off:off + i * o] = rng.uniform(-a, a, i * o)
            off += i * o
    lo = np.asarray(room_lo, np.float64)
    hi = np.asarray(room_hi, np.float64)
    crng = np.random.Generator(np.random.PCG64(rs[1]))
    centers = (lo + (hi - lo) * crng.uniform(0.1, 0.9, (n_groups, 3))).astype(np.float32)
    return out, centers


def reflection_frame(seed: int = 6, H: int = 720, W: int = 1280, n_groups: int = 48, scale: int = 2,
                     room_lo=(0.0, 0.0, 0.0), room_hi=(4.0, 4.0, 1.0), no_hit: float = 0.05) -> ReflectionFrame:
    """Synthetic inputs of the reflection pass at H x W (downscaled by ``scale``): the pixel groups
    are the Voronoi cells of ``n_groups`` seeded image points (screen regions, as a group of tiles
    covers on a real frame), each cell at a seeded depth in [0.8, 4] plus a gentle slope; a
    ``no_hit`` fraction of cells has no surface (net -1, inverse_near 1e-8, the reference's init)."""
    rs = np.random.SeedSequence(seed).spawn(3)
    rng = np.random.Generator(np.random.PCG64(rs[0]))
    cam = room_camera(seed, room_lo, room_hi, H, W)
    rays = camera_rays_np(cam).reshape(H, W, 3)
    DH, DW = H // scale, W // scale
    n_cells = n_groups + max(1, int(round(no_hit * n_groups)))
    pts = rng.uniform(0.0, 1.0, (n_cells, 2)) * np.array([DH, DW])
    yy, xx = np.meshgrid(np.arange(DH, dtype=np.float64), np.arange(DW, dtype=np.float64), indexing="ij")
    d2 = (yy[..., None] - pts[:, 0]) ** 2 + (xx[..., None] - pts[:, 1]) ** 2
    cell = np.argmin(d2, axis=-1)
    group_of_cell = np.concatenate([rng.permutation(n_groups), -np.ones(n_cells - n_groups, np.int64)])
    net = group_of_cell[cell].astype(np.int16)
    depth0 = rng.uniform(0.8, 4.0, n_cells)
    slope = rng.uniform(-1.0, 1.0, (n_cells, 2)) / max(DH, DW)
    depth = depth0[cell] + slope[cell, 0] * (yy - pts[cell, 0]) + slope[cell, 1] * (xx - pts[cell, 1])
    inv_near = (1.0 / np.maximum(depth, 0.3)).astype(np.float32)
    inv_near[net < 0] = np.float32(1e-8)
    params, centers = make_reflection_groups(seed + 1000, n_groups, room_lo, room_hi)
    far = float(np.linalg.norm(np.asarray(room_hi, np.float64) - np.asarray(room_lo, np.float64)))
    return ReflectionFrame(cam.origin.astype(np.float32), rays, inv_near, net, params, centers, far, 0.01, scale)


def reflection_blocks(net_flat: np.ndarray, max_pixels_per_block: int = 64):
    """numpy form of the reference's pixel regrouping (rendering/src/__init__.py:272-291): stable sort
    of the downscaled pixels by group; each group's pixels cut into ceil(n/64) blocks of sizes that
    differ by at most one.  Returns query_pixel_indices i32[P], net i16[B], pixel_starts i32[B+1]
    (groups include -1, whose blocks the MLP skips)."""
    net_flat = np.asarray(net_flat, np.int16).ravel()
    P = net_flat.size
    order = np.argsort(net_flat, kind="stable").astype(np.int32)
    ks = net_flat[order]
    heads = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]])) if P else np.zeros(0, np.int64)
    starts = np.concatenate([heads, [P]]).astype(np.int64)
    nets, pstarts = [], []
    for g in range(heads.size):
        a, b = starts[g], starts[g + 1]
        nb = -(-(b - a) // max_pixels_per_block)
        base, add = divmod(b - a, nb)
        p = a
        for k in range(nb):
            nets.append(ks[a])
            pstarts.append(p)
            p += base + 1 if k < add else base
    pstarts.append(P)
    return order, np.asarray(nets, np.int16), np.asarray(pstarts, np.int32)


# ---------------------------------------------------------------------------- octree format
# The reference's own diffuse scene format (rendering/src/__init__.py:111-121, consumed by
# rendering_diffuse_octree_fp16): per tile 4^3 blocks; an occupied block is a node holding an
# 18^3 dilated cell grid (16^3 voxels plus a one-voxel border copied from the neighbours).

OCT_BLOCKS = 4
OCT_NV = 16


@dataclasses.dataclass
class OctreeScene:
    tile_shape: tuple
    tile_size: float
    scene_min: np.ndarray        # f32[3]
    tile_index_map: np.ndarray   # i32[prod(tile_shape)] dense tile -> tile id, -1 = empty
    centers: np.ndarray          # f32[T, 3]
    group_map: np.ndarray        # i32[T] tile -> group
    block_index_map: np.ndarray  # i32[T * 64] block -> node, -1 = empty
    nodes_index_map: np.ndarray  # i16[nodes * 18^3] cell -> voxel within the node, -1 = empty
    nodes_sample_flag: np.ndarray  # bool[nodes * 18^3]
    voxels_start: np.ndarray     # i64[nodes]
    data_voxels: np.ndarray      # f16[V, 4] (r, g, b, sigma)
    voxel_size: float
    sample_step: float
    num_voxel: int = OCT_NV + 2

    @property
    def n_tiles(self):
        return self.centers.shape[0]

    @property
    def n_nodes(self):
        return self.voxels_start.shape[0]


def make_octree_scene(seed: int, tile_shape=(4, 4, 1), tile_size=1.0, scene_min=(0.0, 0.0, 0.0),
                      block_occupancy=0.3, voxel_occupancy=0.5, n_groups=4, sample_step_scale=0.2,
                      field=None, flag_border=False) -> OctreeScene:
    """Seeded scene in the reference's octree format.  Blocks are occupied ~ Bernoulli(block_occupancy),
    voxels of an occupied block ~ Bernoulli(voxel_occupancy); rgb ~ U(0,1), sigma = exp(N(0,1)) /
    voxel_size (FP16).  A cell is sampled when it is an interior cell whose voxel is occupied.  The
    dilated border holds the neighbouring voxels of the same tile (zero beyond the tile).
    ``field(tile, ijk) -> f32[...,4]`` overrides the voxel values (used by closed-form tests); with
    it every voxel of an occupied block is occupied and the border beyond the tile holds the field's
    values at voxel coordinates -1 and 64 (the field extends past the tile).  ``flag_border`` also
    flags the border cells, so that a sample whose cell rounds onto the border is taken, not skipped."""
    rs = np.random.SeedSequence(seed).spawn(4)
    rng_b = np.random.Generator(np.random.PCG64(rs[0]))
    rng_v = np.random.Generator(np.random.PCG64(rs[1]))
    rng_d = np.random.Generator(np.random.PCG64(rs[2]))
    rng_g = np.random.Generator(np.random.PCG64(rs[3]))
    tile_shape = tuple(int(v) for v in tile_shape)
    nt = int(np.prod(tile_shape))
    smin = np.asarray(scene_min, np.float32)
    R = OCT_BLOCKS * OCT_NV
    vs = float(np.float32(tile_size) / np.float32(R))
    centers = np.empty((nt, 3), np.float32)
    for t in range(nt):
        ix, iy, iz = t // (tile_shape[1] * tile_shape[2]), (t // tile_shape[2]) % tile_shape[1], t % tile_shape[2]
        centers[t] = smin + (np.array([ix, iy, iz], np.float32) + np.float32(0.5)) * np.float32(tile_size)
    M = OCT_NV + 2
    bmap = np.full(nt * OCT_BLOCKS ** 3, -1, np.int32)
    nodes_idx, nodes_flag, starts, data = [], [], [], []
    n_vox = 0
    ijk = np.stack(np.meshgrid(np.arange(M), np.arange(M), np.arange(M), indexing="ij"), -1).reshape(-1, 3)
    for t in range(nt):
        occ_b = rng_b.random(OCT_BLOCKS ** 3) < block_occupancy
        vox_occ = np.zeros((R + 2,) * 3, bool)   # tile voxels with a zero border
        vox_val = np.zeros((R + 2,) * 3 + (4,), np.float32)
        for b in np.flatnonzero(occ_b):
            bx, by, bz = b // 16, (b // 4) % 4, b % 4
            sl = tuple(slice(1 + OCT_NV * c, 1 + OCT_NV * (c + 1)) for c in (bx, by, bz))
            if field is None:
                vox_occ[sl] = rng_v.random((OCT_NV,) * 3) < voxel_occupancy
            else:
                vox_occ[sl] = True
        if field is not None:
            vox_occ[:] = True
        occ_idx = np.argwhere(vox_occ)
        if field is None:
            vals = np.empty((occ_idx.shape[0], 4), np.float32)
            vals[:, :3] = rng_d.random((occ_idx.shape[0], 3))
            vals[:, 3] = np.exp(rng_d.standard_normal(occ_idx.shape[0])) / vs
        else:
            vals = np.asarray(field(t, occ_idx - 1), np.float32)
        vox_val[tuple(occ_idx.T)] = vals
        for b in np.flatnonzero(occ_b):
            bx, by, bz = b // 16, (b // 4) % 4, b % 4
            cells = ijk + np.array([OCT_NV * bx, OCT_NV * by, OCT_NV * bz])   # padded tile coordinates
            occ = vox_occ[tuple(cells.T)]
            idx = np.full(M ** 3, -1, np.int16)
            idx[occ] = np.arange(int(occ.sum()), dtype=np.int16)
            interior = np.all((ijk >= 1) & (ijk <= OCT_NV), axis=1)
            bmap[t * OCT_BLOCKS ** 3 + b] = len(starts)
            nodes_idx.append(idx)
            nodes_flag.append((interior | flag_border) & occ)
            starts.append(n_vox)
            data.append(vox_val[tuple(cells[occ].T)])
            n_vox += int(occ.sum())
    groups = rng_g.integers(0, n_groups, nt).astype(np.int32)
    data_v = (np.concatenate(data) if data else np.zeros((1, 4), np.float32)).astype(np.float16)
    if data_v.shape[0] == 0:
        data_v = np.zeros((1, 4), np.float16)
    return OctreeScene(tile_shape, float(tile_size), smin, np.arange(nt, dtype=np.int32), centers, groups, bmap,
                       np.concatenate(nodes_idx) if nodes_idx else np.full(M ** 3, -1, np.int16),
                       np.concatenate(nodes_flag) if nodes_flag else np.zeros(M ** 3, bool),
                       np.asarray(starts, np.int64) if starts else np.zeros(1, np.int64), data_v, vs,
                       float(np.float32(sample_step_scale) * np.float32(vs)))


# ---------------------------------------------------------------------------- dense training grids
# The training tiles' own format (training/scripts/tile/tile.py:415-432, written per tile as
# Voxel-T{tile}-E{epoch}.npy and Node-T{tile}-E{epoch}.npy): a dilated grid of NUM_VOXEL + 2 cells per
# side (one cell of overlap with the neighbours), float4 (r, g, b, sigma) in x, y, z, channel order and
# an int16 flag per cell: 1 sampled, 0 a value kept for interpolation only, -1 empty (zero value).

@dataclasses.dataclass
class DenseScene:
    tile_shape: tuple
    tile_size: float
    scene_min: np.ndarray        # f32[3]
    num_voxel: int               # NUM_VOXEL (voxels per tile side, without the overlap)
    tile_ids: np.ndarray         # i32[T] dense tile cell of each stored tile (x*Ty*Tz + y*Tz + z)
    centers: np.ndarray          # f32[T, 3]
    voxels: np.ndarray           # f32[T, nv, nv, nv, 4], nv = num_voxel + 2
    nodes: np.ndarray            # i16[T, nv, nv, nv]
    groups: np.ndarray           # i32[T]

    @property
    def voxel_size(self) -> float:
        return float(np.float32(self.tile_size) / np.float32(self.num_voxel))

    @property
    def tile_index_map(self) -> np.ndarray:
        m = np.full(int(np.prod(self.tile_shape)), -1, np.int32)
        m[self.tile_ids] = np.arange(self.tile_ids.size, dtype=np.int32)
        return m


def make_dense_scene(seed: int, tile_shape=(2, 2, 1), tile_size=1.0, scene_min=(0.0, 0.0, 0.0), num_voxel=64,
                     block=8, block_occupancy=0.25, voxel_occupancy=0.6, n_groups=2, drop_tiles=0) -> DenseScene:
    """Seeded scene in the training tiles' dense format.  A global voxel grid over the whole scene is
    occupied by blocks ~ Bernoulli(block_occupancy), voxels of an occupied block ~ Bernoulli(
    voxel_occupancy); occupied voxels are flagged 1, their 26-neighbours that are not occupied 0 (they
    keep a value, as the training bakes values wherever the flag is not -1), the rest -1 with zero
    value; rgb ~ U(0, 1), sigma = exp(N(0, 1)) / voxel_size.  Each tile stores the window of its
    voxels plus one cell of overlap on every side; overlap cells are flagged 0 where they hold a value
    (the training tiles keep only their interior flags, tile.py:405-413).  ``drop_tiles`` tiles are left
    out (empty tiles of a floor plan)."""
    rs = np.random.SeedSequence(seed).spawn(4)
    rng_b, rng_v, rng_d, rng_t = (np.random.Generator(np.random.PCG64(r)) for r in rs)
    tile_shape = tuple(int(v) for v in tile_shape)
    NV = int(num_voxel)
    vs = float(np.float32(tile_size) / np.float32(NV))
    G = np.asarray(tile_shape) * NV
    nb = G // block
    occ_b = rng_b.random(tuple(nb)) < block_occupancy
    occ = np.repeat(np.repeat(np.repeat(occ_b, block, 0), block, 1), block, 2)
    occ &= rng_v.random(tuple(G)) < voxel_occupancy
    pad = np.zeros(tuple(G + 2), bool)
    pad[1:-1, 1:-1, 1:-1] = occ
    dil = np.zeros_like(pad)
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                dil |= np.roll(np.roll(np.roll(pad, dx, 0), dy, 1), dz, 2)
    dil[0, :, :] = dil[-1, :, :] = False
    dil[:, 0, :] = dil[:, -1, :] = False
    dil[:, :, 0] = dil[:, :, -1] = False
    flag = np.where(pad, 1, np.where(dil, 0, -1)).astype(np.int16)
    val = np.zeros(tuple(G + 2) + (4,), np.float32)
    has = flag >= 0
    k = int(has.sum())
    val[has, :3] = rng_d.random((k, 3))
    val[has, 3] = np.exp(rng_d.standard_normal(k)) / vs
    nt = int(np.prod(tile_shape))
    keep = np.sort(rng_t.permutation(nt)[: nt - int(drop_tiles)]).astype(np.int32)
    smin = np.asarray(scene_min, np.float32)
    nv = NV + 2
    vox = np.empty((keep.size, nv, nv, nv, 4), np.float32)
    nodes = np.empty((keep.size, nv, nv, nv), np.int16)
    centers = np.empty((keep.size, 3), np.float32)
    for j, t in enumerate(keep):
        ix, iy, iz = t // (tile_shape[1] * tile_shape[2]), (t // tile_shape[2]) % tile_shape[1], t % tile_shape[2]
        sl = tuple(slice(c * NV, c * NV + nv) for c in (ix, iy, iz))   # padded global coordinates
        vox[j] = val[sl]
        f = flag[sl].copy()
        border = np.ones((nv, nv, nv), bool)
        border[1:-1, 1:-1, 1:-1] = False
        f[border & (f == 1)] = 0          # overlap cells are never sampled by this tile
        nodes[j] = f
        centers[j] = smin + (np.array([ix, iy, iz], np.float32) + np.float32(0.5)) * np.float32(tile_size)
    groups = (np.arange(keep.size) % max(1, n_groups)).astype(np.int32)
    return DenseScene(tile_shape, float(tile_size), smin, NV, keep, centers, vox, nodes, groups)
