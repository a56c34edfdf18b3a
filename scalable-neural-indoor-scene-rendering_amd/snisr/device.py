"""Torch-facing host layer: device-resident scenes and the render entry points.

PyTorch provides device memory and the current stream; all compute runs in the HIP
kernels of libsnisr_hip.so through the C ABI (include/snisr.h).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _native as N
from .scene import Scene


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_cuda(t: torch.Tensor, name: str, dtype=None):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if dtype is not None and t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class DeviceScene:
    """A ``Scene`` uploaded to HBM with its C descriptor and packed MLP copies."""

    def __init__(self, sc: Scene, device="cuda", mfma_dtype: int | None = None):
        dev = torch.device(device)
        self.host = sc
        t = lambda a, dt=None: torch.from_numpy(np.ascontiguousarray(a)).to(dev) if dt is None else \
            torch.from_numpy(np.ascontiguousarray(a)).to(dev, dt)
        self.tile_index_map = t(sc.tile_index_map)
        self.tile_centers = t(sc.tile_centers)
        self.tile_res = t(sc.tile_res)
        self.tile_brick_base = t(sc.tile_brick_base)
        self.brick_map = t(sc.brick_map)
        self.sigma = t(sc.sigma) if sc.n_bricks else torch.zeros((1, 512), device=dev)
        self.feats = t(sc.feats) if sc.n_bricks else torch.zeros((1, 512, sc.feat_dim), device=dev,
                                                                 dtype=torch.float16 if sc.feat_half else torch.float32)
        self.mlp_params = t(sc.mlp_params)
        self.mlp_f32t = None
        self.mlp_packed = None
        self.desc = self._make_desc()
        self.pack_f32t()
        if mfma_dtype is not None:
            self.pack_mfma(mfma_dtype)

    def _make_desc(self) -> N.SceneDesc:
        sc = self.host
        d = N.SceneDesc()
        d.tile_shape[:] = [int(v) for v in sc.tile_shape]
        d.scene_min[:] = [float(v) for v in sc.scene_min]
        d.tile_size = sc.tile_size
        d.n_tiles = sc.n_tiles
        d.n_bricks = sc.n_bricks
        d.feat_dim = sc.feat_dim
        d.feat_dtype = N.SNISR_F16 if self.feats.dtype == torch.float16 else N.SNISR_F32
        d.mlp_in, d.mlp_width, d.mlp_layers = sc.mlp_in, sc.mlp_width, sc.mlp_layers
        d.mlp_stride = self.mlp_params.shape[1]
        d.brick_map_len = self.brick_map.numel()
        d.tile_index_map = self.tile_index_map.data_ptr()
        d.tile_centers = self.tile_centers.data_ptr()
        d.tile_res = self.tile_res.data_ptr()
        d.tile_brick_base = self.tile_brick_base.data_ptr()
        d.brick_map = self.brick_map.data_ptr()
        d.sigma = self.sigma.data_ptr()
        d.feats = self.feats.data_ptr()
        d.mlp_params = self.mlp_params.data_ptr()
        return d

    def set_params(self, mlp_params: torch.Tensor | None = None, feats: torch.Tensor | None = None,
                   sigma: torch.Tensor | None = None):
        """Point the descriptor at new device tensors of the same shapes (training / perturbed copies)."""
        if mlp_params is not None:
            _require_cuda(mlp_params, "mlp_params", torch.float32)
            assert mlp_params.shape == self.mlp_params.shape
            self.mlp_params = mlp_params
            self.desc.mlp_params = mlp_params.data_ptr()
            self.pack_f32t()
            if self.mlp_packed is not None:
                self.pack_mfma(self.desc.mlp_packed_dtype)
        if feats is not None:
            _require_cuda(feats, "feats")
            assert feats.shape == self.feats.shape
            self.feats = feats
            self.desc.feats = feats.data_ptr()
            self.desc.feat_dtype = N.SNISR_F16 if feats.dtype == torch.float16 else N.SNISR_F32
        if sigma is not None:
            _require_cuda(sigma, "sigma", torch.float32)
            assert sigma.shape == self.sigma.shape
            self.sigma = sigma
            self.desc.sigma = sigma.data_ptr()

    def pack_f32t(self):
        L = N.lib()
        nbytes = L.snisr_pack_mlp_size(C.byref(self.desc), N.SNISR_F32)
        self.mlp_f32t = torch.empty(nbytes // 4, dtype=torch.float32, device=self.mlp_params.device)
        N.check(L.snisr_pack_mlp(C.byref(self.desc), N.SNISR_F32, _ptr(self.mlp_f32t), _stream()), "pack_mlp(f32)")
        self.desc.mlp_f32t = self.mlp_f32t.data_ptr()

    def pack_mfma(self, dtype: int = N.SNISR_F16):
        L = N.lib()
        nbytes = L.snisr_pack_mlp_size(C.byref(self.desc), dtype)
        if nbytes <= 0:
            raise RuntimeError(f"MFMA packing unsupported for this MLP shape: {N.lib().snisr_last_error().decode()}")
        self.mlp_packed = torch.empty(nbytes, dtype=torch.uint8, device=self.mlp_params.device)
        N.check(L.snisr_pack_mlp(C.byref(self.desc), dtype, _ptr(self.mlp_packed), _stream()), "pack_mlp(mfma)")
        self.desc.mlp_packed = self.mlp_packed.data_ptr()
        self.desc.mlp_packed_dtype = dtype

    def march_params(self, max_tiles=None, sample_step=None, early_stop=None, trans_th=None) -> N.MarchParams:
        sc = self.host
        p = N.MarchParams()
        p.max_tiles = sc.max_tiles if max_tiles is None else max_tiles
        p.sample_step = sc.sample_step if sample_step is None else sample_step
        p.early_stop = sc.early_stop if early_stop is None else early_stop
        p.trans_th = sc.trans_th if trans_th is None else trans_th
        return p


def _rays_desc(origins, dirs, t_hit=None, visited=None, image_width=0):
    _require_cuda(origins, "origins", torch.float32)
    _require_cuda(dirs, "dirs", torch.float32)
    r = N.RaysDesc()
    n = dirs.numel() // 3
    r.origins = origins.data_ptr()
    r.origin_stride = 0 if origins.numel() == 3 else 3
    if r.origin_stride == 3 and origins.numel() != 3 * n:
        raise ValueError("origins must be [3] or [N,3]")
    r.image_width = image_width
    r.n = n
    r.dirs = dirs.data_ptr()
    if t_hit is not None:
        _require_cuda(t_hit, "t_hit", torch.float32)
        r.t_hit = t_hit.data_ptr()
    if visited is not None:
        _require_cuda(visited, "visited", torch.int32)
        r.visited = visited.data_ptr()
    return r


def _attach_events(o: N.RenderOut, events):
    """Point the output descriptor at torch.cuda.Event handles (events must have been recorded once)."""
    if not events:
        return None
    arr = (C.c_void_p * len(events))(*[C.c_void_p(e.cuda_event) for e in events])
    o.n_stage_events = len(events)
    o.stage_events = C.cast(arr, C.c_void_p)
    return arr


def _render_outputs(n, device, rgb=True, brick_cap=0):
    out = dict(
        rgb=torch.empty((n, 3), dtype=torch.float32, device=device) if rgb else None,
        T=torch.empty(n, dtype=torch.float32, device=device),
        n_samples=torch.empty(n, dtype=torch.int32, device=device),
        depth=torch.empty(n, dtype=torch.float32, device=device),
    )
    if brick_cap:
        out["bricks"] = torch.empty((n, brick_cap), dtype=torch.int32, device=device)
    o = N.RenderOut()
    o.rgb = out["rgb"].data_ptr() if rgb else None
    o.T = out["T"].data_ptr()
    o.n_samples = out["n_samples"].data_ptr()
    o.depth = out["depth"].data_ptr()
    if brick_cap:
        o.bricks = out["bricks"].data_ptr()
        o.brick_cap = brick_cap
    return out, o


def render_fp32(ds: DeviceScene, origins, dirs, t_hit=None, visited=None, image_width=0, rgb=True, brick_cap=0,
                events=None, **march_kw):
    """One-thread-per-ray FP32 render (snisr_render_fp32)."""
    r = _rays_desc(origins, dirs, t_hit, visited, image_width)
    p = ds.march_params(**march_kw)
    if visited is not None and visited.shape[-1] != p.max_tiles:
        raise ValueError("visited must be [N, max_tiles]")
    out, o = _render_outputs(r.n, dirs.device, rgb, brick_cap)
    keep = _attach_events(o, events)  # noqa: F841
    N.check(N.lib().snisr_render_fp32(C.byref(ds.desc), C.byref(p), C.byref(r), C.byref(o), _stream()),
            "render_fp32")
    return out


class MfmaWorkspace:
    """Caller-owned workspace of snisr_render_mfma, grown on demand.  It keeps the record slots per
    ray (``cpr``) and the sorted-sample capacity it was last sized for; requests only grow them."""

    def __init__(self):
        self.buf = None
        self.cpr = 0
        self.samples = 0
        self.chunks = 1
        self.max_count = 0   # largest per-ray sample count seen by a checked render

    def get(self, ds: DeviceScene, n: int, records_per_ray: int = 0, samples: int = 0, chunks: int = 1):
        cpr = max(int(records_per_ray), self.cpr) or 64
        cap = max(int(samples), self.samples) or 16 * max(n, 1)
        k = max(int(chunks), self.chunks, 1)
        tn = N.Tuning(0, 512, 512, cpr, k)
        need = N.lib().snisr_render_mfma_workspace_size_tuned(C.byref(ds.desc), n, C.byref(tn), cap)
        if need < 0:
            raise RuntimeError(N.lib().snisr_last_error().decode())
        if self.buf is None or self.buf.numel() < need:
            self.buf = None
            self.buf = torch.empty(need, dtype=torch.uint8, device=ds.sigma.device)
        self.cpr, self.samples, self.chunks = cpr, cap, k
        return self.buf, self.buf.numel()


def mfma_stats(buf, with_max: bool = False):
    """(total samples, overflow flag[, largest per-ray count]) of the last render on this workspace
    (synchronises)."""
    total, over, mx = C.c_int64(0), C.c_int32(0), C.c_int32(0)
    N.check(N.lib().snisr_render_mfma_stats(_ptr(buf), C.byref(total), C.byref(over), C.byref(mx), _stream()),
            "render_mfma_stats")
    return (total.value, over.value, mx.value) if with_max else (total.value, over.value)


def records_for(max_count: int, headroom: float = 1.25) -> int:
    """Per-ray record slots for a workload whose largest ray has ``max_count`` samples."""
    return max(8, -(-int(max_count * headroom + 1) // 8) * 8)


def schedule_records(t: dict | None, max_count: int = 0) -> int:
    """Record slots per ray a schedule asks for: ``records`` "tight" (the largest per-ray count seen
    + 25%), "loose" (twice it) or a number; 0 when the schedule leaves it to the workspace."""
    r = (t or {}).get("records")
    if r in ("tight", "loose"):
        return records_for(max_count, 1.25 if r == "tight" else 2.0) if max_count else 0
    return int(r) if r else 0


def tuning_struct(t: dict | None, records_per_ray: int = 0):
    """snisr_tuning from a dict {ray_order, mlp_threads, sort_grid, chunks} and the record slots per
    ray."""
    t = t or {}
    return N.Tuning(int(t.get("ray_order", 0)), int(t.get("mlp_threads", 512)), int(t.get("sort_grid", 512)),
                    int(records_per_ray), int(t.get("chunks", 1)))


def render_mfma(ds: DeviceScene, origins, dirs, t_hit=None, visited=None, image_width=0, brick_cap=0,
                workspace: MfmaWorkspace | None = None, capacity: int = 0, sample_capacity: int = 0,
                check: bool = True, events=None, tuning: dict | None = None, outputs: dict | None = None,
                **march_kw):
    """Render with the per-tile MLP on MFMA (snisr_render_mfma_tuned).  ``capacity``: sample record
    slots per ray, ``sample_capacity``: samples of the frame (0 = keep the workspace's).  With
    ``check`` the capacities are verified after the call and the render repeated with room for the
    largest ray and the whole frame.  ``outputs`` (from a previous call) is written in place."""
    if ds.mlp_packed is None:
        raise RuntimeError("scene has no MFMA-packed MLP; call DeviceScene.pack_mfma()")
    r = _rays_desc(origins, dirs, t_hit, visited, image_width)
    p = ds.march_params(**march_kw)
    if outputs is None:
        out, o = _render_outputs(r.n, dirs.device, True, brick_cap)
    else:
        out, o = outputs, N.RenderOut()
        o.rgb, o.T = out["rgb"].data_ptr(), out["T"].data_ptr()
        o.n_samples, o.depth = out["n_samples"].data_ptr(), out["depth"].data_ptr()
    keep = _attach_events(o, events)  # noqa: F841
    ws = workspace or MfmaWorkspace()
    rec = schedule_records(tuning, ws.max_count)
    chunks = int((tuning or {}).get("chunks", 1))
    buf, nbytes = ws.get(ds, r.n, max(int(capacity), rec), sample_capacity, chunks)
    tn = tuning_struct(tuning, rec or ws.cpr)
    N.check(N.lib().snisr_render_mfma_tuned(C.byref(ds.desc), C.byref(p), C.byref(r), C.byref(o), C.byref(tn),
                                            _ptr(buf), nbytes, _stream()), "render_mfma")
    if check:
        total, over, mx = mfma_stats(buf, with_max=True)
        if over:
            ws.max_count = max(ws.max_count, mx)
            # the schedule's slots (e.g. twice the largest count) or the largest count + 25%
            rec = max(schedule_records(tuning, ws.max_count), records_for(mx))
            buf, nbytes = ws.get(ds, r.n, rec, int(total * 1.25) + 64, chunks)
            tn = tuning_struct(tuning, rec if (tuning or {}).get("records") else ws.cpr)
            N.check(N.lib().snisr_render_mfma_tuned(C.byref(ds.desc), C.byref(p), C.byref(r), C.byref(o),
                                                    C.byref(tn), _ptr(buf), nbytes, _stream()), "render_mfma")
            total, over, mx = mfma_stats(buf, with_max=True)
            assert not over
        out["total_samples"] = total
        out["max_samples_per_ray"] = mx
        ws.max_count = max(ws.max_count, mx)
    return out


def compute_rays(intrinsic: torch.Tensor, c2w: torch.Tensor, H: int, W: int, out: torch.Tensor | None = None):
    _require_cuda(intrinsic, "intrinsic", torch.float32)
    _require_cuda(c2w, "c2w", torch.float32)
    if out is None:
        out = torch.empty((H, W, 3), dtype=torch.float32, device=intrinsic.device)
    _require_cuda(out, "rays_d", torch.float32)
    N.check(N.lib().snisr_compute_rays(_ptr(intrinsic), _ptr(c2w), H, W, _ptr(out), _stream()), "compute_rays")
    return out


def trace_tiles(ds: DeviceScene, origin: torch.Tensor, rays_d: torch.Tensor, K: int, out=None):
    _require_cuda(origin, "origin", torch.float32)
    _require_cuda(rays_d, "rays_d", torch.float32)
    n = rays_d.numel() // 3
    if out is None:
        out = torch.empty((n, K), dtype=torch.int32, device=rays_d.device)
    _require_cuda(out, "visitedTiles", torch.int32)
    N.check(N.lib().snisr_trace_tiles(C.byref(ds.desc), _ptr(origin), _ptr(rays_d), n, K, _ptr(out), _stream()),
            "trace_tiles")
    return out


class DeviceMesh:
    """Proxy mesh as a BVH in HBM (built on the host by snisr_bvh_build)."""

    def __init__(self, vertices: np.ndarray, faces: np.ndarray, device="cuda"):
        L = N.lib()
        V = np.ascontiguousarray(vertices, np.float32)
        F = np.ascontiguousarray(faces, np.int32)
        nb, tb = C.c_int64(0), C.c_int64(0)
        N.check(L.snisr_bvh_build_bounds(F.shape[0], C.byref(nb), C.byref(tb)), "bvh_build_bounds")
        nodes = np.empty(nb.value // 4, np.float32)
        tris = np.empty(max(tb.value, 64) // 4, np.float32)
        nn, depth = C.c_int32(0), C.c_int32(0)
        N.check(L.snisr_bvh_build(V.ctypes.data_as(C.c_void_p), V.shape[0], F.ctypes.data_as(C.c_void_p), F.shape[0],
                                  nodes.ctypes.data_as(C.c_void_p), nb.value, tris.ctypes.data_as(C.c_void_p),
                                  tris.nbytes, C.byref(nn), C.byref(depth)), "bvh_build")
        self.n_nodes, self.depth, self.n_tris = nn.value, depth.value, F.shape[0]
        self.nodes = torch.from_numpy(nodes[: 16 * self.n_nodes].copy()).to(device)
        self.tris = torch.from_numpy(tris).to(device)

    def first_hit(self, origins, dirs, image_width=0, t_out=None, id_out=None):
        r = _rays_desc(origins, dirs, image_width=image_width)
        t = t_out if t_out is not None else torch.empty(r.n, dtype=torch.float32, device=dirs.device)
        ids = id_out if id_out is not None else torch.empty(r.n, dtype=torch.int32, device=dirs.device)
        N.check(N.lib().snisr_mesh_first_hit(_ptr(self.nodes), self.n_nodes, _ptr(self.tris), self.n_tris, self.depth,
                                             C.byref(r), _ptr(t), _ptr(ids), _stream()), "mesh_first_hit")
        return t, ids


class TileFaceMesh:
    """A proxy mesh with per-tile per-block face lists (the reference's BConFaceIdx / BConFaceNum,
    training/scripts/tools.py:364-379), built on the host by snisr_tile_face_lists, for the per-tile
    first / double hits of ray_firstHit_inTile_kernel / ray_doubleHit_inTile_kernel
    (training/src/cuda_func.cu:130-444)."""

    def __init__(self, vertices: np.ndarray, faces: np.ndarray, centers: np.ndarray, tile_size: float,
                 num_block: int = 32, device="cuda", face_idx=None, face_num=None):
        L = N.lib()
        V = np.ascontiguousarray(vertices, np.float32)
        F = np.ascontiguousarray(faces, np.int32)
        ctr = np.ascontiguousarray(centers, np.float32).reshape(-1, 3)
        nb = int(num_block)
        if face_num is None:
            fnum = np.empty((ctr.shape[0] * nb ** 3, 2), np.int32)
            total = C.c_int64(0)
            args = (V.ctypes.data_as(C.c_void_p), V.shape[0], F.ctypes.data_as(C.c_void_p), F.shape[0],
                    ctr.ctypes.data_as(C.c_void_p), ctr.shape[0], float(tile_size), nb,
                    fnum.ctypes.data_as(C.c_void_p))
            N.check(L.snisr_tile_face_lists(*args, None, 0, C.byref(total)), "tile_face_lists")
            fidx = np.empty(max(total.value, 1), np.int32)
            N.check(L.snisr_tile_face_lists(*args, fidx.ctypes.data_as(C.c_void_p), fidx.size, C.byref(total)),
                    "tile_face_lists")
            fidx = fidx[:total.value]
        else:   # lists loaded from files, as the reference's load_preprocess reads them
            fidx = np.ascontiguousarray(face_idx, np.int32).reshape(-1)
            fnum = np.ascontiguousarray(face_num, np.int32).reshape(-1, 2)
        self.face_idx_host, self.face_num_host = fidx, fnum
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)
        self.vertices, self.faces, self.centers = t(V), t(F), t(ctr)
        self.face_idx = t(fidx if fidx.size else np.zeros(1, np.int32))
        self.face_num = t(fnum)
        d = N.TileMeshDesc()
        d.vertices, d.faces, d.n_faces = self.vertices.data_ptr(), self.faces.data_ptr(), F.shape[0]
        d.centers, d.face_idx, d.n_face_idx = self.centers.data_ptr(), self.face_idx.data_ptr(), fidx.size
        d.face_num, d.tile_size, d.num_block = self.face_num.data_ptr(), float(tile_size), nb
        self.desc = d

    def hits(self, rays_o: torch.Tensor, rays_d: torch.Tensor, visited: torch.Tensor, mode: int = 0,
             hits: torch.Tensor | None = None, hit_face: torch.Tensor | None = None):
        """mode 0: (hits f32[n, K], faces i32[n, K]); mode 1: hits f32[n, K, 2].  ``hits`` may be given
        (mode 0 leaves entries without a hit as they were, like the reference)."""
        _require_cuda(rays_o, "rays_o", torch.float32)
        _require_cuda(rays_d, "rays_d", torch.float32)
        _require_cuda(visited, "visited", torch.int32)
        n = rays_d.numel() // 3
        K = visited.shape[-1]
        if rays_o.numel() != 3 * n or visited.numel() != n * K:
            raise ValueError("rays_o / rays_d must be [n, 3] and visited [n, K]")
        dev = rays_d.device
        if hits is None:
            hits = torch.full((n, K, 2) if mode == 1 else (n, K), -1.0, dtype=torch.float32, device=dev)
        _require_cuda(hits, "hits", torch.float32)
        if hits.numel() != n * K * (2 if mode == 1 else 1):
            raise ValueError("hits must hold n*K (mode 0) or n*K*2 (mode 1) values")
        if mode == 0 and hit_face is None:
            hit_face = torch.empty((n, K), dtype=torch.int32, device=dev)
        if hit_face is not None:
            _require_cuda(hit_face, "hit_face", torch.int32)
        N.check(N.lib().snisr_mesh_hit_tiles(C.byref(self.desc), _ptr(rays_o), _ptr(rays_d), n, K, _ptr(visited),
                                             int(mode), _ptr(hits), _ptr(hit_face), _stream()), "mesh_hit_tiles")
        return hits if mode == 1 else (hits, hit_face)


class NativeScene:
    """A scene uploaded by the library itself (snisr_scene_create: device copies owned by the handle,
    FP32 and optionally MFMA weight packing), from the host arrays of a ``Scene``."""

    def __init__(self, sc: Scene, mfma_dtype: int = N.SNISR_F32):
        self._keep = [np.ascontiguousarray(a) for a in (sc.tile_index_map, sc.tile_centers, sc.tile_res,
                                                        sc.tile_brick_base, sc.brick_map, sc.sigma, sc.feats,
                                                        sc.mlp_params)]
        tim, ctr, res, base, bmap, sig, feat, mlp = self._keep
        d = N.SceneDesc()
        d.tile_shape[:] = [int(v) for v in sc.tile_shape]
        d.scene_min[:] = [float(v) for v in sc.scene_min]
        d.tile_size = sc.tile_size
        d.n_tiles, d.n_bricks, d.feat_dim = sc.n_tiles, sc.n_bricks, sc.feat_dim
        d.feat_dtype = N.SNISR_F16 if feat.dtype == np.float16 else N.SNISR_F32
        d.mlp_in, d.mlp_width, d.mlp_layers = sc.mlp_in, sc.mlp_width, sc.mlp_layers
        d.mlp_stride = mlp.shape[1]
        d.brick_map_len = bmap.size
        for name, a in (("tile_index_map", tim), ("tile_centers", ctr), ("tile_res", res),
                        ("tile_brick_base", base), ("brick_map", bmap), ("sigma", sig), ("feats", feat),
                        ("mlp_params", mlp)):
            setattr(d, name, a.ctypes.data)
        self.handle = C.c_void_p()
        N.check(N.lib().snisr_scene_create(C.byref(d), int(mfma_dtype), C.byref(self.handle)), "scene_create")
        self.host, self.mfma_dtype = sc, mfma_dtype

    @property
    def desc(self):
        return N.lib().snisr_scene_get_desc(self.handle).contents

    def close(self):
        if self.handle:
            N.lib().snisr_scene_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class NativeGraph:
    """The library's frame driver (snisr_graph_*): camera rays, optional BVH bound and the render of
    H x W frames captured once in a HIP graph; ``render(cam)`` replays it for a new camera and returns
    the device outputs as tensors that alias the driver's buffers (overwritten by the next replay)."""

    def __init__(self, scene: NativeScene, cam, mlp_dtype: int = N.SNISR_F16, mesh: "DeviceMesh | None" = None,
                 sample_headroom: float = 0.05, records_per_ray: int = 0, sample_capacity: int = 0, **march_kw):
        g = N.GraphDesc()
        g.H, g.W, g.mlp_dtype = cam.H, cam.W, int(mlp_dtype)
        g.records_per_ray, g.sample_capacity = int(records_per_ray), int(sample_capacity)
        if mesh is not None:
            g.bvh_nodes, g.bvh_tris = mesh.nodes.data_ptr(), mesh.tris.data_ptr()
            g.bvh_n_nodes, g.bvh_n_tris, g.bvh_depth = mesh.n_nodes, mesh.n_tris, mesh.depth
        g.sample_headroom = float(sample_headroom)
        sc = scene.host
        g.params.max_tiles = march_kw.get("max_tiles", sc.max_tiles)
        g.params.sample_step = march_kw.get("sample_step", sc.sample_step)
        g.params.early_stop = march_kw.get("early_stop", sc.early_stop)
        g.params.trans_th = march_kw.get("trans_th", sc.trans_th)
        self.scene, self.mesh, self.H, self.W = scene, mesh, cam.H, cam.W
        self.handle = C.c_void_p()
        *cam_args, keep = self._camera(cam)  # noqa: F841 (keep: the arrays behind the pointers)
        N.check(N.lib().snisr_graph_create(scene.handle, C.byref(g), *cam_args, C.byref(self.handle)),
                "graph_create")
        ptrs = [C.c_void_p() for _ in range(5)]
        N.check(N.lib().snisr_graph_outputs(self.handle, *[C.byref(p) for p in ptrs]), "graph_outputs")
        self.ptrs = [p.value for p in ptrs]

    @staticmethod
    def _camera(cam):
        a = [np.ascontiguousarray(v, np.float32) for v in (cam.intrinsic, cam.c2w, cam.origin)]
        return [x.ctypes.data_as(C.c_void_p) for x in a] + [a]

    def render(self, cam=None):
        if cam is None:
            args = [None, None, None]
        else:
            *args, keep = self._camera(cam)  # noqa: F841
        N.check(N.lib().snisr_graph_render(self.handle, *args, _stream()), "graph_render")

    def outputs(self):
        """rgb [H*W,3], T, n_samples, depth (and t_hit when a mesh is used): device tensors aliasing the
        driver's buffers (through __cuda_array_interface__), overwritten by the next replay."""
        N.check(N.lib().snisr_graph_synchronize(self.handle), "graph_synchronize")
        n = self.H * self.W
        out = {}
        for name, p, shape, ts in (("rgb", self.ptrs[0], (n, 3), "<f4"), ("T", self.ptrs[1], (n,), "<f4"),
                                   ("n_samples", self.ptrs[2], (n,), "<i4"), ("depth", self.ptrs[3], (n,), "<f4"),
                                   ("t_hit", self.ptrs[4], (n,), "<f4")):
            if p:
                out[name] = torch.as_tensor(_DeviceArray(p, shape, ts), device="cuda")
        return out

    def stats(self):
        total, over, cpr, cap = C.c_int64(0), C.c_int32(0), C.c_int32(0), C.c_int64(0)
        N.check(N.lib().snisr_graph_stats(self.handle, C.byref(total), C.byref(over), C.byref(cpr), C.byref(cap)),
                "graph_stats")
        return dict(total=total.value, overflow=over.value, records_per_ray=cpr.value, sample_capacity=cap.value)

    def close(self):
        if self.handle:
            N.lib().snisr_graph_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DeviceArray:
    """A device buffer of the library described by __cuda_array_interface__ (no copy)."""

    def __init__(self, ptr: int, shape, typestr: str):
        self.__cuda_array_interface__ = {"shape": tuple(shape), "typestr": typestr, "data": (int(ptr), False),
                                         "version": 2, "strides": None}
