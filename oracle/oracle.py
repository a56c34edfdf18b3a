"""ORACLE — test infrastructure only (imported by tests/, the smoke check and bench.py's
cpu_baseline leg).  ctypes front end of ``liboracle_snisr.so``, the host C++/OpenMP
restatement of the reference renderer's per-ray logic (snisr_oracle.cpp cites the reference
file:line of every function)."""
from __future__ import annotations

import ctypes as C
import ctypes as C_
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_snisr.so")
_lib = None


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "liboracle_snisr.so"], check=True)
    return _LIB_PATH


class _SceneC(C.Structure):
    _fields_ = [("tile_shape", C.c_int * 3), ("scene_min", C.c_float * 3), ("tile_size", C.c_float),
                ("tile_index_map", C.c_void_p), ("n_tiles", C.c_int), ("tile_centers", C.c_void_p),
                ("tile_res", C.c_void_p), ("tile_brick_base", C.c_void_p), ("brick_map", C.c_void_p),
                ("sigma", C.c_void_p), ("feats", C.c_void_p), ("feat_dim", C.c_int), ("feat_half", C.c_int),
                ("mlp_in", C.c_int), ("mlp_width", C.c_int), ("mlp_layers", C.c_int),
                ("mlp_params", C.c_void_p), ("mlp_stride", C.c_longlong)]


class _ParamsC(C.Structure):
    _fields_ = [("max_tiles", C.c_int), ("sample_step", C.c_float), ("early_stop", C.c_float),
                ("trans_th", C.c_float), ("brick_debug", C.c_int)]


def lib():
    """The oracle library; SNISR_ORACLE_LIB names another build of it (the sanitizer build)."""
    global _lib
    if _lib is None:
        path = os.environ.get("SNISR_ORACLE_LIB") or _LIB_PATH
        if path == _LIB_PATH and not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(path)
        _lib.orc_expf_c.restype = C.c_float
        _lib.orc_expf_c.argtypes = [C.c_float]
        _lib.orc_num_threads.restype = C.c_int
        _lib.orc_half_to_float_c.restype = C.c_float
        _lib.orc_half_to_float_c.argtypes = [C.c_uint16]
        _lib.orc_round_half_f64.restype = C.c_double
        _lib.orc_round_half_f64.argtypes = [C.c_double]
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class OracleScene:
    """Holds contiguous copies of a ``snisr.scene.Scene``'s arrays and the C descriptor."""

    def __init__(self, sc):
        self.keep = dict(
            tim=np.ascontiguousarray(sc.tile_index_map, np.int32),
            tc=np.ascontiguousarray(sc.tile_centers, np.float32),
            tr=np.ascontiguousarray(sc.tile_res, np.int32),
            tb=np.ascontiguousarray(sc.tile_brick_base, np.int32),
            bm=np.ascontiguousarray(sc.brick_map, np.int32),
            sg=np.ascontiguousarray(sc.sigma, np.float32),
            ft=np.ascontiguousarray(sc.feats),
            mp=np.ascontiguousarray(sc.mlp_params, np.float32),
        )
        k = self.keep
        c = _SceneC()
        c.tile_shape[:] = [int(v) for v in sc.tile_shape]
        c.scene_min[:] = [float(v) for v in sc.scene_min]
        c.tile_size = sc.tile_size
        c.tile_index_map = _p(k["tim"])
        c.n_tiles = sc.n_tiles
        c.tile_centers = _p(k["tc"])
        c.tile_res = _p(k["tr"])
        c.tile_brick_base = _p(k["tb"])
        c.brick_map = _p(k["bm"])
        c.sigma = _p(k["sg"])
        c.feats = _p(k["ft"])
        c.feat_dim = sc.feat_dim
        c.feat_half = 1 if k["ft"].dtype == np.float16 else 0
        c.mlp_in, c.mlp_width, c.mlp_layers = sc.mlp_in, sc.mlp_width, sc.mlp_layers
        c.mlp_params = _p(k["mp"])
        c.mlp_stride = k["mp"].shape[1]
        self.c = c
        self.sc = sc

    def params(self, brick_debug=0, max_tiles=None, early_stop=None, sample_step=None):
        sc = self.sc
        p = _ParamsC()
        p.max_tiles = sc.max_tiles if max_tiles is None else max_tiles
        p.sample_step = sc.sample_step if sample_step is None else sample_step
        p.early_stop = sc.early_stop if early_stop is None else early_stop
        p.trans_th = sc.trans_th
        p.brick_debug = brick_debug
        return p


def expf(x: float) -> float:
    return lib().orc_expf_c(float(x))


def half_to_float(bits: int) -> float:
    return lib().orc_half_to_float_c(int(bits))


def num_threads() -> int:
    return lib().orc_num_threads()


def compute_rays(H, W, intrinsic, c2w, dtype=np.float32):
    out = np.empty((H * W, 3), dtype)
    fn = lib().orc_compute_rays_f32 if dtype == np.float32 else lib().orc_compute_rays_f64
    fn(C.c_int(H), C.c_int(W), _p(np.ascontiguousarray(intrinsic, np.float32)),
       _p(np.ascontiguousarray(c2w, np.float32)), _p(out))
    return out


def ray_aabb(o, d, c, h):
    out = np.empty(2, np.float32)
    a = [np.ascontiguousarray(v, np.float32) for v in (o, d, c, h)]
    lib().orc_ray_aabb_f32(*[_p(v) for v in a], _p(out))
    return out


def trace_tiles(osc: OracleScene, origin, dirs, K):
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    out = np.empty((dirs.shape[0], K), np.int32)
    lib().orc_trace_tiles_f32(C.byref(osc.c), _p(np.ascontiguousarray(origin, np.float32)), _p(dirs),
                              C.c_longlong(dirs.shape[0]), C.c_int(K), _p(out))
    return out


def render(osc: OracleScene, origins, dirs, t_hit=None, visited=None, ray_index=None, precision="f32",
           with_mlp=True, brick_debug=0, **pkw):
    """Per-ray render.  ``origins`` is f32[3] (shared) or f32[N,3].  Returns a dict of
    rgb f32[M,3], T f32[M], n_samples i32[M], depth f32[M] (+ bricks i32[M,brick_debug])."""
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    N = dirs.shape[0]
    origins = np.ascontiguousarray(origins, np.float32)
    ostride = 0 if origins.size == 3 else 3
    if ray_index is not None:
        ray_index = np.ascontiguousarray(ray_index, np.int64)
        M = ray_index.shape[0]
    else:
        M = N
    p = osc.params(brick_debug=brick_debug, **pkw)
    rgb = np.zeros((M, 3), np.float32)
    T = np.zeros(M, np.float32)
    ns = np.zeros(M, np.int32)
    depth = np.zeros(M, np.float32)
    bricks = np.full((M, brick_debug), -1, np.int32) if brick_debug else None
    if t_hit is not None:
        t_hit = np.ascontiguousarray(t_hit, np.float32)
    if visited is not None:
        visited = np.ascontiguousarray(visited, np.int32)
        assert visited.shape == (N, p.max_tiles)
    fn = lib().orc_render_f32 if precision == "f32" else lib().orc_render_f64
    fn(C.byref(osc.c), C.byref(p), _p(origins), C.c_int(ostride), _p(dirs), _p(t_hit), _p(visited),
       C.c_longlong(M), _p(ray_index), _p(rgb), _p(T), _p(ns), _p(depth), _p(bricks), C.c_int(1 if with_mlp else 0))
    out = dict(rgb=rgb, T=T, n_samples=ns, depth=depth)
    if bricks is not None:
        out["bricks"] = bricks
    return out


def trilinear(osc: OracleScene, tile, f):
    f = np.ascontiguousarray(f, np.float32).reshape(-1, 3)
    M = f.shape[0]
    sig = np.empty(M, np.float32)
    feat = np.empty((M, osc.sc.feat_dim), np.float32)
    lib().orc_trilinear_f32(C.byref(osc.c), C.c_int(tile), _p(f), C.c_longlong(M), _p(sig), _p(feat))
    return sig, feat


def mlp(osc: OracleScene, tile, x):
    x = np.ascontiguousarray(x, np.float64).reshape(-1, osc.sc.mlp_in)
    out = np.empty((x.shape[0], 3), np.float64)
    lib().orc_mlp_f64(C.byref(osc.c), C.c_int(tile), _p(x), C.c_longlong(x.shape[0]), _p(out))
    return out


def view_encoding(d):
    d = np.ascontiguousarray(d, np.float64).reshape(-1, 3)
    out = np.empty((d.shape[0], 15), np.float64)
    lib().orc_view_encoding_f64(_p(d), C.c_longlong(d.shape[0]), _p(out))
    return out


def normalize(d):
    d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
    out = np.empty_like(d)
    lib().orc_normalize_f32(_p(d), C.c_longlong(d.shape[0]), _p(out))
    return out


def mesh_hit(vertices, faces, origins, dirs):
    V = np.ascontiguousarray(vertices, np.float32)
    Fc = np.ascontiguousarray(faces, np.int32)
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    N = dirs.shape[0]
    o = np.ascontiguousarray(np.broadcast_to(np.asarray(origins, np.float32).reshape(-1, 3), (N, 3)))
    t = np.empty(N, np.float32)
    ids = np.empty(N, np.int32)
    lib().orc_mesh_hit_f32(_p(V), _p(Fc), C.c_int(Fc.shape[0]), _p(o), _p(dirs), C.c_longlong(N), _p(t), _p(ids))
    return t, ids


def inverse_z_sampling(origin, dirs, inv_near, inv_bound, offset, spr):
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    n = dirs.shape[0]
    pts = np.empty((n * spr, 3), np.float32)
    st = np.empty(n * spr, np.float32)
    lib().orc_inverse_z_sampling_f32(_p(np.ascontiguousarray(origin, np.float32)), _p(dirs),
                                     _p(np.ascontiguousarray(inv_near, np.float32).ravel()),
                                     _p(np.ascontiguousarray(inv_bound, np.float32).ravel()),
                                     C.c_longlong(n), C.c_int(offset), C.c_int(spr), _p(pts), _p(st))
    return pts, st


def integrate_points(sample_results, steps, early, T, color):
    """In-place on T (f32[n]) and color (f32[n,3])."""
    res = np.ascontiguousarray(sample_results, np.float32)
    st = np.ascontiguousarray(steps, np.float32)
    n = T.size
    lib().orc_integrate_points_f32(_p(res), _p(st), C.c_longlong(n), C.c_float(early), _p(T), _p(color))


def assign_pixels(net, pixel_starts, block_starts):
    net = np.ascontiguousarray(net, np.int16)
    ps = np.ascontiguousarray(pixel_starts, np.int32)
    bs = np.ascontiguousarray(block_starts, np.int32)
    nb = int(bs[-1])
    new_net = np.empty(nb, np.int16)
    new_starts = np.empty(nb, np.int32)
    lib().orc_assign_pixels(_p(net), _p(ps), _p(bs), C.c_int(net.shape[0]), _p(new_net), _p(new_starts))
    return new_net, new_starts


REF_MLP_PARAMS = 22556  # halves per group: per layer bias[out] then W[in][out]; 63-64-64-64-64-64-28


def inference_mlp(spr, net_idx, pixel_starts, query_pixel_indices, transparency, group_centers, samples, rays_d,
                  params_f16, inverse_far, trans_stop, results, quant16=False):
    """FP64 restatement of inference_mlp_kernel; writes results f32[H*W*spr, 4] in place.
    rays_d is f32[H, W, 3] (any strides of a numpy view are honoured)."""
    net = np.ascontiguousarray(net_idx, np.int16)
    ps = np.ascontiguousarray(pixel_starts, np.int32)
    q = np.ascontiguousarray(query_pixel_indices, np.int32)
    T = np.ascontiguousarray(transparency, np.float32).ravel()
    ctr = np.ascontiguousarray(group_centers, np.float32)
    smp = np.ascontiguousarray(samples, np.float32)
    rd = np.asarray(rays_d, np.float32)
    H, W = rd.shape[0], rd.shape[1]
    if rd.strides[2] != 4:
        rd = np.ascontiguousarray(rd)
    prm = np.ascontiguousarray(params_f16, np.float16).view(np.uint16)
    n_groups = prm.size // REF_MLP_PARAMS
    assert results.dtype == np.float32 and results.flags.c_contiguous
    lib().orc_inference_mlp_f64(C.c_int(spr), C.c_longlong(net.size), C.c_float(inverse_far), C.c_float(trans_stop),
                                _p(net), _p(T), _p(q), _p(ps), _p(ctr), C.c_int(n_groups), _p(smp),
                                C.c_void_p(rd.ctypes.data), C.c_longlong(rd.strides[0] // 4),
                                C.c_longlong(rd.strides[1] // 4), C.c_int(W), C.c_longlong(H * W), _p(prm),
                                C.c_int(1 if quant16 else 0), _p(results))


def sh2_basis(dirs):
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    out = np.empty((d.shape[0], 9), np.float64)
    lib().orc_sh2_basis_f64(_p(d), C.c_longlong(d.shape[0]), _p(out))
    return out


def round_half(v: float) -> float:
    return lib().orc_round_half_f64(C.c_double(v))


def train_grad(osc: OracleScene, origins, dirs, g, t_hit=None, precision="f64", **pkw):
    """Colours C[N,3] and the gradients of L = sum_i <g_i, C_i> with respect to sigma, features
    and the MLP parameters (the oracle's adjoint of the training render)."""
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(-1, 3)
    N = dirs.shape[0]
    origins = np.ascontiguousarray(origins, np.float32)
    ostride = 0 if origins.size == 3 else 3
    dt = np.float64 if precision == "f64" else np.float32
    sc = osc.sc
    g = np.ascontiguousarray(g, dt).reshape(N, 3)
    Cc = np.zeros((N, 3), dt)
    dsig = np.zeros(sc.sigma.shape, dt)
    dfeat = np.zeros(sc.feats.shape, dt)
    dmlp = np.zeros(sc.mlp_params.shape, dt)
    if t_hit is not None:
        t_hit = np.ascontiguousarray(t_hit, np.float32)
    p = osc.params(**pkw)
    fn = lib().orc_train_grad_f64 if precision == "f64" else lib().orc_train_grad_f32
    fn(C_.byref(osc.c), C_.byref(p), _p(origins),
       C_.c_int(ostride), _p(dirs), _p(t_hit), C_.c_longlong(N), _p(g), _p(Cc),
       _p(dsig), _p(dfeat), _p(dmlp))
    return dict(rgb=Cc, dsigma=dsig, dfeats=dfeat, dmlp=dmlp)


def render_octree(osc, origin, rays_d, visited, trans_th=0.8, early_stop=0.01, soft_hit=False,
                  inverse_near=None, net_idx=None):
    """FP32 restatement of rendering_diffuse_octree_fp16 on a ``snisr.scene.OctreeScene``.
    Returns frame f32[n,3], inverse_near f32[n], net_idx i16[n], n_samples i32[n]; inverse_near and
    net_idx start from the given arrays (default 1e-8 and -1, the reference's init)."""
    d = np.ascontiguousarray(rays_d, np.float32).reshape(-1, 3)
    n = d.shape[0]
    vis = np.ascontiguousarray(visited, np.int32).reshape(n, -1)
    K = vis.shape[1]
    frame = np.empty((n, 3), np.float32)
    inv = np.full(n, 1e-8, np.float32) if inverse_near is None else np.array(inverse_near, np.float32).ravel()
    net = np.full(n, -1, np.int16) if net_idx is None else np.array(net_idx, np.int16).ravel()
    ns = np.empty(n, np.int32)
    arrs = [np.ascontiguousarray(osc.group_map, np.int32), np.ascontiguousarray(osc.data_voxels, np.float16),
            np.ascontiguousarray(osc.block_index_map, np.int32), np.ascontiguousarray(osc.nodes_index_map, np.int16),
            np.ascontiguousarray(osc.nodes_sample_flag, np.uint8), np.ascontiguousarray(osc.voxels_start, np.int64),
            np.ascontiguousarray(osc.centers, np.float32)]
    lib().orc_render_octree_f32(*[_p(a) for a in arrs], C.c_float(osc.tile_size), C.c_float(osc.voxel_size),
                                C.c_float(osc.sample_step), C.c_int(osc.num_voxel),
                                _p(np.ascontiguousarray(origin, np.float32)), _p(d), _p(vis), C.c_longlong(n),
                                C.c_int(K), C.c_float(trans_th), C.c_float(early_stop), C.c_int(1 if soft_hit else 0),
                                _p(frame), _p(inv), _p(net), _p(ns))
    return frame, inv, net, ns


# ---------------------------------------------------------------------------- training samplers
def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def sample_sparse(rays_o, rays_d, node_flags, center, tile_size, num_voxel, z_vals, dists):
    """In place on z_vals / dists f32[n, N] (rows of rays that miss or see no flagged voxel are left)."""
    o, d = _f32(rays_o), _f32(rays_d)
    fl = np.ascontiguousarray(node_flags, np.int16)
    lib().orc_sample_sparse(_p(o), _p(d), C.c_longlong(o.shape[0]), _p(fl), C.c_float(center[0]),
                            C.c_float(center[1]), C.c_float(center[2]), C.c_float(tile_size), C.c_int(num_voxel),
                            C.c_int(z_vals.shape[1]), _p(z_vals), _p(dists))


def sample_uniform(rays_o, rays_d, center, tile_size, z_vals):
    o, d = _f32(rays_o), _f32(rays_d)
    lib().orc_sample_uniform(_p(o), _p(d), C.c_longlong(o.shape[0]), C.c_float(center[0]), C.c_float(center[1]),
                             C.c_float(center[2]), C.c_float(tile_size), C.c_int(z_vals.shape[1]), _p(z_vals))


def sample_bg(rays_o, rays_d, bg_zval, tile_center, tile_size, sample_range, z_vals):
    o, d, bg, tc = _f32(rays_o), _f32(rays_d), _f32(bg_zval), _f32(tile_center)
    lib().orc_sample_bg(_p(o), _p(d), _p(bg), C.c_longlong(o.shape[0]), _p(tc), C.c_float(tile_size),
                        C.c_float(sample_range), C.c_int(z_vals.shape[1]), _p(z_vals))


def sample_reflection(rays_o, rays_d, node_flags, voxels, center, tile_size, far, num_voxel, T_th, z_vals):
    o, d = _f32(rays_o), _f32(rays_d)
    fl = np.ascontiguousarray(node_flags, np.int16)
    vx = _f32(voxels)
    lib().orc_sample_reflection(_p(o), _p(d), C.c_longlong(o.shape[0]), _p(fl), _p(vx), C.c_float(center[0]),
                                C.c_float(center[1]), C.c_float(center[2]), C.c_float(tile_size), C.c_float(far),
                                C.c_int(num_voxel), C.c_float(T_th), C.c_int(z_vals.shape[1]), _p(z_vals))


def group_boundary(net_idx, ksize, boundary=None):
    """bool[H, W] (in place on `boundary` when given)."""
    net = np.ascontiguousarray(net_idx, np.int16)
    H, W = net.shape
    out = np.zeros((H, W), np.uint8) if boundary is None else boundary
    lib().orc_group_boundary(_p(net), C.c_int(H), C.c_int(W), C.c_int(ksize), _p(out))
    return out


def _region_blocks(net, max_px):
    """Stable sort of the pixels by group and <= max_px-pixel blocks per group (sizes differing by at
    most one, larger first), the reference's regrouping (rendering/src/__init__.py:272-291)."""
    order = np.argsort(net, kind="stable").astype(np.int32)
    ks = net[order]
    heads = np.flatnonzero(np.concatenate([[True], ks[1:] != ks[:-1]])) if net.size else np.zeros(0, np.int64)
    ends = np.concatenate([heads[1:], [net.size]])
    bnet, bps = [], []
    for a, b in zip(heads, ends):
        nb = -(-(b - a) // max_px)
        base, add = divmod(b - a, nb)
        p = a
        for k in range(nb):
            bnet.append(ks[a])
            bps.append(p)
            p += base + (1 if k < add else 0)
    bps.append(net.size)
    return order, np.asarray(bnet, np.int16), np.asarray(bps, np.int32)


def frender_subset(osc, grid: OracleScene, params_f16, centers, origin, dirs, H, W, far, hi_idx, K=30, scale=2,
                   trans_th=0.8, early_stop_D=0.01, early_stop_R=0.01, max_px=64, quant16=False):
    """The FRender.render frame (rendering/src/__init__.py:247-334) at a subset of its full-resolution
    pixels ``hi_idx`` (flat y*W + x), without rendering the whole frame: the octree diffuse pass of
    those pixels, the reflection pass (4 x 16 inverse-depth samples through the group MLP in FP64,
    FP16-rounded activations only with ``quant16``) of the downscaled pixels their bilinear x2
    upsampling reads (align_corners=True), and diffuse + reflection clamped to [0, 1].  ``osc`` is an
    octree scene (snisr.scene.OctreeScene fields), ``grid`` an OracleScene with its tile grid for the
    tile tracer, ``dirs`` the frame's f32[H*W, 3] camera directions.
    Returns dict(frame f64[M,3], diffuse f32[M,3], spec_lo f32[L,3], lo_idx (flat DH*DW indices))."""
    DH, DW = H // scale, W // scale
    dirs = np.ascontiguousarray(dirs, np.float32).reshape(H * W, 3)
    hi_idx = np.asarray(hi_idx, np.int64)
    hy, hx = hi_idx // W, hi_idx % W
    # torch's align_corners=True source coordinate: (in - 1) / (out - 1) * dst, in float32
    sy = (np.float32(DH - 1) / np.float32(H - 1)) * hy.astype(np.float32)
    sx = (np.float32(DW - 1) / np.float32(W - 1)) * hx.astype(np.float32)
    y0 = np.floor(sy).astype(np.int64)
    x0 = np.floor(sx).astype(np.int64)
    y1, x1 = np.minimum(y0 + 1, DH - 1), np.minimum(x0 + 1, DW - 1)
    ly, lx = (sy - y0).astype(np.float64), (sx - x0).astype(np.float64)
    corners = [y0 * DW + x0, y0 * DW + x1, y1 * DW + x0, y1 * DW + x1]
    lo_idx = np.unique(np.concatenate(corners))
    n_lo = lo_idx.size
    lo_full = (lo_idx // DW) * scale * W + (lo_idx % DW) * scale      # ray of the downscaled pixel
    d_lo = np.ascontiguousarray(dirs[lo_full])
    org = np.ascontiguousarray(origin, np.float32)
    # diffuse pass of the downscaled pixels (first-hit inverse depth and group) and of the subset
    _, inv_near, net, _ = render_octree(osc, org, d_lo, trace_tiles(grid, org, d_lo, K), trans_th, early_stop_D)
    d_hi = np.ascontiguousarray(dirs[hi_idx])
    diff_hi, _, _, _ = render_octree(osc, org, d_hi, trace_tiles(grid, org, d_hi, K), trans_th, early_stop_D)
    # reflection pass on the downscaled pixels as a 1 x n_lo image (FRender._reflection_and_compose)
    sample_far = (np.float32(1.0) / inv_near + np.float32(far)).astype(np.float32)
    inv_bound = ((np.float32(1.0) / sample_far - inv_near) / np.float32(63)).astype(np.float32)
    order, bnet, bps = _region_blocks(net, max_px)
    T = np.ones(n_lo, np.float32)
    col = np.zeros((n_lo, 3), np.float32)
    for step in range(0, 64, 16):
        pts, st = inverse_z_sampling(org, d_lo, inv_near, inv_bound, step, 16)
        res = np.zeros((n_lo * 16, 4), np.float32)
        inference_mlp(16, bnet, bps, order, T, centers, pts, d_lo.reshape(1, n_lo, 3), params_f16, 1.0 / far,
                      early_stop_R, res, quant16=quant16)
        integrate_points(res, st, np.float32(early_stop_R), T, col)
    pos = {int(v): k for k, v in enumerate(lo_idx)}
    s = [col[np.fromiter((pos[int(v)] for v in c), np.int64, c.size)].astype(np.float64) for c in corners]
    lxe, lye = lx[:, None], ly[:, None]
    up = (1 - lye) * ((1 - lxe) * s[0] + lxe * s[1]) + lye * ((1 - lxe) * s[2] + lxe * s[3])
    frame = np.clip(diff_hi.astype(np.float64) + up, 0.0, 1.0)
    return dict(frame=frame, diffuse=diff_hi, spec_lo=col, lo_idx=lo_idx, net_lo=net, inv_near_lo=inv_near)


def dense_overlap(voxels, center, tile_size, voxel_size, points):
    """OVERLAP interpolation (training/src/include/interpolation.h:120-134) of one dense grid
    f32[nv, nv, nv, 4] at world points f32[M, 3] -> f32[M, 4]."""
    vox = np.ascontiguousarray(voxels, np.float32)
    M = vox.shape[0]
    p = np.ascontiguousarray(points, np.float32).reshape(-1, 3)
    out = np.empty((p.shape[0], 4), np.float32)
    lib().orc_dense_overlap_f32(_p(vox), C.c_int(M), _p(np.ascontiguousarray(center, np.float32)),
                                C.c_float(tile_size), C.c_float(voxel_size), _p(p), C.c_longlong(p.shape[0]), _p(out))
    return out


def render_dense(voxels, nodes, sparse_to_group, centers, tile_size, voxel_size, sample_step, rays_o, rays_d,
                 visited):
    """FP32 restatement of render_diffuse_kernel (training/src/cuda_func.cu:1687-1802): returns
    diffuse f32[n, 3], depth f32[n], valid i32[n]."""
    vox = np.ascontiguousarray(voxels, np.float32)
    M = vox.shape[1]
    nod = np.ascontiguousarray(nodes, np.int16)
    o = np.ascontiguousarray(rays_o, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(rays_d, np.float32).reshape(-1, 3)
    n = d.shape[0]
    vis = np.ascontiguousarray(visited, np.int32).reshape(n, -1)
    dif = np.empty((n, 3), np.float32)
    dep = np.empty(n, np.float32)
    val = np.empty(n, np.int32)
    lib().orc_render_dense_f32(_p(vox), _p(nod), _p(np.ascontiguousarray(sparse_to_group, np.int32)),
                               _p(np.ascontiguousarray(centers, np.float32)), C.c_float(tile_size),
                               C.c_float(voxel_size), C.c_float(sample_step), C.c_int(M), _p(o), _p(d), _p(vis),
                               C.c_longlong(n), C.c_int(vis.shape[1]), _p(dif), _p(dep), _p(val))
    return dif, dep, val


def mesh_hit_tiles(vertices, faces, centers, tile_size, num_block, rays_o, rays_d, visited, mode=0):
    """Restatement of ray_firstHit_inTile_kernel (mode 0) / ray_doubleHit_inTile_kernel (mode 1),
    training/src/cuda_func.cu:130-444, testing every face near each walked block.  Returns
    (hits f32[n, K] with NaN where unwritten, faces i32[n, K]) for mode 0, hits f32[n, K, 2] for mode 1."""
    V = np.ascontiguousarray(vertices, np.float32)
    Fc = np.ascontiguousarray(faces, np.int32)
    o = np.ascontiguousarray(rays_o, np.float32).reshape(-1, 3)
    d = np.ascontiguousarray(rays_d, np.float32).reshape(-1, 3)
    n = d.shape[0]
    vis = np.ascontiguousarray(visited, np.int32).reshape(n, -1)
    K = vis.shape[1]
    hits = np.full((n, K, 2) if mode == 1 else (n, K), np.nan, np.float32)
    face = np.full((n, K), -7, np.int32)
    lib().orc_mesh_hit_tiles_f32(C.c_int(mode), _p(V), _p(Fc), C.c_int(Fc.shape[0]),
                                 _p(np.ascontiguousarray(centers, np.float32)), C.c_float(tile_size),
                                 C.c_int(num_block), _p(o), _p(d), C.c_longlong(n), C.c_int(K), _p(vis), _p(hits),
                                 _p(face))
    return hits if mode == 1 else (hits, face)


def volume_render(rgb, sigma, dists, init_T=None, g_out=None, g_trans=None, g_alpha=None):
    """FP64 restatement of utils.volume_rendering (training/scripts/utils.py:823-837) and its adjoint.
    rgb [B, N, 3], sigma and dists [B, N]; returns dict(out, trans, alpha) and, with g_out, the
    gradients g_rgb, g_sigma (and g_init when init_T is given)."""
    rgb = np.ascontiguousarray(rgb, np.float64)
    B, n = rgb.shape[0], rgb.shape[1]
    sig = np.ascontiguousarray(sigma, np.float64).reshape(B, n)
    dst = np.ascontiguousarray(dists, np.float64).reshape(B, n)
    it = None if init_T is None else np.ascontiguousarray(init_T, np.float64).reshape(B)
    out, trans, alpha = np.empty((B, 3)), np.empty((B, n)), np.empty((B, n))
    go = None if g_out is None else np.ascontiguousarray(g_out, np.float64).reshape(B, 3)
    gt = None if g_trans is None else np.ascontiguousarray(g_trans, np.float64).reshape(B, n)
    ga = None if g_alpha is None else np.ascontiguousarray(g_alpha, np.float64).reshape(B, n)
    g_rgb, g_sig = np.zeros((B, n, 3)), np.zeros((B, n))
    g_init = np.zeros(B) if it is not None else None
    lib().orc_volume_render_f64(_p(rgb), _p(sig), _p(dst), _p(it) if it is not None else None, C.c_longlong(B),
                                C.c_int(n), _p(go) if go is not None else None, _p(gt) if gt is not None else None,
                                _p(ga) if ga is not None else None, _p(out), _p(trans), _p(alpha), _p(g_rgb),
                                _p(g_sig), _p(g_init) if g_init is not None else None)
    res = dict(out=out, trans=trans, alpha=alpha)
    if go is not None:
        res.update(g_rgb=g_rgb, g_sigma=g_sig)
        if g_init is not None:
            res["g_init"] = g_init
    return res


def sh_sigmoid(coeffs, dirs, deg, g_spec=None):
    """FP64 restatement of sigmoid(eval_sh(deg, ...)) (tile.py:212-219, sh.py:55-111) for coefficient
    rows [n, K] ([basis][rgb] first) at the normalised directions of the rows' rays (dirs [n_rays, 3],
    n / n_rays rows per ray); with g_spec also dL/dcoeffs."""
    q = np.ascontiguousarray(coeffs, np.float64)
    K = q.shape[-1]
    q2 = q.reshape(-1, K)
    n = q2.shape[0]
    d = np.ascontiguousarray(dirs, np.float64).reshape(-1, 3)
    spr = n // d.shape[0]
    spec = np.empty((n, 3))
    gq = np.zeros((n, K))
    gs = None if g_spec is None else np.ascontiguousarray(g_spec, np.float64).reshape(n, 3)
    lib().orc_sh_sigmoid_f64(_p(q2), C.c_int(K), _p(d), C.c_longlong(n), C.c_int(spr), C.c_int(deg),
                             _p(gs) if gs is not None else None, _p(spec), _p(gq))
    return (spec, gq) if gs is not None else spec


def pemlp(params, points, depth, n_out, slope, g_out=None):
    """FP64 networks.MLP (L = 10, width 64) forward; with g_out also the parameter gradient."""
    p = np.ascontiguousarray(params, np.float64)
    x = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    n = x.shape[0]
    out = np.empty((n, n_out))
    go = None if g_out is None else np.ascontiguousarray(g_out, np.float64).reshape(n, n_out)
    gp = np.zeros_like(p) if go is not None else None
    lib().orc_pemlp_f64(_p(p), _p(x), C.c_longlong(n), C.c_int(depth), C.c_int(n_out), C.c_double(slope),
                        _p(go) if go is not None else None, _p(out), _p(gp) if gp is not None else None)
    return (out, gp) if go is not None else out


def flag_cells(flags):
    """Row-major indices of the cells whose node flag is not -1 (Tile.compute_voxels' torch.where)."""
    return np.flatnonzero(np.asarray(flags).reshape(-1) != -1).astype(np.int32)


def bake_points(cells, D, min_corner, voxel_size, center, scale):
    """FP32 MLP inputs of the cells: ((index + 0.5) voxel_size + min_corner - center) / scale per axis,
    in the reference's operation order (tile.py:91-95, tileGroup.py:74)."""
    c = np.asarray(cells, np.int64)
    idx = np.stack([c % D, (c // D) % D, c // (D * D)], 1).astype(np.float32)
    f = np.float32
    p = (idx + f(0.5)) * f(voxel_size) + np.asarray(min_corner, np.float32)
    return ((p - np.asarray(center, np.float32)) / f(scale)).astype(np.float32)


def diffuse_head(raw, g_out=None):
    """FP64 (sigmoid(r0..r2), relu(r3)) of MLP rows; with g_out also dL/draw."""
    r = np.asarray(raw, np.float64).reshape(-1, 4)
    s = 1.0 / (1.0 + np.exp(-r[:, :3]))
    out = np.concatenate([s, np.maximum(r[:, 3:], 0.0)], 1)
    if g_out is None:
        return out
    g = np.asarray(g_out, np.float64).reshape(-1, 4)
    return out, np.concatenate([g[:, :3] * s * (1 - s), np.where(r[:, 3:] > 0, g[:, 3:], 0.0)], 1)


def voxel_query(vox, points, min_corner, voxel_size, g_rgb=None, g_sigma=None):
    """FP64 Tile.query_voxel of a channel-last grid vox[D][D][D][4] at world points; returns
    dict(rgb, sigma[, g_vox])."""
    v = np.ascontiguousarray(vox, np.float64)
    D = v.shape[0]
    x = np.ascontiguousarray(points, np.float64).reshape(-1, 3)
    n = x.shape[0]
    mc = np.ascontiguousarray(min_corner, np.float64)
    rgb, sig = np.empty((n, 3)), np.empty(n)
    want = g_rgb is not None or g_sigma is not None
    gr = None if g_rgb is None else np.ascontiguousarray(g_rgb, np.float64).reshape(n, 3)
    gs = None if g_sigma is None else np.ascontiguousarray(g_sigma, np.float64).reshape(n)
    gv = np.empty_like(v) if want else None
    lib().orc_voxel_query_f64(_p(v), C.c_int(D), _p(x), C.c_longlong(n), _p(mc), C.c_double(voxel_size),
                              _p(gr) if gr is not None else None, _p(gs) if gs is not None else None, _p(rgb), _p(sig),
                              _p(gv) if gv is not None else None)
    r = dict(rgb=rgb, sigma=sig)
    if want:
        r["g_vox"] = gv
    return r
