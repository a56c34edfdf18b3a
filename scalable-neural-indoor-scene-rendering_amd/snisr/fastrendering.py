"""Operator-for-operator stand-in for the reference's ``FASTRENDERING`` extension module.

Same names, argument order and in-place semantics as the pybind11 module of
rendering/src/binding.cpp:13-27 (signatures in rendering/src/include/*.h), so a driver written
like ``FRender.render`` (rendering/src/__init__.py:247-341) can import this module in its place.
Every op runs in the HIP kernels of libsnisr_hip.so through the C ABI (include/snisr.h); there is
no CPU path.

Differences from the reference, all deliberate:

* ops launch on torch's current stream (the reference uses the legacy default stream);
* ``num_thread`` is accepted for signature compatibility and ignored: launch shapes are chosen by
  the library;
* outputs must be contiguous tensors of the stated dtype on the GPU.  The reference writes a
  non-contiguous output through ``.contiguous()`` into a temporary and silently loses the result;
  here that raises ``ValueError``;
* failures raise ``RuntimeError`` carrying the library's message, as ``AT_CUDA_CHECK`` does.
"""
from __future__ import annotations

import ctypes as C
import weakref

import numpy as np
import torch

from . import _native as N
from .device import _ptr, _require_cuda, _stream

REF_MLP_PARAMS = 22556  # halves per group of the reflection MLP (rendering/src/rayreflection_render.cu:138)


def _host_floats(t, n: int, name: str) -> list[float]:
    v = t.detach().reshape(-1).to("cpu", torch.float32).tolist() if isinstance(t, torch.Tensor) else list(t)
    if len(v) != n:
        raise ValueError(f"{name} must hold {n} values")
    return [float(x) for x in v]


def _host_ints(t, n: int, name: str) -> list[int]:
    v = t.detach().reshape(-1).to("cpu", torch.int64).tolist() if isinstance(t, torch.Tensor) else list(t)
    if len(v) != n:
        raise ValueError(f"{name} must hold {n} values")
    return [int(x) for x in v]


def compute_rays(intrinsic: torch.Tensor, c2w_rotation: torch.Tensor, num_thread: int, rays_d: torch.Tensor) -> None:
    """rendering/src/compute_rays.cu:46-68: rays_d[y, x] = R [(x - cx) / f, (y - cy) / f, 1] with
    intrinsic = (cx, cy, 1/f); H, W from rays_d's shape; directions are not normalised."""
    _require_cuda(rays_d, "rays_d", torch.float32)
    if rays_d.dim() != 3 or rays_d.shape[2] != 3:
        raise ValueError("rays_d must be [H, W, 3]")
    intr = intrinsic.to(rays_d.device, torch.float32).contiguous()
    rot = c2w_rotation.to(rays_d.device, torch.float32).contiguous()
    if intr.numel() != 3 or rot.numel() != 9:
        raise ValueError("intrinsic must hold 3 values and c2w_rotation 9")
    H, W = rays_d.shape[0], rays_d.shape[1]
    N.check(N.lib().snisr_compute_rays(_ptr(intr), _ptr(rot), H, W, _ptr(rays_d), _stream()), "compute_rays")


def tracing_tiles(rays_d: torch.Tensor, IndexMap: torch.Tensor, origin: torch.Tensor, tile_shape, scene_min_corner,
                  scene_size, tile_size: float, num_thread: int, visitedTiles: torch.Tensor) -> None:
    """rendering/src/raytile_tracing.cu:63-95: the occupied tiles each ray crosses, in order,
    -1-padded to visitedTiles.shape[-1] entries.  scene_size must equal tile_shape * tile_size
    (as the reference's callers build it); the library derives it from those two."""
    _require_cuda(visitedTiles, "visitedTiles", torch.int32)
    shape = _host_ints(tile_shape, 3, "tile_shape")
    smin = _host_floats(scene_min_corner, 3, "scene_min_corner")
    ssz = np.asarray(_host_floats(scene_size, 3, "scene_size"), np.float32)
    want = np.asarray(shape, np.float32) * np.float32(tile_size)
    if not np.array_equal(ssz, want):
        raise ValueError(f"scene_size {ssz.tolist()} differs from tile_shape * tile_size {want.tolist()}")
    imap = IndexMap.to(visitedTiles.device, torch.int32).contiguous()
    if imap.numel() != shape[0] * shape[1] * shape[2]:
        raise ValueError("IndexMap must hold prod(tile_shape) entries")
    rd = rays_d.to(visitedTiles.device, torch.float32).contiguous()
    org = origin.to(visitedTiles.device, torch.float32).contiguous()
    n = rd.numel() // 3
    K = visitedTiles.shape[-1]
    if visitedTiles.numel() != n * K:
        raise ValueError("visitedTiles must be [H, W, K] with H*W rays")
    d = N.SceneDesc()
    d.tile_shape[:] = shape
    d.scene_min[:] = smin
    d.tile_size = float(tile_size)
    d.tile_index_map = imap.data_ptr()
    N.check(N.lib().snisr_trace_tiles(C.byref(d), _ptr(org), _ptr(rd), n, K, _ptr(visitedTiles), _stream()),
            "tracing_tiles")


def inverse_z_sampling(origin: torch.Tensor, rays_d: torch.Tensor, inverse_near: torch.Tensor,
                       inverse_bound: torch.Tensor, num_thread: int, offset: int, sample_per_ray: int,
                       sample_points: torch.Tensor, integrate_steps: torch.Tensor) -> None:
    """rendering/src/sampling.cu:48-81: samples offset .. offset+sample_per_ray-1 of 64 uniform in
    inverse depth.  rays_d may be a strided view such as ``rays_d[::2, ::2]`` (read in place)."""
    _require_cuda(sample_points, "sample_points", torch.float32)
    _require_cuda(integrate_steps, "integrate_steps", torch.float32)
    H, W = inverse_near.shape[0], inverse_near.shape[1]
    if rays_d.dim() != 3 or tuple(rays_d.shape) != (H, W, 3) or rays_d.dtype != torch.float32 or not rays_d.is_cuda:
        raise ValueError("rays_d must be a float32 [H, W, 3] GPU tensor matching inverse_near")
    if rays_d.stride(2) != 1:
        rays_d = rays_d.contiguous()
    inn = inverse_near.to(torch.float32).contiguous()
    inb = inverse_bound.to(torch.float32).contiguous()
    if inb.numel() != H * W:
        raise ValueError("inverse_bound must hold H*W values")
    if sample_points.numel() != H * W * sample_per_ray * 3 or integrate_steps.numel() != H * W * sample_per_ray:
        raise ValueError("sample_points / integrate_steps sizes do not match H*W*sample_per_ray")
    org = origin.to(rays_d.device, torch.float32).contiguous()
    N.check(N.lib().snisr_inverse_z_sampling(_ptr(org), _ptr(rays_d), rays_d.stride(0), rays_d.stride(1), H, W,
                                             _ptr(inn), _ptr(inb), int(offset), int(sample_per_ray),
                                             _ptr(sample_points), _ptr(integrate_steps), _stream()),
            "inverse_z_sampling")


def integrate_points(sample_results: torch.Tensor, intergrate_steps: torch.Tensor, num_thread: int,
                     early_step: float, transparency: torch.Tensor, colors_buffer: torch.Tensor) -> None:
    """rendering/src/volume_rendering.cu:60-87: composites 16 samples per pixel into
    colors_buffer [H, W, 3] and transparency (H*W values), in place."""
    _require_cuda(transparency, "transparency", torch.float32)
    _require_cuda(colors_buffer, "colors_buffer", torch.float32)
    npix = colors_buffer.shape[0] * colors_buffer.shape[1]
    res = sample_results.to(torch.float32).contiguous()
    st = intergrate_steps.to(torch.float32).contiguous()
    if res.numel() != npix * 64 or st.numel() != npix * 16 or transparency.numel() != npix:
        raise ValueError("integrate_points expects 16 samples per pixel")
    N.check(N.lib().snisr_integrate_points(_ptr(res), _ptr(st), npix, float(early_step), _ptr(transparency),
                                           _ptr(colors_buffer), _stream()), "integrate_points")


_sort_ws: dict = {}


def sort_by_key(keys_tensor: torch.Tensor, values_tensor: torch.Tensor, starts_tensor: torch.Tensor) -> int:
    """rendering/src/reorder.cu:11-23: stable sort of int16 keys carrying int32 values, then the
    run heads; returns the number of distinct keys m (keys[:m], starts[:m] hold the runs)."""
    _require_cuda(keys_tensor, "keys", torch.int16)
    _require_cuda(values_tensor, "values", torch.int32)
    _require_cuda(starts_tensor, "starts", torch.int32)
    n = keys_tensor.numel()
    if values_tensor.numel() != n or starts_tensor.numel() != n:
        raise ValueError("keys, values and starts must have the same length")
    need = N.lib().snisr_sort_by_key_workspace_size(n)
    if need < 0:
        raise RuntimeError("sort_by_key: workspace query failed")
    dev = keys_tensor.device
    ws = _sort_ws.get(dev)
    if ws is None or ws.numel() < need:
        ws = torch.empty(max(need, 1 << 16), dtype=torch.uint8, device=dev)
        _sort_ws[dev] = ws
    m = C.c_int32(0)
    N.check(N.lib().snisr_sort_by_key(_ptr(keys_tensor), _ptr(values_tensor), _ptr(starts_tensor), n, _ptr(ws),
                                      ws.numel(), C.byref(m), _stream()), "sort_by_key")
    return m.value


def compute_group_boundary(netIdx: torch.Tensor, boundary: torch.Tensor, ksize: int, num_thread: int) -> None:
    """rendering/src/compute_group_boundary.cu:53-75: boundary (bool [H, W, 1] or [H, W]) is set where
    the ksize x ksize window around a pixel holds another group id of netIdx (int16 [H, W])."""
    _require_cuda(netIdx, "netIdx", torch.int16)
    _require_cuda(boundary, "boundary", torch.bool)
    if netIdx.dim() != 2 or boundary.numel() != netIdx.numel():
        raise ValueError("netIdx must be [H, W] and boundary hold H*W flags")
    H, W = netIdx.shape
    N.check(N.lib().snisr_compute_group_boundary(_ptr(netIdx), H, W, int(ksize), _ptr(boundary), _stream()),
            "compute_group_boundary")


def assign_pixels(netIdxs: torch.Tensor, pixel_starts: torch.Tensor, block_starts: torch.Tensor, num_thread: int,
                  new_netIdxs: torch.Tensor, new_pixel_starts: torch.Tensor) -> None:
    """rendering/src/reorder.cu:69-94: splits each group's pixel range into
    block_starts[i+1]-block_starts[i] chunks (sizes differing by at most one, larger first)."""
    _require_cuda(new_netIdxs, "new_netIdxs", torch.int16)
    _require_cuda(new_pixel_starts, "new_pixel_starts", torch.int32)
    net = netIdxs.to(torch.int16).contiguous()
    ps = pixel_starts.to(torch.int32).contiguous()
    bs = block_starts.to(torch.int32).contiguous()
    n = net.numel()
    if ps.numel() < n + 1 or bs.numel() < n + 1:
        raise ValueError("pixel_starts and block_starts need numNet + 1 entries")
    nb = min(new_netIdxs.numel(), new_pixel_starts.numel())   # the reference sizes them block_starts[-1]
    N.check(N.lib().snisr_assign_pixels(_ptr(net), _ptr(ps), _ptr(bs), n, nb, _ptr(new_netIdxs),
                                        _ptr(new_pixel_starts), _stream()), "assign_pixels")


_packed_cache: dict = {}


def pack_inference_params(params: torch.Tensor) -> torch.Tensor:
    """The MFMA fragment layout of the reference's per-group parameters (22,556 halves per group),
    cached per parameter tensor object and version.  The cache holds a weak reference to the tensor,
    so a new tensor that reuses a freed tensor's memory (and id) never finds the old packing."""
    _require_cuda(params, "params", torch.float16)
    if params.numel() % REF_MLP_PARAMS:
        raise ValueError(f"params must hold a multiple of {REF_MLP_PARAMS} halves")
    G = params.numel() // REF_MLP_PARAMS
    hit = _packed_cache.get(id(params))
    if hit is not None and hit[0]() is params and hit[1] == params._version:
        return hit[2]
    nbytes = N.lib().snisr_inference_mlp_packed_size(G)
    packed = torch.empty(max(nbytes, 16), dtype=torch.uint8, device=params.device)
    N.check(N.lib().snisr_pack_inference_mlp(_ptr(params), G, _ptr(packed), _stream()), "pack_inference_mlp")
    for k in [k for k, v in _packed_cache.items() if v[0]() is None]:
        del _packed_cache[k]
    _packed_cache[id(params)] = (weakref.ref(params), params._version, packed)
    return packed


def inference_mlp(max_samples_per_ray: int, num_block: int, num_thread: int, inverse_far: float, trans_stop: float,
                  netIdxs: torch.Tensor, transparency: torch.Tensor, query_pixel_indices: torch.Tensor,
                  pixel_starts: torch.Tensor, group_centers: torch.Tensor, samples: torch.Tensor,
                  rays_d: torch.Tensor, params: torch.Tensor, sample_results: torch.Tensor, n_blocks_device: torch.Tensor | None = None) -> None:
    """rendering/src/rayreflection_render.cu:303-335: the reflection MLP of every live sample of the
    num_block pixel blocks; writes (rgb, sigma) into sample_results [H, W, spr, 4] at the samples'
    original pixels.  rays_d may be a strided [H, W, 3] view such as ``rays_d[::2, ::2]``."""
    _require_cuda(sample_results, "sample_results", torch.float32)
    if rays_d.dim() != 3 or rays_d.shape[2] != 3 or rays_d.dtype != torch.float32 or not rays_d.is_cuda:
        raise ValueError("rays_d must be a float32 [H, W, 3] GPU tensor")
    H, W = rays_d.shape[0], rays_d.shape[1]
    spr = int(max_samples_per_ray)
    if samples.numel() != H * W * spr * 3 or sample_results.numel() != H * W * spr * 4:
        raise ValueError("samples / sample_results must hold H*W*max_samples_per_ray points / results")
    if rays_d.stride(2) != 1:
        rays_d = rays_d.contiguous()
    net = netIdxs.to(torch.int16).contiguous()
    ps = pixel_starts.to(torch.int32).contiguous()
    q = query_pixel_indices.to(torch.int32).contiguous()
    T = transparency.to(torch.float32).contiguous()
    ctr = group_centers.to(torch.float32).contiguous()
    smp = samples.to(torch.float32).contiguous()
    nb = int(num_block)
    if net.numel() < nb or ps.numel() < nb + 1:
        raise ValueError("netIdxs needs num_block entries and pixel_starts num_block + 1")
    if T.numel() != H * W:
        raise ValueError("transparency must hold H*W values")
    packed = pack_inference_params(params)
    G = params.numel() // REF_MLP_PARAMS
    if ctr.numel() < 3 * G:
        raise ValueError("group_centers must be [num_groups, 3]")
    Q = N.MlpQuery()
    Q.samples_per_ray, Q.n_groups, Q.n_blocks = spr, G, nb
    Q.inverse_far, Q.trans_stop = float(inverse_far), float(trans_stop)
    Q.net_idx, Q.pixel_starts, Q.query_pixel_indices = net.data_ptr(), ps.data_ptr(), q.data_ptr()
    Q.n_query = q.numel()
    Q.transparency, Q.group_centers, Q.samples = T.data_ptr(), ctr.data_ptr(), smp.data_ptr()
    Q.rays_d, Q.d_row_stride, Q.d_col_stride = rays_d.data_ptr(), rays_d.stride(0), rays_d.stride(1)
    Q.H, Q.W = H, W
    Q.packed_params, Q.sample_results = packed.data_ptr(), sample_results.data_ptr()
    if n_blocks_device is not None:  # the block count stays on the device (regroup_pixels)
        _require_cuda(n_blocks_device, "n_blocks_device", torch.int32)
        Q.n_blocks_device = n_blocks_device.data_ptr()
    N.check(N.lib().snisr_inference_mlp(C.byref(Q), _stream()), "inference_mlp")


def rendering_diffuse_octree_fp16(rays_d: torch.Tensor, visitedTiles: torch.Tensor, origin: torch.Tensor,
                                  groupMap: torch.Tensor, data_voxels: torch.Tensor, block_IndexMap: torch.Tensor,
                                  nodes_IndexMap: torch.Tensor, nodes_sampleFlag: torch.Tensor,
                                  voxels_start: torch.Tensor, centers: torch.Tensor, tile_size: float,
                                  voxel_size: float, sample_step: float, num_voxel: int, num_thread: int,
                                  trans_th: float, early_stop: float, soft_hit: bool, frame_diffuse: torch.Tensor,
                                  inverse_near: torch.Tensor, netIdxs: torch.Tensor,
                                  n_samples: torch.Tensor | None = None) -> None:
    """rendering/src/raydiffuse_render_octree_fp16.cu:238-287: diffuse colour of every ray over its
    visited tiles of the octree scene (frame_diffuse [H, W, 3]); the first sample with T < trans_th
    writes inverse_near and netIdxs (the caller initialises them, the reference with 1e-8 and -1).
    Interpolation and compositing run in FP32 (the reference: FP16).  ``n_samples`` (optional, int32
    [H, W]) receives the per-ray sample count."""
    _require_cuda(frame_diffuse, "frame_diffuse", torch.float32)
    _require_cuda(inverse_near, "inverse_near", torch.float32)
    _require_cuda(netIdxs, "netIdxs", torch.int16)
    dev = frame_diffuse.device
    rd = rays_d.to(dev, torch.float32).contiguous()
    n = rd.numel() // 3
    vis = visitedTiles.to(dev, torch.int32).contiguous()
    K = vis.shape[-1] if vis.dim() else 0
    if vis.numel() != n * K or frame_diffuse.numel() != 3 * n or inverse_near.numel() != n or netIdxs.numel() != n:
        raise ValueError("rays_d [H,W,3], visitedTiles [H,W,K], frame_diffuse [H,W,3], inverse_near/netIdxs [H,W]")
    if n_samples is not None:
        _require_cuda(n_samples, "n_samples", torch.int32)
    keep = [t.to(dev).contiguous() for t in (groupMap, data_voxels, block_IndexMap, nodes_IndexMap, voxels_start,
                                             centers, origin)]
    gm, dv, bm, nm, vs_, ct, org = keep
    if gm.dtype != torch.int32 or dv.dtype != torch.float16 or bm.dtype != torch.int32 or nm.dtype != torch.int16 \
            or vs_.dtype != torch.int64 or ct.dtype != torch.float32:
        raise ValueError("dtypes: groupMap i32, data_voxels f16, block_IndexMap i32, nodes_IndexMap i16, "
                         "voxels_start i64, centers f32")
    flags = nodes_sampleFlag.to(dev).contiguous()
    if flags.dtype == torch.bool:
        flags = flags.view(torch.uint8)
    org = org.to(torch.float32)
    O = N.OctreeDesc()
    O.group_map, O.data_voxels, O.block_index_map = gm.data_ptr(), dv.data_ptr(), bm.data_ptr()
    O.nodes_index_map, O.nodes_sample_flag, O.voxels_start = nm.data_ptr(), flags.data_ptr(), vs_.data_ptr()
    O.centers = ct.data_ptr()
    O.tile_size, O.voxel_size, O.sample_step, O.num_voxel = float(tile_size), float(voxel_size), float(sample_step), \
        int(num_voxel)
    width = rays_d.shape[1] if rays_d.dim() == 3 else 0
    N.check(N.lib().snisr_render_octree(C.byref(O), _ptr(org), _ptr(rd), _ptr(vis), n, K, width, float(trans_th),
                                        float(early_stop), 1 if soft_hit else 0, _ptr(frame_diffuse),
                                        _ptr(inverse_near), _ptr(netIdxs), _ptr(n_samples), _stream()),
            "rendering_diffuse_octree_fp16")


def regroup_max_blocks(n: int, pixels_per_block: int) -> int:
    return int(N.lib().snisr_regroup_max_blocks(int(n), int(pixels_per_block)))


def regroup_pixels(netIdxs: torch.Tensor, pixels_per_block: int, query: torch.Tensor, block_net: torch.Tensor,
                   block_starts: torch.Tensor, n_blocks: torch.Tensor, workspace: torch.Tensor) -> None:
    """The regrouping of FRender.render (sort_by_key, block_starts, assign_pixels: rendering/src/
    __init__.py:272-291) on the device without host synchronisation (snisr_regroup_pixels): query =
    pixel indices sorted by group, block_net / block_starts the <= pixels_per_block blocks of each group
    (-1 / n past the last block, block_net sized for regroup_max_blocks), n_blocks [1] the block count.
    workspace: uint8, snisr_regroup_workspace_size bytes."""
    net = netIdxs.reshape(-1)
    _require_cuda(net, "netIdxs", torch.int16)
    for t, name, dt in ((query, "query", torch.int32), (block_net, "block_net", torch.int16),
                        (block_starts, "block_starts", torch.int32), (n_blocks, "n_blocks", torch.int32),
                        (workspace, "workspace", torch.uint8)):
        _require_cuda(t, name, dt)
    n = net.numel()
    mb = block_net.numel()
    if query.numel() < n or block_starts.numel() < mb + 1 or n_blocks.numel() < 1:
        raise ValueError("query needs n entries, block_starts len(block_net) + 1, n_blocks 1")
    N.check(N.lib().snisr_regroup_pixels(_ptr(net), n, int(pixels_per_block), mb, _ptr(query), _ptr(block_net),
                                         _ptr(block_starts), _ptr(n_blocks), _ptr(workspace), workspace.numel(),
                                         _stream()), "regroup_pixels")
