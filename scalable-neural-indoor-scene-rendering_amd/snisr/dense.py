"""The training side's plain FP32 marcher over dense per-tile grids (render_diffuse_kernel,
training/src/cuda_func.cu:1687-1802) on the library's HIP kernel (snisr_render_dense).

The grids are the training tiles' own: per tile a dilated grid of NUM_VOXEL + 2 cells per side,
float4 (r, g, b, sigma) in x, y, z, channel order and an int16 cell flag (1 sampled, 0 holds a value
for interpolation only, -1 empty), exactly what ``Voxel-T{tile}-E{epoch}.npy`` and
``Node-T{tile}-E{epoch}.npy`` hold (training/scripts/tile/tile.py:415-432).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _native as N
from .device import _ptr, _require_cuda, _stream


def render_diffuse(rays_o: torch.Tensor, rays_d: torch.Tensor, visitedTiles: torch.Tensor,
                   SparseToGroup: torch.Tensor, tile_size: float, num_voxel: int, voxel_size: float,
                   sample_step: float, centers: torch.Tensor, voxels: torch.Tensor, nodes: torch.Tensor,
                   diffuses: torch.Tensor, depths: torch.Tensor, valid: torch.Tensor) -> None:
    """In place on diffuses f32[B, 3], depths f32[B] and valid i32[B], with the kernel's argument
    meaning: rays_o / rays_d f32[B, 3] (used as given), visitedTiles i32[B, K], SparseToGroup i32[T]
    (tile -> grid slot or -1), centers f32[T, 3], voxels f32[G, nv, nv, nv, 4], nodes i16[G, nv, nv, nv]
    with nv = num_voxel (the dilated cells per side)."""
    for t, name, dt in ((rays_o, "rays_o", torch.float32), (rays_d, "rays_d", torch.float32),
                        (visitedTiles, "visitedTiles", torch.int32), (SparseToGroup, "SparseToGroup", torch.int32),
                        (centers, "centers", torch.float32), (voxels, "voxels", torch.float32),
                        (nodes, "nodes", torch.int16), (diffuses, "diffuses", torch.float32),
                        (depths, "depths", torch.float32), (valid, "valid", torch.int32)):
        _require_cuda(t, name, dt)
    n = rays_o.numel() // 3
    K = visitedTiles.shape[-1] if visitedTiles.dim() else 0
    nv = int(num_voxel)
    cells = nv ** 3
    if rays_d.numel() != 3 * n or visitedTiles.numel() != n * K:
        raise ValueError("rays_o / rays_d must be [B, 3] and visitedTiles [B, K]")
    if diffuses.numel() != 3 * n or depths.numel() != n or valid.numel() != n:
        raise ValueError("diffuses [B, 3], depths [B] and valid [B]")
    if voxels.numel() % (4 * cells) or nodes.numel() != voxels.numel() // 4:
        raise ValueError("voxels must be [G, nv^3, 4] and nodes [G, nv^3] with nv = num_voxel")
    if centers.numel() != 3 * SparseToGroup.numel():
        raise ValueError("centers must be [T, 3] for the T tiles of SparseToGroup")
    g = N.DenseGridDesc()
    g.voxels, g.nodes = voxels.data_ptr(), nodes.data_ptr()
    g.sparse_to_group, g.centers = SparseToGroup.data_ptr(), centers.data_ptr()
    g.tile_size, g.voxel_size, g.sample_step, g.num_voxel = float(tile_size), float(voxel_size), \
        float(sample_step), nv
    N.check(N.lib().snisr_render_dense(C.byref(g), _ptr(rays_o), _ptr(rays_d), _ptr(visitedTiles), n, K,
                                       _ptr(diffuses), _ptr(depths), _ptr(valid), _stream()), "render_diffuse")
