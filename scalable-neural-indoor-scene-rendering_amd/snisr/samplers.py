"""The reference's training samplers under their CUDA_EXT names (training/src/binding.cpp:4-12,
training/src/Sample.cpp:97-158): ``Sample_sparse``, ``Sample_reflection``, ``Sample_uniform`` and
``Sample_bg`` with the same arguments and in-place outputs, on the HIP kernels of
``libsnisr_hip.so`` (snisr_sample_*).  Shapes are checked as Sample.cpp's TORCH_CHECKs do
(rays [B,3], z_vals [B,N], node_flags 3-d, voxels 4-d); inputs must be contiguous CUDA tensors
(CHECK_INPUT), outputs contiguous float32 CUDA tensors."""
from __future__ import annotations

import torch

from . import _native as N
from .device import _ptr, _require_cuda, _stream


def _check_rays(rays_start: torch.Tensor, rays_dir: torch.Tensor, z_vals: torch.Tensor):
    for t, name in ((rays_start, "rays_start"), (rays_dir, "rays_dir")):
        _require_cuda(t, name, torch.float32)
        if t.dim() != 2 or t.shape[1] != 3:
            raise ValueError(f"{name} must be [B, 3]")
    _require_cuda(z_vals, "z_vals", torch.float32)
    if z_vals.dim() != 2 or z_vals.shape[0] != rays_start.shape[0] or rays_dir.shape[0] != rays_start.shape[0]:
        raise ValueError("rays_start, rays_dir and z_vals must agree in B; z_vals must be [B, N]")


def _flags(node_flags: torch.Tensor, num_voxel: int) -> torch.Tensor:
    _require_cuda(node_flags, "node_flags", torch.int16)
    if node_flags.dim() != 3 or node_flags.numel() != num_voxel ** 3:
        raise ValueError("node_flags must be [num_voxel, num_voxel, num_voxel] int16")
    return node_flags


def Sample_sparse(rays_start, rays_dir, node_flags, cx, cy, cz, tile_size, num_voxel, z_vals, dists):
    """training/src/Sample.cpp:125-148 (kernel Sample_kernel.cu:181-434)."""
    _check_rays(rays_start, rays_dir, z_vals)
    _require_cuda(dists, "dists", torch.float32)
    if dists.shape != z_vals.shape:
        raise ValueError("dists must have the shape of z_vals")
    fl = _flags(node_flags, int(num_voxel))
    N.check(N.lib().snisr_sample_sparse(_ptr(rays_start), _ptr(rays_dir), rays_start.shape[0], _ptr(fl), float(cx),
                                        float(cy), float(cz), float(tile_size), int(num_voxel), z_vals.shape[1],
                                        _ptr(z_vals), _ptr(dists), _stream()), "Sample_sparse")


def Sample_reflection(rays_start, rays_dir, node_flags, voxels, cx, cy, cz, tile_size, far, num_voxel, T_th, z_vals):
    """training/src/Sample.cpp:47-76 (kernel Sample_kernel.cu:16-179).  The densities are read as a
    [num_voxel]^3 x 4 grid (x-major), as the kernel's OVERLAP interpolation indexes them
    (interpolation.h:120-134).  The reference's training step passes its dilated (num_voxel + 2)^3
    grid here (tile.py:743-744 through utils.py:786-793), which that indexing reads with num_voxel
    strides; a larger grid is therefore accepted and read the same way."""
    _check_rays(rays_start, rays_dir, z_vals)
    fl = _flags(node_flags, int(num_voxel))
    _require_cuda(voxels, "voxels", torch.float32)
    if voxels.dim() != 4 or voxels.shape[-1] != 4 or voxels.numel() < 4 * int(num_voxel) ** 3:
        raise ValueError("voxels must be a 4-d float32 [..., 4] grid of at least num_voxel^3 cells")
    N.check(N.lib().snisr_sample_reflection(_ptr(rays_start), _ptr(rays_dir), rays_start.shape[0], _ptr(fl),
                                            _ptr(voxels), float(cx), float(cy), float(cz), float(tile_size),
                                            float(far), int(num_voxel), float(T_th), z_vals.shape[1], _ptr(z_vals),
                                            _stream()), "Sample_reflection")


def Sample_uniform(rays_start, rays_dir, cx, cy, cz, tile_size, z_vals):
    """training/src/Sample.cpp:149-158 (kernel Sample_kernel.cu:436-479)."""
    _check_rays(rays_start, rays_dir, z_vals)
    N.check(N.lib().snisr_sample_uniform(_ptr(rays_start), _ptr(rays_dir), rays_start.shape[0], float(cx), float(cy),
                                         float(cz), float(tile_size), z_vals.shape[1], _ptr(z_vals), _stream()),
            "Sample_uniform")


def Sample_bg(rays_start, rays_dir, bg_zval, tile_center, tile_size, sample_range, z_vals):
    """training/src/Sample.cpp:78-97 (kernel Sample_kernel.cu:481-542): bg_zval [B, 1] (-1 = no
    background sample), tile_center [3] on the device."""
    _check_rays(rays_start, rays_dir, z_vals)
    _require_cuda(bg_zval, "bg_zval", torch.float32)
    _require_cuda(tile_center, "tile_center", torch.float32)
    if bg_zval.dim() != 2 or bg_zval.shape[0] != rays_start.shape[0] or tile_center.dim() != 1 or tile_center.numel() < 3:
        raise ValueError("bg_zval must be [B, 1] and tile_center [3]")
    N.check(N.lib().snisr_sample_bg(_ptr(rays_start), _ptr(rays_dir), _ptr(bg_zval), rays_start.shape[0],
                                    _ptr(tile_center), float(tile_size), float(sample_range), z_vals.shape[1],
                                    _ptr(z_vals), _stream()), "Sample_bg")
