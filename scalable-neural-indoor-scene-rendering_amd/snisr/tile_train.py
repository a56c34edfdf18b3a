"""The reference's tile training step on the HIP ops: the sampling wrappers of
training/scripts/utils.py:786-820 and the forward of Tile.train_batch (training/scripts/tile/tile.py:
631-770) for one tile, so that a training loop written like the reference's (loss, backward,
optimizer step) runs on the library.

- ``sparse_sampling``, ``sampling_background``, ``reflection_sampling``, ``uniform_sampling``: the
  utils.py wrappers (the flag and voxel grids permuted to x-major as they do), on ``snisr.samplers``.
- ``TileState`` holds what Tile keeps between steps: the dilated node flags, their occupied cells
  (recomputed when the flags change), the tile and group geometry, and the last baked grid.
- ``train_batch_forward(tile, surf, refl, batch_data, train_flag, ...)`` returns the prediction and
  the intermediate terms Tile.train_batch hands to its loss (criterion.py), differentiable with
  respect to both MLPs' parameters.  train_flag 0: diffuse only (the bake, grid samples, background
  samples, chained composites); 1: specular only (the diffuse part without gradients, reflection
  samples through the reflection MLP, SH, composite, clamp); 2: both.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import torch

from . import samplers as SMP
from . import train_ref as TR


@torch.no_grad()
def sparse_sampling(rays_o, rays_d, node_flags, num_voxel, center, size, nsamples):
    """utils.sparse_sampling (utils.py:797-806): node_flags [D, H, W] (z, y, x) of the tile's
    num_voxel^3 voxels; returns (z_vals, dists) [B, nsamples], -1 where the ray takes no sample."""
    z = torch.full((rays_o.shape[0], nsamples), -1.0, dtype=torch.float32, device=rays_o.device)
    d = torch.full_like(z, -1.0)
    SMP.Sample_sparse(rays_o, rays_d, node_flags.permute(2, 1, 0).contiguous(), *[float(c) for c in center],
                      float(size), int(num_voxel), z, d)
    return z, d


@torch.no_grad()
def uniform_sampling(rays_o, rays_d, center, size, nsamples):
    """utils.uniform_sampling (utils.py:808-813)."""
    z = torch.full((rays_o.shape[0], nsamples), -1.0, dtype=torch.float32, device=rays_o.device)
    SMP.Sample_uniform(rays_o, rays_d, *[float(c) for c in center], float(size), z)
    return z


@torch.no_grad()
def sampling_background(rays_o, rays_d, bgdepth, tile_center, tile_size, nsamples, sample_range):
    """utils.sampling_background (utils.py:815-820): bgdepth [B, 1] (-1: no background)."""
    z = torch.full((rays_o.shape[0], nsamples), -1.0, dtype=torch.float32, device=rays_o.device)
    tc = torch.as_tensor(tile_center, dtype=torch.float32, device=rays_o.device).reshape(3)
    SMP.Sample_bg(rays_o, rays_d, bgdepth.contiguous(), tc, float(tile_size), float(sample_range), z)
    return z


@torch.no_grad()
def reflection_sampling(rays_o, rays_d, node_flags, voxels, num_voxel, center, size, nsamples, far, thresh):
    """utils.reflection_sampling (utils.py:786-793): node_flags [D, H, W] of the num_voxel^3 voxels;
    voxels the tile's baked grid as this library keeps it, [D', D', D', 4] channel-last (z, y, x)
    (the reference passes voxels[0], its [4, D', D', D'] view, and permutes it to x-major)."""
    z = torch.full((rays_o.shape[0], nsamples), -1.0, dtype=torch.float32, device=rays_o.device)
    SMP.Sample_reflection(rays_o, rays_d, node_flags.permute(2, 1, 0).contiguous(),
                          voxels.detach().permute(2, 1, 0, 3).contiguous(), *[float(c) for c in center], float(size),
                          float(far), int(num_voxel), float(thresh), z)
    return z


@dataclass
class TileState:
    """A tile's training state (Tile.__init__ / build_training_context, tile.py:19-105): node flags of
    the dilated grid (int16 [NV + 2]^3, -1 = empty), its geometry and the group's normalisation."""
    nodes_flag: torch.Tensor
    num_voxel: int
    voxel_size: float
    tile_center: tuple
    group_center: tuple
    group_size: float
    far: float
    _cells: torch.Tensor | None = field(default=None, repr=False)
    _flags_seen: tuple | None = field(default=None, repr=False)
    voxels: torch.Tensor | None = field(default=None, repr=False)

    @property
    def tile_size(self) -> float:
        return self.num_voxel * self.voxel_size

    @property
    def dilate(self) -> int:
        return self.num_voxel + 2

    @property
    def min_corner(self) -> tuple:
        """tile_center - tile_size / 2 - voxel_size (tile.py:80), in FP32."""
        c = torch.tensor(self.tile_center, dtype=torch.float32)
        return tuple((c - self.tile_size / 2.0 - self.voxel_size).tolist())

    def cells(self) -> torch.Tensor:
        """Occupied cells of the dilated grid, recomputed when the flags changed: another tensor
        assigned (pruning replaces it, tile.py:379) or the same one written in place."""
        f = self.nodes_flag
        seen = self._flags_seen
        if self._cells is None or seen is None or seen[0] is not f or seen[1] != f._version:
            self._cells = TR.flag_cells(f)
            self._flags_seen = (f, f._version)
        return self._cells


def train_batch_forward(tile: TileState, surf: "TR.PEMLP", refl: "TR.PEMLP", batch_data: torch.Tensor,
                        train_flag: int, nsamples: int, bgsamples: int, rsamples: int, bg_range: float,
                        patch: int, deg: int = 2):
    """The forward of Tile.train_batch (tile.py:631-770) for batch_data [B, patch, patch, 14]
    (rays_o, rays_d, bgdepth, gtcolor, gt_diffuse, mask).  Returns None when no ray of the batch
    samples the tile (the reference returns early), else a dict with pred [B, P, P, 3], gtcolor,
    gt_diffuse, mask (with the rays that take no sample cleared), diffuse_intile, diffuse and, for
    the specular flags, coeffi [B*P*P, rsamples, 3, (deg+1)^2] and rgb_vd."""
    if train_flag not in (0, 1, 2):
        raise ValueError("train_flag must be 0, 1 or 2")
    B = batch_data.shape[0]
    rays_o = batch_data[..., :3].reshape(-1, 3).contiguous()
    rays_d = batch_data[..., 3:6].reshape(-1, 3).contiguous()
    bgdepth = batch_data[..., 6:7].reshape(-1, 1).contiguous()
    gtcolor = batch_data[..., 7:10].reshape(-1, 3).contiguous()
    gt_diffuse = batch_data[..., 10:13].contiguous()
    mask = batch_data[..., 13:14].reshape(-1, 1).clone()
    nv, inner = tile.num_voxel, tile.nodes_flag[1:-1, 1:-1, 1:-1]
    z_vals, sparse_dists = sparse_sampling(rays_o, rays_d, inner, nv, tile.tile_center, tile.tile_size, nsamples)
    invalid = torch.any(z_vals == -1, dim=-1)
    mask[invalid] = 0
    if float(mask.sum()) == 0.0:
        return None
    invalid = mask[:, 0] == 0
    mc, D = tile.min_corner, tile.dilate

    def diffuse_branch():
        bg_z = sampling_background(rays_o, rays_d, bgdepth, tile.tile_center, tile.tile_size, bgsamples, bg_range)
        bg_sample = rays_o[:, None, :] + bg_z[..., None] * rays_d[:, None, :]
        bg_rgb, bg_sigma = TR.inference_diffuse(surf, bg_sample, tile.group_center, tile.group_size)
        tile.voxels = TR.bake_voxels(surf, tile.cells(), D, mc, tile.voxel_size, tile.group_center, tile.group_size)
        samples = rays_o[:, None, :] + z_vals[..., None] * rays_d[:, None, :]
        rgb, sigma = TR.query_voxel(tile.voxels, samples, mc, tile.voxel_size)
        total_z = torch.cat([z_vals, bg_z], dim=-1)
        dists = torch.cat([total_z[..., 1:] - total_z[..., :-1],
                           torch.full((rays_o.shape[0], 1), 1e10, device=rays_o.device)], dim=-1)
        dists[:, :nsamples] = sparse_dists
        dists[invalid] = 0
        d_in, left, _ = TR.volume_rendering(rgb, sigma, dists[:, :nsamples, None].contiguous())
        d_bg, _, _ = TR.volume_rendering(bg_rgb, bg_sigma, dists[:, nsamples:, None].contiguous(), left[:, -1:, :])
        d_in = d_in.reshape(B, patch, patch, 3)
        d_bg = d_bg.reshape(B, patch, patch, 3)
        bg_mask = bgdepth.reshape(B, patch, patch, 1) != -1
        return d_in, d_in + d_bg * bg_mask

    if train_flag in (0, 2):
        diffuse_intile, diffuse = diffuse_branch()
    else:
        with torch.no_grad():
            diffuse_intile, diffuse = diffuse_branch()
    out = dict(gtcolor=gtcolor.reshape(B, patch, patch, 3), gt_diffuse=gt_diffuse, mask=mask.reshape(B, patch, patch, 1),
               diffuse_intile=diffuse_intile, diffuse=diffuse, coeffi=None)
    if train_flag == 0:
        out["pred"] = diffuse
        return out
    rz = reflection_sampling(rays_o, rays_d, inner, tile.voxels, nv, tile.tile_center, tile.tile_size, rsamples,
                             tile.far, 1.0)
    rmask = torch.all(rz != -1, dim=-1)
    rsamples_xyz = rays_o[:, None, :] + rz[..., None] * rays_d[:, None, :]
    raw = refl(TR.normalize_points(rsamples_xyz, tile.group_center, tile.far))
    reflection = TR.sh_sigmoid(raw, rays_d, deg)
    rsigma = torch.relu(raw[..., -1:])
    nb = (deg + 1) ** 2
    out["coeffi"] = raw[..., :-1].reshape(*raw.shape[:-1], 9, 3)[..., :nb, :].transpose(-1, -2)
    rd = torch.cat([rz[..., 1:] - rz[..., :-1], torch.full((rays_o.shape[0], 1), 1e10, device=rays_o.device)], dim=-1)
    rd[invalid] = 0
    rgb_vd, _, _ = TR.volume_rendering(reflection, rsigma, rd[..., None].contiguous())
    rgb_vd = rgb_vd.reshape(B, patch, patch, 3)
    rmask = rmask.reshape(B, patch, patch, 1)
    base = diffuse.detach() if train_flag == 1 else diffuse
    out["rgb_vd"] = rgb_vd
    out["pred"] = torch.clamp(base + rmask * rgb_vd, 0, 1)
    return out
