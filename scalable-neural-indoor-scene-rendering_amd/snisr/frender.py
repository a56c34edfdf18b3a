"""Frame driver in the form of the reference's ``FRender.render`` (rendering/src/__init__.py:247-341),
calling the FASTRENDERING stand-in operators: rays, tile tracing, the octree diffuse pass, pixel
regrouping by group, the screen-space group boundary, four chunks of 16 inverse-depth reflection
samples through the per-group MLP with compositing, bilinear upsampling of the reflection layer,
diffuse + reflection clamped to [0, 1], and optionally the U-Net refinement (load_cnn).
"""
from __future__ import annotations

import math

import numpy as np
import torch
import torch.nn.functional as F

from . import fastrendering as FR
from .scene import Camera, OctreeScene


class FRender:
    """Holds the scene tensors on the device, like the reference's render_prepare
    (rendering/src/__init__.py:95-146).  ``params``: reflection MLPs, f16[G * 22556]; ``group_centers``:
    f32[G, 3]."""

    def __init__(self, osc: OctreeScene, params: np.ndarray, group_centers: np.ndarray, H=720, W=1280,
                 max_tracing_tile=30, reflection_scale=2, device="cuda"):
        dev = torch.device(device)
        self.device = dev
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.H, self.W = H, W
        self.max_tracing_tile = max_tracing_tile
        self.reflection_scale = reflection_scale
        self.DH, self.DW = H // reflection_scale, W // reflection_scale
        self.downscale_numPixel = self.DH * self.DW
        self.num_thread = 256
        self.max_pixels_per_block = 64
        self.trans_decay = 0.8
        self.early_stop_R = 0.01
        self.early_stop_D = 0.01
        self.tile_IndexMap = t(osc.tile_index_map)
        self.tile_shape = torch.tensor(osc.tile_shape, dtype=torch.int32)
        self.scene_min_corner = torch.from_numpy(osc.scene_min)
        self.tile_size = osc.tile_size
        self.scene_size = torch.from_numpy(np.asarray(osc.tile_shape, np.float32) * np.float32(osc.tile_size))
        self.tile2group = t(osc.group_map)
        self.data_voxels = t(osc.data_voxels)
        self.block_IndexMap = t(osc.block_index_map)
        self.nodes_IndexMap = t(osc.nodes_index_map)
        self.nodes_sampleFlag = t(osc.nodes_sample_flag)
        self.voxels_start = t(osc.voxels_start)
        self.centers = t(osc.centers)
        self.voxel_size = osc.voxel_size
        self.sample_step = osc.sample_step
        self.block_num_voxel = osc.num_voxel
        self.net_params = t(params)
        self.group_centers = t(group_centers)
        self.far = float(np.linalg.norm(osc.tile_size * np.asarray(osc.tile_shape, np.float64)))
        self.inverse_far = 1.0 / self.far
        self.ksize_boundary = 3
        self.max_depth = float(np.max(np.asarray(osc.tile_shape, np.float64) * osc.tile_size))
        self.render_mode = 1
        self.enable_Unet = True
        self.cnn = None
        self.frame = None

    def set_mode(self, mode_type: int):
        """Render mode of the reference (rendering/src/__init__.py:199-200, 296-303, 330-332): 1 the
        full frame, 2 the diffuse layer only, 3 the upsampled reflection layer only, 4 depth
        (1 / inverse_near / max_depth in all three channels).  A captured frame is recaptured."""
        if mode_type not in (1, 2, 3, 4):
            raise ValueError(f"render mode must be 1, 2, 3 or 4, got {mode_type}")
        self.render_mode = int(mode_type)
        self._graph = None

    def set_unet(self, flag: bool):
        """Switches the refinement network on or off (rendering/src/__init__.py:201-202); it runs only
        once loaded with load_cnn."""
        self.enable_Unet = bool(flag)
        self._graph = None

    def load_cnn(self, state_dict=None, half: bool = True, backend: str = "hip"):
        """Enables the screen-space refinement network (the reference's CNN / Unet,
        rendering/src/__init__.py:79-90); ``state_dict`` as saved by the reference (None keeps the
        random initialisation).  backend "hip": the library's FP16 MFMA convolutions (HipUnet);
        "torch": the module through torch (FP16 if ``half``)."""
        from .cnn import CNN, HipUnet
        net = CNN()
        if state_dict is not None:
            net.load_state_dict(state_dict)
        if backend == "hip":
            self.cnn = HipUnet(net, device=self.device)
            return
        if backend != "torch":
            raise ValueError(f"backend must be 'hip' or 'torch', got {backend!r}")
        net = net.to(self.device).eval()
        self.cnn = net.half() if half else net

    @classmethod
    def from_render_data(cls, path, sample_step_scale: float = 0.2, **kw) -> "FRender":
        """FRender of a renderData.npz scene (the reference's render_prepare,
        rendering/src/__init__.py:98-135)."""
        from .render_data import load_render_data
        rd = load_render_data(path, sample_step_scale)
        fr = cls(rd.scene, rd.net_params, rd.group_centers, **kw)
        fr.group_names = rd.group_names
        return fr

    def set_camera(self, cam: Camera):
        """New camera; copied into the existing device tensors, so a captured frame (render_graphed)
        renders the new view on its next replay."""
        for name, v in (("intrinsic", cam.intrinsic), ("c2w", cam.c2w), ("origin", cam.origin)):
            t = torch.from_numpy(np.ascontiguousarray(v))
            cur = getattr(self, name, None)
            if cur is not None and cur.shape == t.shape and cur.dtype == t.dtype:
                cur.copy_(t.to(self.device))
            else:
                setattr(self, name, t.to(self.device))

    @torch.no_grad()
    def render(self):
        """One frame in the reference's call sequence, including its regrouping through sort_by_key /
        assign_pixels with the two host synchronisations (the returned group count and block count)."""
        rays_d, inverse_near, netIdxs = self._diffuse()
        s = self.reflection_scale
        npx = self.downscale_numPixel
        dev = self.device
        query = torch.arange(0, npx, 1, device=dev, dtype=torch.int32)
        pixel_starts = torch.arange(0, npx, 1, device=dev, dtype=torch.int32)
        numNet = FR.sort_by_key(netIdxs, query, pixel_starts)
        tail = torch.full((1,), npx, dtype=torch.int32, device=dev)
        per_net = torch.cat([pixel_starts[1:numNet], tail]) - pixel_starts[:numNet]
        pixel_starts = torch.cat([pixel_starts[:numNet], tail])
        block_starts = torch.cat([torch.zeros(1, dtype=torch.int32, device=dev),
                                  torch.cumsum(torch.ceil(per_net / self.max_pixels_per_block), dim=0).int()])
        netIdxs = netIdxs[:numNet]
        n_blocks = int(block_starts[-1])
        new_net = torch.empty(n_blocks, device=dev, dtype=torch.int16)
        new_starts = torch.empty(n_blocks, device=dev, dtype=torch.int32)
        FR.assign_pixels(netIdxs, pixel_starts, block_starts, self.num_thread, new_net, new_starts)
        new_starts = torch.cat([new_starts, tail])
        self.n_reflection_blocks = n_blocks
        return self._reflection_and_compose(rays_d, inverse_near, n_blocks, new_net, query, new_starts, None)

    @torch.no_grad()
    def render_device(self):
        """The same frame with no host synchronisation: the regrouping runs on the device
        (regroup_pixels) and inference_mlp reads the block count from device memory.  The output
        equals render()'s."""
        rays_d, inverse_near, netIdxs = self._diffuse()
        mb, query, block_net, block_starts, nblk = self._regroup(netIdxs)
        return self._reflection_and_compose(rays_d, inverse_near, mb, block_net, query, block_starts, nblk)

    def _regroup(self, netIdxs):
        """Device-side regrouping into buffers kept across frames (so captured kernels can read them)."""
        npx, P = self.downscale_numPixel, self.max_pixels_per_block
        dev = self.device
        if getattr(self, "_rg", None) is None:
            from . import _native as N
            mb = FR.regroup_max_blocks(npx, P)
            self._rg = (mb, torch.empty(npx, dtype=torch.int32, device=dev),
                        torch.empty(mb, dtype=torch.int16, device=dev),
                        torch.empty(mb + 1, dtype=torch.int32, device=dev),
                        torch.empty(1, dtype=torch.int32, device=dev),
                        torch.empty(int(N.lib().snisr_regroup_workspace_size(npx)), dtype=torch.uint8, device=dev))
        mb, query, block_net, block_starts, nblk, ws = self._rg
        FR.regroup_pixels(netIdxs, P, query, block_net, block_starts, nblk, ws)
        self.n_blocks_device = nblk
        return mb, query, block_net, block_starts, nblk

    def render_graphed(self):
        """render_device captured once in a HIP graph and replayed: the diffuse pass, the device-side
        regrouping (its sort and scans clear their counters with their own kernels, no memset) and the
        reflection / composition pass are one graph.  The returned frame tensor is the graph's output
        buffer, overwritten by the next replay; set_camera between replays renders the new view."""
        if getattr(self, "_graph", None) is None:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):  # warm-up: packs the MLP parameters, sizes the workspaces
                self.render_device()
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                out = self.render_device()
            self._graph = (g, out)
        g, out = self._graph
        g.replay()
        self.frame = out
        return out

    def _diffuse(self):
        dev = self.device
        H, W, s = self.H, self.W, self.reflection_scale
        # compute_rays writes every direction and tracing_tiles every one of the K entries per pixel
        # (-1 padded), so these two need no fill (the reference's render() fills them with -1)
        rays_d = torch.empty((H, W, 3), device=dev, dtype=torch.float32)
        visited = torch.empty((H, W, self.max_tracing_tile), device=dev, dtype=torch.int32)
        self.frame_diffuse = torch.full((H, W, 3), -1, device=dev, dtype=torch.float32)
        inverse_near = torch.full((H, W), 1e-8, device=dev, dtype=torch.float32)
        netIdxs = torch.full((H, W), -1, device=dev, dtype=torch.int16)
        FR.compute_rays(self.intrinsic, self.c2w, self.num_thread, rays_d)
        FR.tracing_tiles(rays_d, self.tile_IndexMap, self.origin, self.tile_shape, self.scene_min_corner,
                         self.scene_size, self.tile_size, self.num_thread, visited)
        FR.rendering_diffuse_octree_fp16(rays_d, visited, self.origin, self.tile2group, self.data_voxels,
                                         self.block_IndexMap, self.nodes_IndexMap, self.nodes_sampleFlag,
                                         self.voxels_start, self.centers, self.tile_size, self.voxel_size,
                                         self.sample_step, self.block_num_voxel, self.num_thread, self.trans_decay,
                                         self.early_stop_D, False, self.frame_diffuse, inverse_near, netIdxs)
        del visited
        self.large_inverse_near = inverse_near.clone()
        self.large_netIdx = netIdxs.clone()
        inverse_near = inverse_near[::s, ::s].contiguous()
        netIdxs = netIdxs[::s, ::s].flatten()
        return rays_d, inverse_near, netIdxs

    def _reflection_and_compose(self, rays_d, inverse_near, n_blocks, new_net, query, new_starts, n_blocks_device):
        dev = self.device
        H, W, s = self.H, self.W, self.reflection_scale
        self.boundary = torch.zeros((H, W, 1), dtype=torch.bool, device=dev)
        FR.compute_group_boundary(self.large_netIdx, self.boundary, self.ksize_boundary, self.num_thread * 2)
        if self.render_mode == 2:
            self.frame = self.frame_diffuse
            return self.frame
        if self.render_mode == 4:
            self.frame = (1 / self.large_inverse_near[..., None] / self.max_depth).repeat(1, 1, 3)
            return self.frame
        sample_far = 1.0 / inverse_near + self.far
        inverse_bound = (1.0 / sample_far - inverse_near) / 63
        transparency = torch.ones((self.DH, self.DW, 1), device=dev, dtype=torch.float32)
        spec = torch.zeros((self.DH, self.DW, 3), device=dev, dtype=torch.float32)
        rays_ds = rays_d[::s, ::s]
        # inverse_z_sampling writes every sample point and step, so those buffers are allocated once
        # without a fill; inference_mlp writes only the samples it evaluates, so the results are
        # cleared before every round of 16 samples
        sample_results = torch.empty((self.DH, self.DW, 16, 4), device=dev, dtype=torch.float32)
        samples = torch.empty((self.DH, self.DW, 16, 3), device=dev, dtype=torch.float32)
        steps = torch.empty((self.DH, self.DW, 16), device=dev, dtype=torch.float32)
        for step in range(0, 64, 16):
            sample_results.zero_()
            FR.inverse_z_sampling(self.origin, rays_ds, inverse_near, inverse_bound, self.num_thread, step, 16,
                                  samples, steps)
            FR.inference_mlp(16, n_blocks, self.num_thread, self.inverse_far, self.early_stop_R, new_net,
                             transparency, query, new_starts, self.group_centers, samples, rays_ds, self.net_params,
                             sample_results, n_blocks_device=n_blocks_device)
            FR.integrate_points(sample_results, steps, self.num_thread * 2, self.early_stop_R, transparency, spec)
        self.frame_speculars_lowres = spec
        self.frame_speculars = F.interpolate(spec[None].permute(0, 3, 1, 2), scale_factor=s, mode="bilinear",
                                             align_corners=True)[0].permute(1, 2, 0)
        if self.render_mode == 3:
            self.frame = self.frame_speculars
            return self.frame
        self.frame = torch.clamp(self.frame_diffuse + self.frame_speculars, 0, 1)
        if self.cnn is not None and self.enable_Unet:
            from .cnn import HipUnet, refine
            if isinstance(self.cnn, HipUnet):
                self.frame = self.cnn(self.frame.contiguous(), self.large_inverse_near.contiguous(),
                                      self.boundary.contiguous())
            else:
                self.frame = refine(self.cnn, self.frame, self.large_inverse_near, self.boundary)
        return self.frame
