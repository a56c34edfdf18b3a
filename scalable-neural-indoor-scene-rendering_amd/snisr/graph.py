"""HIP-graph form of the MFMA render: the whole launch sequence of snisr_render_mfma_tuned (ray
schedule, count march, scans, emit march, record sort, persistent MLP, reduction) captured once for
fixed input / output / workspace tensors and replayed.  The library enqueues no memset and no host
synchronisation in that sequence, and every counter it needs is cleared by its own kernels, so a
replay is the complete render.  Update the input tensors in place between replays; the outputs
are overwritten by each replay."""
from __future__ import annotations

import numpy as np
import torch

from . import device as D


class RenderGraph:
    def __init__(self, ds: D.DeviceScene, origins, dirs, t_hit=None, image_width=0, tuning=None, workspace=None,
                 capacity_factor: float = 1.5):
        self.ds, self.origins, self.dirs, self.t_hit = ds, origins, dirs, t_hit
        self.image_width, self.tuning = image_width, tuning
        self.ws = workspace or D.MfmaWorkspace()
        # size the workspace for this workload with headroom for inputs updated in place (replays
        # cannot grow it; overflowed() reports a replay that ran out of record space)
        first = D.render_mfma(ds, origins, dirs, t_hit=t_hit, image_width=image_width, workspace=self.ws,
                              check=True, tuning=tuning)
        self.capacity = D.records_for(first["max_samples_per_ray"], capacity_factor)
        self.sample_capacity = int(first["total_samples"] * capacity_factor) + 64
        self.ws.get(ds, dirs.numel() // 3, self.capacity, self.sample_capacity)
        self.outputs = {k: first[k] for k in ("rgb", "T", "n_samples", "depth")}
        self.graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._launch()            # warm-up on the capture stream
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        with torch.cuda.graph(self.graph):
            self._launch()

    def _launch(self):
        D.render_mfma(self.ds, self.origins, self.dirs, t_hit=self.t_hit, image_width=self.image_width,
                      workspace=self.ws, capacity=self.capacity, sample_capacity=self.sample_capacity,
                      check=False, tuning=self.tuning,
                      outputs=self.outputs)

    def overflowed(self) -> bool:
        """True if the last replay ran out of record slots or sample capacity (synchronises)."""
        return D.mfma_stats(self.ws.buf)[1] != 0

    def replay(self):
        self.graph.replay()
        return self.outputs


class FrameGraph:
    """One camera step captured in a HIP graph: the camera rays of every view (snisr_compute_rays),
    the proxy-mesh nearest hit bounding the march (snisr_mesh_first_hit, when a mesh is given) and the
    MFMA render of all views as one ray batch.  The workspace is sized from an eager render of the
    same views (record slots for the largest ray with 25% headroom, sorted-sample room for the
    frame's samples plus ``sample_headroom``); ``set_camera`` updates a view in place between
    replays, and ``overflowed`` reports a replay that outgrew the workspace (its rgb is NaN)."""

    def __init__(self, ds: D.DeviceScene, mesh, cams, tuning=None, sample_headroom: float = 0.05):
        dev = ds.sigma.device
        self.ds, self.mesh, self.tuning = ds, mesh, tuning
        self.V, self.H, self.W = len(cams), cams[0].H, cams[0].W
        self.n = self.V * self.H * self.W
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.intr = [t(c.intrinsic) for c in cams]
        self.c2w = [t(c.c2w) for c in cams]
        self.origin = t(cams[0].origin) if self.V == 1 else \
            t(np.repeat(np.stack([c.origin for c in cams]), self.H * self.W, axis=0))
        self.rays = torch.empty((self.V, self.H, self.W, 3), dtype=torch.float32, device=dev)
        self.t_hit = torch.empty(self.n, dtype=torch.float32, device=dev) if mesh is not None else None
        self.tri = torch.empty(self.n, dtype=torch.int32, device=dev) if mesh is not None else None
        self.ws = D.MfmaWorkspace()
        self._trace()
        first = D.render_mfma(ds, self.origin, self.rays, t_hit=self.t_hit, image_width=self.W, workspace=self.ws,
                              check=True, tuning=tuning)
        self.max_samples_per_ray = first["max_samples_per_ray"]
        self.frame_samples = first["total_samples"]
        self.capacity = D.records_for(self.max_samples_per_ray)
        self.sample_capacity = int(self.frame_samples * (1.0 + sample_headroom)) + 64
        self.ws.get(ds, self.n, self.capacity, self.sample_capacity)
        self.outputs = {k: first[k] for k in ("rgb", "T", "n_samples", "depth")}
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            self._step()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self._step()

    def _trace(self):
        for v in range(self.V):
            D.compute_rays(self.intr[v], self.c2w[v], self.H, self.W, out=self.rays[v])
        if self.mesh is not None:
            self.mesh.first_hit(self.origin, self.rays, image_width=self.W, t_out=self.t_hit, id_out=self.tri)

    def _step(self, events=None):
        self._trace()
        D.render_mfma(self.ds, self.origin, self.rays, t_hit=self.t_hit, image_width=self.W, workspace=self.ws,
                      capacity=self.capacity, sample_capacity=self.sample_capacity, check=False, tuning=self.tuning,
                      outputs=self.outputs, events=events)

    def eager(self, events=None):
        """The same step without the graph (``events``: 7 stage events of the render)."""
        self._step(events)
        return self.outputs

    def set_camera(self, v: int, cam):
        self.intr[v].copy_(torch.from_numpy(np.ascontiguousarray(cam.intrinsic)))
        self.c2w[v].copy_(torch.from_numpy(np.ascontiguousarray(cam.c2w)))
        o = torch.from_numpy(np.ascontiguousarray(cam.origin)).to(self.origin.device)
        if self.V == 1:
            self.origin.copy_(o)
        else:
            self.origin[v * self.H * self.W:(v + 1) * self.H * self.W].copy_(o.expand(self.H * self.W, 3))

    def overflowed(self) -> bool:
        return D.mfma_stats(self.ws.buf)[1] != 0

    def replay(self):
        self.graph.replay()
        return self.outputs
