"""HIP-graph form of the MFMA render: the whole launch sequence of snisr_render_mfma_tuned (ray
schedule, count march, scans, emit march, record sort, persistent MLP, reduction) captured once for
fixed input / output / workspace tensors and replayed.  The library enqueues no memset and no host
synchronisation in that sequence, and every counter it needs is cleared by its own kernels, so a
replay is the complete render.  Update the input tensors in place between replays; the outputs
are overwritten by each replay."""
from __future__ import annotations

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
