"""Self-written autotuner for the MFMA render's launch schedule (snisr_tuning).

For one workload (scene shape, ray count, ray kind) it renders with every candidate schedule, times
each with HIP events on the current stream, checks that every schedule produced bit-identical pixels
and sample counts (the schedule must never change results), and keeps the fastest.  Results are
cached in memory and, when ``cache_path`` is given, in a JSON file keyed by the workload signature.

Candidates: ray_order in {0 (index / pixel-packet order), 1 (counting sort by first occupied tile
and direction octant), 2 (first tile, octant and origin cell), 3 (first tile and a finer origin
cell)} x mlp_threads in {256, 512} x sort_grid in {256, 1024} x record slots per ray in {the
workload's largest count + 25% ("tight"), twice the largest count ("loose")}, and for ray_order 0
the rays split into 2 or 4 chunks whose march overlaps the previous chunk's sort and MLP.
"""
from __future__ import annotations

import itertools
import json
import os

import torch

from . import device as D

CANDIDATES = [dict(ray_order=o, mlp_threads=m, sort_grid=g, records=r)
              for o, m, g, r in itertools.product((0, 1, 2, 3), (256, 512), (256, 1024), ("tight", "loose"))] + \
             [dict(ray_order=0, mlp_threads=m, sort_grid=g, records=r, chunks=k)
              for k, m, g, r in itertools.product((2, 4), (256, 512), (256, 1024), ("tight", "loose"))]
_memo: dict = {}


def signature(ds: D.DeviceScene, n_rays: int, image_width: int, has_t_hit: bool, per_ray_origins: bool) -> str:
    d = ds.desc
    return (f"n={n_rays};tiles={d.n_tiles};F={d.feat_dim};fdt={d.feat_dtype};w={d.mlp_width};L={d.mlp_layers};"
            f"mfma={d.mlp_packed_dtype};img={int(image_width > 0)};thit={int(has_t_hit)};org={int(per_ray_origins)}")


def tune_render(ds: D.DeviceScene, origins, dirs, t_hit=None, image_width=0, workspace=None, reps=3,
                cache_path: str | None = None, candidates=None):
    """Returns (best schedule dict, {schedule key: mean ms}).  Sizes the workspace first."""
    n = dirs.numel() // 3
    sig = signature(ds, n, image_width, t_hit is not None, origins.numel() != 3)
    if sig in _memo:
        return _memo[sig]
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as f:
            disk = json.load(f)
        if sig in disk:
            _memo[sig] = (disk[sig]["best"], disk[sig]["ms"])
            return _memo[sig]
    ws = workspace or D.MfmaWorkspace()
    ref = D.render_mfma(ds, origins, dirs, t_hit=t_hit, image_width=image_width, workspace=ws, check=True)
    loose = D.schedule_records(dict(records="loose"), ref["max_samples_per_ray"])
    # one workspace holding the largest slot count of the candidates and the whole batch's samples
    ws.get(ds, n, loose, int(ref["total_samples"] * 1.25) + 64,
           max(int(c.get("chunks", 1)) for c in (candidates or CANDIDATES)))
    table = {}
    for cand in candidates or CANDIDATES:
        out = D.render_mfma(ds, origins, dirs, t_hit=t_hit, image_width=image_width, workspace=ws, check=False,
                            tuning=cand)
        if not (torch.equal(out["rgb"], ref["rgb"]) and torch.equal(out["n_samples"], ref["n_samples"])):
            raise RuntimeError(f"schedule {cand} changed the render: the schedule must not affect results")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        for _ in range(reps):
            D.render_mfma(ds, origins, dirs, t_hit=t_hit, image_width=image_width, workspace=ws, check=False,
                          tuning=cand, outputs=out)
        ev[1].record()
        ev[1].synchronize()
        table[json.dumps(cand, sort_keys=True)] = ev[0].elapsed_time(ev[1]) / reps
    best = json.loads(min(table, key=table.get))
    _memo[sig] = (best, table)
    if cache_path:
        disk = {}
        if os.path.exists(cache_path):
            with open(cache_path) as f:
                disk = json.load(f)
        disk[sig] = {"best": best, "ms": table}
        with open(cache_path, "w") as f:
            json.dump(disk, f, indent=1, sort_keys=True)
    return best, table
