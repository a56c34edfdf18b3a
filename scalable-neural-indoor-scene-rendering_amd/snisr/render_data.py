"""The reference's on-disk scene format, ``renderData.npz``, read and written on the host.

The reference loads it in ``FRender.render_prepare`` (rendering/src/__init__.py:98-135): the octree
scene (tile grid, per-tile 4^3 block map, per-node dilated cell maps, sample flags, voxel starts and
half4 voxels), the reflection MLPs of every group (22,556 halves each) and the group centres.
``load_render_data`` turns such a file into an ``OctreeScene`` plus the reflection parameters, with
the derived scalars computed as the reference computes them (tile_size = num_voxel * voxel_size,
block grid num_voxel / 2^octree_depth + 2, sample_step = sample_step_scale * voxel_size, in double
precision; the operators round them to FP32 like the pybind float arguments).  ``save_render_data``
writes the same keys, so scenes round-trip and synthetic scenes can be handed to the reference.
``pack_reflection_mlp`` packs a trained reflection MLP (the state dict of the reference's
``MLP`` / ``GeneralMLP``, training/scripts/networks.py:17-41, as saved in Coeffi-TShared-E*.pt by
training/scripts/tile/tileGroup.py:546-549) into that 22,556-half layout: per linear layer the bias,
then the weight as [in][out] (the load order of rayreflection_render.cu:197-287).

Files are read with ``numpy.load(allow_pickle=False)``: ``group_names`` must be a plain string
array.  Nothing here touches the GPU.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .scene import REF_MLP_DIMS, REF_MLP_PARAMS, OctreeScene

OCTREE_DEPTH = 2   # 4^3 blocks per tile: the block DDA of raydiffuse_render_octree_fp16.cu:6


@dataclasses.dataclass
class RenderData:
    scene: OctreeScene
    net_params: np.ndarray      # f16[G * 22556]
    group_centers: np.ndarray   # f32[G, 3]
    group_names: np.ndarray     # str[G]

    @property
    def n_groups(self) -> int:
        return self.group_centers.shape[0]


def _scalar(d, key):
    v = np.asarray(d[key]).reshape(-1)
    if v.size < 1:
        raise ValueError(f"renderData: '{key}' is empty")
    return v[0]


def load_render_data(path, sample_step_scale: float = 0.2) -> RenderData:
    """Reads a renderData.npz written by the reference's pipeline (or by ``save_render_data``).
    ``sample_step_scale`` is FRender's (rendering/src/__init__.py:41, 0.2 by default)."""
    with np.load(path, allow_pickle=False) as d:
        depth = int(_scalar(d, "octree_depth"))
        if depth != OCTREE_DEPTH:
            raise ValueError(f"renderData: octree_depth {depth} (the octree operator walks 4^3 blocks, depth 2)")
        nv = int(_scalar(d, "num_voxel"))
        if nv % (1 << depth) != 0:
            raise ValueError(f"renderData: num_voxel {nv} is not a multiple of {1 << depth}")
        vs = np.float64(_scalar(d, "voxel_size"))
        block_nv = nv // (1 << depth) + 2
        if not 2 <= block_nv <= 32:
            raise ValueError(f"renderData: block grid {block_nv}^3 outside [2, 32]^3")
        tile_shape = tuple(int(v) for v in np.asarray(d["tile_shape"]).reshape(-1)[:3])
        nt_grid = int(np.prod(tile_shape))
        tile_index_map = np.ascontiguousarray(d["tile_IndexMap"], np.int32).reshape(-1)
        centers = np.ascontiguousarray(d["centers"], np.float32).reshape(-1, 3)
        group_map = np.ascontiguousarray(d["tile2group"], np.int32).reshape(-1)
        block_index_map = np.ascontiguousarray(d["block_IndexMap"], np.int32).reshape(-1)
        nodes_index_map = np.ascontiguousarray(d["nodes_IndexMap"], np.int16).reshape(-1)
        nodes_sample_flag = np.ascontiguousarray(d["nodes_sampleFlag"], bool).reshape(-1)
        voxels_start = np.ascontiguousarray(d["voxels_start"], np.int64).reshape(-1)
        data_voxels = np.ascontiguousarray(d["data_voxels"], np.float16).reshape(-1, 4)
        net_params = np.ascontiguousarray(d["net_params"], np.float16).reshape(-1)
        group_centers = np.ascontiguousarray(d["group_centers"], np.float32).reshape(-1, 3)
        names = np.asarray(d["group_names"]) if "group_names" in d.files else np.array([], dtype=str)
        scene_min = np.ascontiguousarray(d["scene_min_corner"], np.float32).reshape(-1)[:3]
    nt = centers.shape[0]
    cells = block_nv ** 3
    checks = [
        (tile_index_map.size == nt_grid, f"tile_IndexMap has {tile_index_map.size} entries, tile grid {nt_grid}"),
        (tile_index_map.max(initial=-1) < nt, "tile_IndexMap refers to a tile beyond centers"),
        (group_map.size == nt, f"tile2group has {group_map.size} entries for {nt} tiles"),
        (block_index_map.size == nt * 64, f"block_IndexMap has {block_index_map.size} entries for {nt} tiles x 64"),
        (nodes_index_map.size == voxels_start.size * cells, "nodes_IndexMap is not nodes x block grid"),
        (nodes_sample_flag.size == nodes_index_map.size, "nodes_sampleFlag and nodes_IndexMap differ in size"),
        (block_index_map.max(initial=-1) < voxels_start.size, "block_IndexMap refers to a node beyond voxels_start"),
        (net_params.size == group_centers.shape[0] * REF_MLP_PARAMS,
         f"net_params has {net_params.size} halves for {group_centers.shape[0]} groups x {REF_MLP_PARAMS}"),
        (group_map.max(initial=-1) < group_centers.shape[0], "tile2group refers to a group beyond group_centers"),
    ]
    for ok, msg in checks:
        if not ok:
            raise ValueError("renderData: " + msg)
    tile_size = float(np.float64(nv) * vs)
    osc = OctreeScene(tile_shape=tile_shape, tile_size=tile_size, scene_min=scene_min,
                      tile_index_map=tile_index_map, centers=centers, group_map=group_map,
                      block_index_map=block_index_map, nodes_index_map=nodes_index_map,
                      nodes_sample_flag=nodes_sample_flag, voxels_start=voxels_start, data_voxels=data_voxels,
                      voxel_size=float(vs), sample_step=float(np.float64(sample_step_scale) * vs),
                      num_voxel=block_nv)
    return RenderData(osc, net_params, group_centers, names)


def save_render_data(path, osc: OctreeScene, net_params: np.ndarray, group_centers: np.ndarray,
                     group_names=None) -> None:
    """Writes ``osc`` and the reflection MLPs with the keys and dtypes ``render_prepare`` reads."""
    nv = (osc.num_voxel - 2) << OCTREE_DEPTH
    g = np.asarray(group_centers, np.float32).reshape(-1, 3)
    names = np.asarray(group_names if group_names is not None else [f"group-{i}" for i in range(g.shape[0])],
                       dtype=str)
    np.savez(path,
             octree_depth=np.array([OCTREE_DEPTH], np.int64),
             num_voxel=np.array([nv], np.int64),
             voxel_size=np.array([osc.voxel_size], np.float32),
             scene_min_corner=np.asarray(osc.scene_min, np.float32),
             tile_shape=np.asarray(osc.tile_shape, np.int32),
             net_params=np.asarray(net_params, np.float16).reshape(-1),
             group_centers=g,
             centers=np.asarray(osc.centers, np.float32),
             tile2group=np.asarray(osc.group_map, np.int32),
             tile_IndexMap=np.asarray(osc.tile_index_map, np.int32),
             block_IndexMap=np.asarray(osc.block_index_map, np.int32),
             nodes_IndexMap=np.asarray(osc.nodes_index_map, np.int16),
             nodes_sampleFlag=np.asarray(osc.nodes_sample_flag, bool),
             voxels_start=np.asarray(osc.voxels_start, np.int64),
             data_voxels=np.asarray(osc.data_voxels, np.float16),
             group_names=names)


def pack_reflection_mlp(state_dict) -> np.ndarray:
    """22,556 halves of one group's reflection MLP from its state dict: the six linear layers in
    order (keys ending in ``.weight`` / ``.bias``, e.g. ``mlp.mlp.0.weight`` … ``mlp.mlp.10.bias``
    of the reference's ``MLP``), each packed as bias[out] then weight^T [in][out].  Accepts torch
    tensors or numpy arrays."""
    def arr(v):
        return np.asarray(v.detach().cpu().numpy() if hasattr(v, "detach") else v, np.float32)

    weights = [(k, arr(v)) for k, v in state_dict.items() if k.endswith(".weight")]
    biases = {k[: -len(".bias")]: arr(v) for k, v in state_dict.items() if k.endswith(".bias")}
    if len(weights) != len(REF_MLP_DIMS):
        raise ValueError(f"expected {len(REF_MLP_DIMS)} linear layers, found {len(weights)}")
    out = []
    for (k, w), (i, o) in zip(weights, REF_MLP_DIMS):
        b = biases.get(k[: -len(".weight")])
        if w.shape != (o, i) or b is None or b.shape != (o,):
            raise ValueError(f"layer {k}: expected weight ({o}, {i}) and bias ({o},), got {w.shape}")
        out += [b, w.T.reshape(-1)]
    flat = np.concatenate(out).astype(np.float16)
    assert flat.size == REF_MLP_PARAMS
    return flat


# ---------------------------------------------------------------------------- training checkpoints
# The training side writes, per tile, its dense dilated grid Voxel-T{tile}-E{epoch}.npy (f32
# [nv, nv, nv, 4] in x, y, z, channel order, nv = NUM_VOXEL + 2) and cell flags Node-T{tile}-E{epoch}.npy
# (int [nv, nv, nv], 1 sampled, 0 value only, -1 empty) (training/scripts/tile/tile.py:415-432), and
# per group the reflection MLP's state dict Coeffi-TShared-E{epoch}.pt and its centre group_center.npy
# (training/scripts/tile/tileGroup.py:65-72, 531-552).  The exporter below turns them into the
# renderData.npz octree format: every tile cut into 4^3 blocks of NUM_VOXEL / 4 voxels; a block with a
# sampled voxel becomes a node holding the (NUM_VOXEL / 4 + 2)^3 window of the tile's dilated grid
# that starts at the block (its cells plus one cell of overlap on every side); a node cell holding a
# value (flag >= 0) gets the next index of the node's voxel list and is sampled when its flag is 1.


def octree_from_dense(voxels: np.ndarray, nodes: np.ndarray, tile_cells, tile_shape, tile_size: float, scene_min,
                      groups, sample_step_scale: float = 0.2) -> OctreeScene:
    """OctreeScene of dense training grids: voxels f32[T, nv, nv, nv, 4], nodes int[T, nv, nv, nv] of the
    T tiles at dense tile cells ``tile_cells`` (x*Ty*Tz + y*Tz + z) of ``tile_shape``; groups i32[T]."""
    voxels = np.asarray(voxels, np.float32)
    nodes = np.asarray(nodes)
    T, nv = voxels.shape[0], voxels.shape[1]
    if voxels.shape != (T, nv, nv, nv, 4) or nodes.shape != (T, nv, nv, nv):
        raise ValueError(f"voxels must be [T, nv, nv, nv, 4] and nodes [T, nv, nv, nv], got {voxels.shape} / "
                         f"{nodes.shape}")
    NV = nv - 2
    nb = 1 << OCTREE_DEPTH
    if NV % nb:
        raise ValueError(f"NUM_VOXEL {NV} is not a multiple of {nb}")
    bs = NV // nb
    M = bs + 2
    tile_shape = tuple(int(v) for v in tile_shape)
    cells = np.asarray(tile_cells, np.int64)
    if cells.shape != (T,) or cells.min(initial=0) < 0 or cells.max(initial=-1) >= int(np.prod(tile_shape)):
        raise ValueError("tile_cells must hold one dense tile cell per tile")
    tim = np.full(int(np.prod(tile_shape)), -1, np.int32)
    tim[cells] = np.arange(T, dtype=np.int32)
    smin = np.asarray(scene_min, np.float32).reshape(3)
    ts = np.float32(tile_size)
    ix, iy, iz = cells // (tile_shape[1] * tile_shape[2]), (cells // tile_shape[2]) % tile_shape[1], cells % tile_shape[2]
    centers = (smin + (np.stack([ix, iy, iz], 1).astype(np.float32) + np.float32(0.5)) * ts).astype(np.float32)
    bmap = np.full(T * nb ** 3, -1, np.int32)
    n_idx, n_flag, starts, data = [], [], [], []
    count = 0
    for t in range(T):
        f = nodes[t]
        v = voxels[t]
        for b in range(nb ** 3):
            bx, by, bz = b // (nb * nb), (b // nb) % nb, b % nb
            inner = f[1 + bs * bx:1 + bs * (bx + 1), 1 + bs * by:1 + bs * (by + 1), 1 + bs * bz:1 + bs * (bz + 1)]
            if not (inner == 1).any():
                continue
            win = (slice(bs * bx, bs * bx + M), slice(bs * by, bs * by + M), slice(bs * bz, bs * bz + M))
            wf = f[win].reshape(-1)
            has = wf >= 0
            k = int(has.sum())
            if k > 32767:
                raise ValueError("a node holds more voxels than an int16 index reaches")
            idx = np.full(M ** 3, -1, np.int16)
            idx[has] = np.arange(k, dtype=np.int16)
            bmap[t * nb ** 3 + b] = len(starts)
            n_idx.append(idx)
            n_flag.append(wf == 1)
            starts.append(count)
            data.append(v[win].reshape(-1, 4)[has])
            count += k
    data_v = (np.concatenate(data) if data else np.zeros((1, 4), np.float32)).astype(np.float16)
    if data_v.shape[0] == 0:
        data_v = np.zeros((1, 4), np.float16)
    vs = float(np.float64(tile_size) / NV)
    return OctreeScene(tile_shape=tile_shape, tile_size=float(tile_size), scene_min=smin, tile_index_map=tim,
                       centers=centers, group_map=np.asarray(groups, np.int32).reshape(T),
                       block_index_map=bmap,
                       nodes_index_map=np.concatenate(n_idx) if n_idx else np.full(M ** 3, -1, np.int16),
                       nodes_sample_flag=np.concatenate(n_flag) if n_flag else np.zeros(M ** 3, bool),
                       voxels_start=np.asarray(starts, np.int64) if starts else np.zeros(1, np.int64),
                       data_voxels=data_v, voxel_size=vs, sample_step=float(np.float64(sample_step_scale) * vs),
                       num_voxel=M)


def export_checkpoint(out_path, groups, tile_shape, tile_size: float, scene_min, epoch: int,
                      sample_step_scale: float = 0.2) -> RenderData:
    """Writes renderData.npz from training outputs.  ``groups``: one dict per group with ``name``,
    ``logdir`` (holding Coeffi-TShared-E{epoch}.pt and optionally group_center.npy), ``tile_dir``
    (holding the Voxel-T / Node-T files) and ``tiles``: a list of (file tile index, dense tile cell).
    The state dict is read with torch.load(weights_only=True); without group_center.npy the centre is
    the mean of the group's tile centres (tileGroup.py:65)."""
    import os

    import torch
    vox, nod, cells, gidx, params, gcent, names = [], [], [], [], [], [], []
    smin = np.asarray(scene_min, np.float32).reshape(3)
    for g, grp in enumerate(groups):
        tc = []
        for tile_idx, cell in grp["tiles"]:
            v = np.load(os.path.join(grp["tile_dir"], f"Voxel-T{tile_idx}-E{epoch}.npy"), allow_pickle=False)
            f = np.load(os.path.join(grp["tile_dir"], f"Node-T{tile_idx}-E{epoch}.npy"), allow_pickle=False)
            vox.append(np.asarray(v, np.float32))
            nod.append(np.asarray(f).astype(np.int16))
            cells.append(int(cell))
            gidx.append(g)
            c = int(cell)
            ijk = np.array([c // (tile_shape[1] * tile_shape[2]), (c // tile_shape[2]) % tile_shape[1],
                            c % tile_shape[2]], np.float32)
            tc.append(smin + (ijk + np.float32(0.5)) * np.float32(tile_size))
        sd = torch.load(os.path.join(grp["logdir"], f"Coeffi-TShared-E{epoch}.pt"), map_location="cpu",
                        weights_only=True)
        params.append(pack_reflection_mlp(sd))
        cpath = os.path.join(grp["logdir"], "group_center.npy")
        gcent.append(np.load(cpath, allow_pickle=False).reshape(3) if os.path.exists(cpath) else np.mean(tc, axis=0))
        names.append(grp.get("name", f"group-{g}"))
    order = np.argsort(cells, kind="stable")          # tiles in dense cell order
    osc = octree_from_dense(np.stack(vox)[order], np.stack(nod)[order], np.asarray(cells)[order], tile_shape,
                            tile_size, scene_min, np.asarray(gidx, np.int32)[order], sample_step_scale)
    rd = RenderData(osc, np.concatenate(params).astype(np.float16), np.asarray(gcent, np.float32),
                    np.asarray(names, dtype=str))
    save_render_data(out_path, osc, rd.net_params, rd.group_centers, rd.group_names)
    return rd
