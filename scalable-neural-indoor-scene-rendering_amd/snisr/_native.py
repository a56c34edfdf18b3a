"""ctypes binding of the C ABI in include/snisr.h (libsnisr_hip.so).

Loading fails loudly when the library is missing: there is no CPU fallback behind this
module.  Struct layouts are checked against ``snisr_abi_sizeof`` at load time.
"""
from __future__ import annotations

import ctypes as C
import os

from .build import LIB_PATH

SNISR_F32, SNISR_F16, SNISR_BF16 = 0, 1, 2


class SceneDesc(C.Structure):
    _fields_ = [
        ("tile_shape", C.c_int32 * 3), ("scene_min", C.c_float * 3), ("tile_size", C.c_float),
        ("n_tiles", C.c_int32), ("n_bricks", C.c_int32), ("feat_dim", C.c_int32), ("feat_dtype", C.c_int32),
        ("mlp_in", C.c_int32), ("mlp_width", C.c_int32), ("mlp_layers", C.c_int32),
        ("mlp_packed_dtype", C.c_int32), ("mlp_stride", C.c_int64), ("brick_map_len", C.c_int64),
        ("tile_index_map", C.c_void_p), ("tile_centers", C.c_void_p), ("tile_res", C.c_void_p),
        ("tile_brick_base", C.c_void_p), ("brick_map", C.c_void_p), ("sigma", C.c_void_p),
        ("feats", C.c_void_p), ("mlp_params", C.c_void_p), ("mlp_f32t", C.c_void_p), ("mlp_packed", C.c_void_p),
    ]


class MarchParams(C.Structure):
    _fields_ = [("max_tiles", C.c_int32), ("sample_step", C.c_float), ("early_stop", C.c_float),
                ("trans_th", C.c_float)]


class RaysDesc(C.Structure):
    _fields_ = [("origins", C.c_void_p), ("origin_stride", C.c_int32), ("image_width", C.c_int32),
                ("n", C.c_int64), ("dirs", C.c_void_p), ("t_hit", C.c_void_p), ("visited", C.c_void_p)]


class RenderOut(C.Structure):
    _fields_ = [("rgb", C.c_void_p), ("T", C.c_void_p), ("n_samples", C.c_void_p), ("depth", C.c_void_p),
                ("bricks", C.c_void_p), ("brick_cap", C.c_int32), ("n_stage_events", C.c_int32),
                ("stage_events", C.c_void_p)]


class MlpQuery(C.Structure):
    _fields_ = [("samples_per_ray", C.c_int32), ("n_groups", C.c_int32), ("n_blocks", C.c_int64),
                ("inverse_far", C.c_float), ("trans_stop", C.c_float), ("net_idx", C.c_void_p),
                ("pixel_starts", C.c_void_p), ("query_pixel_indices", C.c_void_p), ("n_query", C.c_int64),
                ("transparency", C.c_void_p), ("group_centers", C.c_void_p), ("samples", C.c_void_p),
                ("rays_d", C.c_void_p), ("d_row_stride", C.c_int64), ("d_col_stride", C.c_int64),
                ("H", C.c_int32), ("W", C.c_int32), ("packed_params", C.c_void_p), ("sample_results", C.c_void_p),
                ("n_blocks_device", C.c_void_p)]


class OctreeDesc(C.Structure):
    _fields_ = [("group_map", C.c_void_p), ("data_voxels", C.c_void_p), ("block_index_map", C.c_void_p),
                ("nodes_index_map", C.c_void_p), ("nodes_sample_flag", C.c_void_p), ("voxels_start", C.c_void_p),
                ("centers", C.c_void_p), ("tile_size", C.c_float), ("voxel_size", C.c_float),
                ("sample_step", C.c_float), ("num_voxel", C.c_int32)]


class DenseGridDesc(C.Structure):
    _fields_ = [("voxels", C.c_void_p), ("nodes", C.c_void_p), ("sparse_to_group", C.c_void_p),
                ("centers", C.c_void_p), ("tile_size", C.c_float), ("voxel_size", C.c_float),
                ("sample_step", C.c_float), ("num_voxel", C.c_int32)]


class TileMeshDesc(C.Structure):
    _fields_ = [("vertices", C.c_void_p), ("faces", C.c_void_p), ("n_faces", C.c_int64), ("centers", C.c_void_p),
                ("face_idx", C.c_void_p), ("n_face_idx", C.c_int64), ("face_num", C.c_void_p),
                ("tile_size", C.c_float), ("num_block", C.c_int32)]


class GraphDesc(C.Structure):
    _fields_ = [("H", C.c_int32), ("W", C.c_int32), ("mlp_dtype", C.c_int32), ("bvh_n_nodes", C.c_int32),
                ("bvh_nodes", C.c_void_p), ("bvh_tris", C.c_void_p), ("bvh_n_tris", C.c_int64),
                ("bvh_depth", C.c_int32), ("records_per_ray", C.c_int32), ("sample_capacity", C.c_int64),
                ("sample_headroom", C.c_float), ("params", MarchParams)]


class Tuning(C.Structure):
    _fields_ = [("ray_order", C.c_int32), ("mlp_threads", C.c_int32), ("sort_grid", C.c_int32),
                ("records_per_ray", C.c_int32), ("chunks", C.c_int32)]


_lib = None

# (name, restype, argtypes) of every exported entry point declared in include/snisr.h
SIGNATURES = [
    ("snisr_last_error", C.c_char_p, []),
    ("snisr_version", C.c_int, []),
    ("snisr_abi_sizeof", C.c_int64, [C.c_int32]),
    ("snisr_compute_rays", C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p]),
    ("snisr_trace_tiles", C.c_int, [C.POINTER(SceneDesc), C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                    C.c_void_p, C.c_void_p]),
    ("snisr_render_fp32", C.c_int, [C.POINTER(SceneDesc), C.POINTER(MarchParams), C.POINTER(RaysDesc),
                                    C.POINTER(RenderOut), C.c_void_p]),
    ("snisr_render_mfma_workspace_size", C.c_int64, [C.POINTER(SceneDesc), C.c_int64, C.c_int64, C.c_int64]),
    ("snisr_render_mfma_workspace_size_tuned", C.c_int64, [C.POINTER(SceneDesc), C.c_int64, C.POINTER(Tuning),
                                                           C.c_int64]),
    ("snisr_render_mfma", C.c_int, [C.POINTER(SceneDesc), C.POINTER(MarchParams), C.POINTER(RaysDesc),
                                    C.POINTER(RenderOut), C.c_void_p, C.c_int64, C.c_void_p]),
    ("snisr_render_mfma_tuned", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(Tuning),
                                          C.c_void_p, C.c_int64, C.c_void_p]),
    ("snisr_render_mfma_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                          C.POINTER(C.c_int32), C.c_void_p]),
    ("snisr_bvh_build_bounds", C.c_int, [C.c_int64, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    ("snisr_bvh_build", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p,
                                  C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    ("snisr_mesh_first_hit", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int32,
                                       C.POINTER(RaysDesc), C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_inverse_z_sampling", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_int32, C.c_int32,
                                           C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    ("snisr_integrate_points", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    ("snisr_sort_by_key_workspace_size", C.c_int64, [C.c_int64]),
    ("snisr_sort_by_key", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                    C.POINTER(C.c_int32), C.c_void_p]),
    ("snisr_sample_sparse", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_float, C.c_float,
                                      C.c_float, C.c_float, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                      C.c_void_p]),
    ("snisr_sample_reflection", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_float,
                                          C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_float,
                                          C.c_int32, C.c_void_p, C.c_void_p]),
    ("snisr_sample_uniform", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.c_int32, C.c_void_p, C.c_void_p]),
    ("snisr_sample_bg", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_float, C.c_float,
                                  C.c_int32, C.c_void_p, C.c_void_p]),
    ("snisr_compute_group_boundary", C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_void_p,
                                               C.c_void_p]),
    ("snisr_regroup_max_blocks", C.c_int64, [C.c_int64, C.c_int32]),
    ("snisr_regroup_workspace_size", C.c_int64, [C.c_int64]),
    ("snisr_regroup_pixels", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("snisr_regroup_pixels_static_max", C.c_int64, []),
    ("snisr_regroup_pixels_static", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                              C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_adam_step", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_int64, C.c_void_p]),
    ("snisr_unet_packed_size", C.c_int64, []),
    ("snisr_unet_pack", C.c_int, [C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.c_void_p, C.c_void_p]),
    ("snisr_unet_workspace_size", C.c_int64, [C.c_int32, C.c_int32]),
    ("snisr_unet_refine", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_int32,
                                    C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("snisr_assign_pixels", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p,
                                      C.c_void_p, C.c_void_p]),
    ("snisr_render_octree", C.c_int, [C.POINTER(OctreeDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                      C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_int32, C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_inference_mlp_packed_size", C.c_int64, [C.c_int64]),
    ("snisr_pack_inference_mlp", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("snisr_inference_mlp", C.c_int, [C.POINTER(MlpQuery), C.c_void_p]),
    ("snisr_train_workspace_size", C.c_int64, [C.POINTER(SceneDesc), C.c_int64, C.c_int64]),
    ("snisr_train_forward", C.c_int, [C.POINTER(SceneDesc), C.POINTER(MarchParams), C.POINTER(RaysDesc),
                                      C.POINTER(RenderOut), C.c_void_p, C.c_int64, C.c_void_p]),
    ("snisr_train_backward", C.c_int, [C.POINTER(SceneDesc), C.POINTER(MarchParams), C.POINTER(RaysDesc),
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                       C.c_void_p, C.c_int32, C.c_void_p]),
    ("snisr_render_dense", C.c_int, [C.POINTER(DenseGridDesc), C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                     C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_mesh_hit_tiles", C.c_int, [C.POINTER(TileMeshDesc), C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                       C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_tile_face_lists", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int32,
                                        C.c_float, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64,
                                        C.POINTER(C.c_int64)]),
    ("snisr_scene_create", C.c_int, [C.POINTER(SceneDesc), C.c_int32, C.POINTER(C.c_void_p)]),
    ("snisr_scene_destroy", C.c_int, [C.c_void_p]),
    ("snisr_scene_get_desc", C.POINTER(SceneDesc), [C.c_void_p]),
    ("snisr_graph_create", C.c_int, [C.c_void_p, C.POINTER(GraphDesc), C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.POINTER(C.c_void_p)]),
    ("snisr_graph_render", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_graph_outputs", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    ("snisr_graph_stats", C.c_int, [C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                    C.POINTER(C.c_int64)]),
    ("snisr_graph_synchronize", C.c_int, [C.c_void_p]),
    ("snisr_graph_destroy", C.c_int, [C.c_void_p]),
    ("snisr_volume_render_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                              C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_volume_render_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_int32,
                                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                               C.c_void_p]),
    ("snisr_sh_sigmoid_forward", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                           C.c_void_p, C.c_void_p]),
    ("snisr_sh_sigmoid_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p,
                                            C.c_void_p, C.c_int32, C.c_void_p]),
    ("snisr_pemlp_param_count", C.c_int64, [C.c_int32, C.c_int32, C.c_int32]),
    ("snisr_pemlp_workspace_size", C.c_int64, [C.c_int64, C.c_int32, C.c_int32, C.c_int32]),
    ("snisr_pemlp_forward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                      C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("snisr_pemlp_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_int32, C.c_float,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("snisr_flag_cells_workspace_size", C.c_int64, [C.c_int64]),
    ("snisr_flag_cells", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p]),
    ("snisr_bake_points", C.c_int, [C.c_void_p, C.c_int64, C.c_int32] + [C.c_float] * 8 + [C.c_void_p, C.c_void_p]),
    ("snisr_normalize_points", C.c_int, [C.c_void_p, C.c_int64] + [C.c_float] * 4 + [C.c_void_p, C.c_void_p]),
    ("snisr_diffuse_head_forward", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p]),
    ("snisr_diffuse_head_backward", C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                                              C.c_void_p]),
    ("snisr_voxel_query_forward", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_int64] + [C.c_float] * 4
     + [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_voxel_query_backward", C.c_int, [C.c_int32, C.c_void_p, C.c_int64] + [C.c_float] * 4
     + [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("snisr_pack_mlp_size", C.c_int64, [C.POINTER(SceneDesc), C.c_int32]),
    ("snisr_pack_mlp", C.c_int, [C.POINTER(SceneDesc), C.c_int32, C.c_void_p, C.c_void_p]),
]


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"snisr: HIP library not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    for which, st in enumerate((SceneDesc, MarchParams, RaysDesc, RenderOut, MlpQuery, OctreeDesc, Tuning,
                                   DenseGridDesc, TileMeshDesc, GraphDesc)):
        got = L.snisr_abi_sizeof(which)
        if got != C.sizeof(st):
            raise RuntimeError(f"snisr: ABI mismatch for {st.__name__}: C {got} vs ctypes {C.sizeof(st)}")
    _lib = L
    return L


SNISR_OK, SNISR_ERR_ARG, SNISR_ERR_SHAPE, SNISR_ERR_ALIGN, SNISR_ERR_DEVICE, SNISR_ERR_CAPACITY, \
    SNISR_ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6
STATUS_NAMES = {-1: "SNISR_ERR_ARG", -2: "SNISR_ERR_SHAPE", -3: "SNISR_ERR_ALIGN", -4: "SNISR_ERR_DEVICE",
                -5: "SNISR_ERR_CAPACITY", -6: "SNISR_ERR_UNSUPPORTED"}


class SnisrError(RuntimeError):
    """A failed C-ABI call; ``status`` is the snisr_status code."""

    def __init__(self, what: str, status: int, msg: str):
        super().__init__(f"{what} failed ({STATUS_NAMES.get(status, status)}): {msg}")
        self.status = status


def check(rc: int, what: str):
    if rc != 0:
        raise SnisrError(what, rc, lib().snisr_last_error().decode())
