"""CPU checks of the C ABI: the library loads, exports every entry point declared in
include/snisr.h, the ctypes struct layouts match, and argument validation rejects bad
descriptors before any device work (no compute calls here)."""
import ctypes as C
import os
import re

import pytest

from snisr import _native as N

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    names = set()
    for f in os.listdir(os.path.join(REPO, "include")):
        if f.endswith(".h"):
            txt = open(os.path.join(REPO, "include", f)).read()
            txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
            names |= set(re.findall(r"\b(snisr_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = N.lib()
    names = declared_entry_points()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    bound = {n for n, _, _ in N.SIGNATURES}
    assert set(declared_entry_points()) <= bound


def test_struct_layouts_match():
    L = N.lib()
    for which, st in enumerate((N.SceneDesc, N.MarchParams, N.RaysDesc, N.RenderOut)):
        assert L.snisr_abi_sizeof(which) == C.sizeof(st)
    assert L.snisr_version() >= 1


def test_null_and_inconsistent_arguments_are_rejected():
    L = N.lib()
    p = N.MarchParams(max_tiles=8, sample_step=0.01, early_stop=1e-3, trans_th=0.8)
    r = N.RaysDesc()
    o = N.RenderOut()
    rc = L.snisr_render_fp32(None, C.byref(p), C.byref(r), C.byref(o), None)
    assert rc == -1 and b"scene" in L.snisr_last_error()
    s = N.SceneDesc()
    s.tile_shape[:] = [1, 1, 1]
    s.tile_size = 1.0
    s.n_tiles = 1
    rc = L.snisr_render_fp32(C.byref(s), C.byref(p), C.byref(r), C.byref(o), None)
    assert rc == -1  # tile tables are NULL
    # a non-NULL but unaligned feature pointer (never dereferenced: validation fails first)
    fake = C.c_void_p(16)
    for name in ("tile_index_map", "tile_centers", "tile_res", "tile_brick_base", "brick_map", "sigma",
                 "mlp_params"):
        setattr(s, name, fake)
    s.n_bricks = 1
    s.feat_dim = 16
    s.mlp_in = 31
    s.mlp_layers = 3
    s.mlp_width = 64
    s.feats = C.c_void_p(18)
    o.rgb = C.c_void_p(16)
    r.n = 1
    r.origins = fake
    r.dirs = fake
    rc = L.snisr_render_fp32(C.byref(s), C.byref(p), C.byref(r), C.byref(o), None)
    assert rc == -3, L.snisr_last_error()
    s.feats = fake
    s.mlp_in = 30
    rc = L.snisr_render_fp32(C.byref(s), C.byref(p), C.byref(r), C.byref(o), None)
    assert rc == -1 and b"mlp_in" in L.snisr_last_error()
    s.mlp_in = 31
    bad = N.MarchParams(max_tiles=8, sample_step=0.0, early_stop=1e-3, trans_th=0.8)
    rc = L.snisr_render_fp32(C.byref(s), C.byref(bad), C.byref(r), C.byref(o), None)
    assert rc == -1 and b"sample_step" in L.snisr_last_error()
    r.origin_stride = 5
    rc = L.snisr_render_fp32(C.byref(s), C.byref(p), C.byref(r), C.byref(o), None)
    assert rc == -1 and b"origin_stride" in L.snisr_last_error()


def test_compute_rays_rejects_null():
    L = N.lib()
    assert L.snisr_compute_rays(None, None, 4, 4, None, None) == -1


def test_python_layer_fails_loudly_without_device_tensors():
    torch = pytest.importorskip("torch")
    from snisr import device as D
    with pytest.raises(ValueError):
        D.compute_rays(torch.zeros(3), torch.zeros(9), 2, 2)


def test_unet_entry_points_validate_before_device_work():
    L = N.lib()
    assert L.snisr_unet_packed_size() > 0
    assert L.snisr_unet_workspace_size(720, 1280) > L.snisr_unet_workspace_size(64, 96) > 0
    assert L.snisr_unet_workspace_size(7, 8) == -1 and b"multiples of 4" in L.snisr_last_error()
    rc = L.snisr_unet_refine(None, None, None, None, 8, 8, None, 0, None, None)
    assert rc != 0 and b"NULL" in L.snisr_last_error()
    rc = L.snisr_unet_refine(None, None, None, None, 6, 8, None, 0, None, None)
    assert rc != 0 and b"multiples of 4" in L.snisr_last_error()
    nul = (C.c_void_p * 10)()
    assert L.snisr_unet_pack(nul, nul, C.c_void_p(256), None) != 0


def test_regroup_refuses_too_few_blocks_before_device_work():
    """snisr_regroup_pixels checks max_blocks against snisr_regroup_max_blocks(n, P) before any launch
    (fewer output blocks would merge the last groups' pixels into one block)."""
    L = N.lib()
    n, P = 1000, 64
    mb = L.snisr_regroup_max_blocks(n, P)
    assert mb == -(-n // P) + n
    ws = L.snisr_regroup_workspace_size(n)
    assert ws > 0
    fake = C.c_void_p(256)
    rc = L.snisr_regroup_pixels(fake, n, P, mb - 1, fake, fake, fake, fake, fake, ws, None)
    assert rc == -5 and b"max_blocks" in L.snisr_last_error()
    rc = L.snisr_regroup_pixels(fake, n, P, mb, fake, fake, fake, fake, fake, ws - 1, None)
    assert rc == -5 and b"workspace" in L.snisr_last_error()
    assert L.snisr_sort_by_key_workspace_size(-1) == -1
