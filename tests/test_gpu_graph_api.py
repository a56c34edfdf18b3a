"""The native scene handle and captured frame driver of the C ABI (snisr_scene_* / snisr_graph_*):
a scene uploaded by the library from host arrays renders, through one HIP graph per frame driver,
exactly what the torch-side path renders (FP16 MFMA with the BVH bound and the plain FP32 path),
across camera changes between replays (FRender.render_prepare / render, rendering/src/__init__.py:
98-341, for the feature scenes)."""
import numpy as np
import pytest
import torch

from snisr import _native as N
from snisr import device as D
from snisr import graph as G
from snisr import scene as S

pytestmark = pytest.mark.gpu


def test_native_graph_matches_torch_path():
    print("native graph: C2 room 360x640, 20k triangles", flush=True)
    sc, cam, mesh = S.config_c2(H=360, W=640, n_triangles=20_000)
    cam.intrinsic[2] *= 2.0
    cams = [cam, S.room_camera(7, (0, 0, 0), (4, 4, 1), 360, 640, focal=375.0)]
    ns = D.NativeScene(sc, N.SNISR_F16)
    dm = D.DeviceMesh(mesh.vertices, mesh.faces)
    ng = D.NativeGraph(ns, cams[0], N.SNISR_F16, mesh=dm, records_per_ray=96, sample_capacity=40 * 360 * 640)
    ds = D.DeviceScene(sc, mfma_dtype=N.SNISR_F16)
    for rep in range(2):
        for c in cams:
            ng.render(c)
            out = {k: v.clone() for k, v in ng.outputs().items()}
            st = ng.stats()
            assert st["overflow"] == 0, st
            ref = G.FrameGraph(ds, dm, [c]).replay()
            torch.cuda.synchronize()
            for k in ("rgb", "T", "n_samples", "depth"):
                assert torch.equal(out[k], ref[k]), (rep, k)
            assert torch.isfinite(out["rgb"]).all()
    # the plain FP32 path through its own graph
    fg32 = D.NativeGraph(ns, cams[1], N.SNISR_F32, mesh=dm)
    fg32.render()
    o32 = fg32.outputs()
    r32 = D.render_fp32(ds, torch.from_numpy(cams[1].origin).cuda(),
                        D.compute_rays(torch.from_numpy(cams[1].intrinsic).cuda(), torch.from_numpy(cams[1].c2w).cuda(),
                                       360, 640), t_hit=o32["t_hit"].clone(), image_width=640)
    for k in ("rgb", "T", "n_samples", "depth"):
        assert torch.equal(o32[k], r32[k]), k
    fg32.close()
    ng.close()
    ns.close()


def test_native_scene_rejects_bad_arguments():
    L = N.lib()
    import ctypes as C
    h = C.c_void_p()
    assert L.snisr_scene_create(None, N.SNISR_F16, C.byref(h)) == -1
    sc, cam, _ = S.config_c1()
    ns = D.NativeScene(sc, N.SNISR_F32)
    g = N.GraphDesc()
    g.H, g.W, g.mlp_dtype = 4, 4, N.SNISR_F16          # no MFMA packing in this scene
    g.params.max_tiles, g.params.sample_step, g.params.early_stop = 4, 0.01, 1e-3
    out = C.c_void_p()
    z = np.zeros(16, np.float32)
    p = z.ctypes.data_as(C.c_void_p)
    assert L.snisr_graph_create(ns.handle, C.byref(g), p, p, p, C.byref(out)) == -1
    assert b"MFMA" in L.snisr_last_error()
    ns.close()
