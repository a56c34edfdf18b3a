"""Pins of the oracle's restatement of the reference's octree-format renderer
(rendering_diffuse_octree_fp16, rendering/src/raydiffuse_render_octree_fp16.cu) by closed forms on
single-tile scenes whose every cell is occupied and sampled:

* a linear colour field with an opaque density: the pixel is the field at the first sample, the
  tile entry point (rays whose first interval is too short to be opaque are left out), because trilinear interpolation reproduces linear fields exactly (any index,
  axis, half-voxel or border mistake moves it by >= 1e-3);
* a constant field: T = exp(-sigma (t_exit - t_entry)) and C = c (1 - T), because the intervals of
  the march tile every block's [near, far) without gap or overlap (border cells flagged too: a
  sample that rounds onto an unflagged border cell is skipped with its interval, by design);
* the first-hit distance: the first sample with T < trans_th lies within one step after the
  continuous crossing t* = t_entry + ln(1 / trans_th) / sigma."""
import math

import numpy as np
import pytest

from snisr import scene as S


def _rays(seed, n, origin_lo=-1.5, origin_hi=2.5):
    rng = np.random.default_rng(seed)
    o = np.array([rng.uniform(origin_lo, -0.2), rng.uniform(0.1, 0.9), rng.uniform(0.1, 0.9)], np.float32)
    tgt = rng.uniform(0.05, 0.95, (n, 3))
    d = (tgt - o).astype(np.float32)
    d *= rng.uniform(0.5, 2.0, (n, 1)).astype(np.float32)   # unnormalised, as compute_rays gives
    return o, d


def _entry_exit(o, d):
    """t range of the unit cube [0,1]^3 in float64."""
    o = o.astype(np.float64)
    d = d.astype(np.float64)
    with np.errstate(divide="ignore"):
        t0 = (0.0 - o) / d
        t1 = (1.0 - o) / d
    lo = np.maximum(np.minimum(t0, t1).max(1), 0.0)
    hi = np.maximum(t0, t1).min(1)
    return lo, hi


def test_linear_field_opaque_returns_entry_point(oracle_mod):
    def field(t, ijk):
        v = np.empty(ijk.shape[:-1] + (4,), np.float32)
        v[..., :3] = (ijk + 0.5) / 128.0     # (2i+1)/256: exact in FP16
        v[..., 3] = 60000.0
        return v
    osc = S.make_octree_scene(3, tile_shape=(1, 1, 1), block_occupancy=1.0, field=field)
    o, d = _rays(1, 3000)
    vis = np.zeros((d.shape[0], 1), np.int32)
    frame, inv, net, ns = oracle_mod.render_octree(osc, o, d, vis, trans_th=0.5, early_stop=0.0)
    lo, hi = _entry_exit(o, d)
    p = o.astype(np.float64) + lo[:, None] * d.astype(np.float64)
    # the first sample is opaque only if its interval (entry to the first block boundary) exceeds
    # 88 / sigma, where exp(-sigma * interval) underflows to 0 in FP32
    d64 = d.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        nxt = np.where(d64 >= 0, (np.floor(p / 0.25) + 1) * 0.25, np.ceil(p / 0.25 - 1) * 0.25)
        first = np.where(d64 != 0, (nxt - p) / d64, np.inf).min(1)
    hit = (hi > lo + 1e-3) & (first > 100.0 / 60000.0)
    assert hit.sum() > 2000
    expect = p / 2.0    # field value (x - corner) / (128 voxel) with 64 voxels per unit
    err = np.abs(frame[hit] - expect[hit]).max()
    assert err < 2e-5, err
    miss = hi < lo - 1e-6
    assert (ns[hit] >= 1).all() and (frame[miss] == 0).all() and (ns[miss] == 0).all()


def test_constant_field_closed_form(oracle_mod):
    sig, c = 2.0, np.array([0.25, 0.5, 0.75], np.float32)
    def field(t, ijk):
        v = np.empty(ijk.shape[:-1] + (4,), np.float32)
        v[..., :3] = c
        v[..., 3] = sig
        return v
    osc = S.make_octree_scene(4, tile_shape=(1, 1, 1), block_occupancy=1.0, field=field, flag_border=True)
    o, d = _rays(2, 3000)
    vis = np.zeros((d.shape[0], 1), np.int32)
    frame, inv, net, ns = oracle_mod.render_octree(osc, o, d, vis, trans_th=0.0, early_stop=0.0)
    lo, hi = _entry_exit(o, d)
    T = np.exp(-sig * np.maximum(hi - lo, 0.0))
    assert np.abs(frame - (1.0 - T)[:, None] * c).max() < 2e-5
    # samples: every step of every block's [near, far) -> ceil per block; total at least the span
    assert (ns >= np.floor(np.maximum(hi - lo, 0) / osc.sample_step) - 1).all()


def test_first_hit_within_one_step_of_continuous_crossing(oracle_mod):
    sig, th = 40.0, 0.8
    def field(t, ijk):
        v = np.zeros(ijk.shape[:-1] + (4,), np.float32)
        v[..., :3] = 0.5
        v[..., 3] = sig
        return v
    osc = S.make_octree_scene(5, tile_shape=(1, 1, 1), block_occupancy=1.0, n_groups=1, field=field,
                               flag_border=True)
    o, d = _rays(3, 2000)
    vis = np.zeros((d.shape[0], 1), np.int32)
    frame, inv, net, ns = oracle_mod.render_octree(osc, o, d, vis, trans_th=th, early_stop=0.0)
    lo, hi = _entry_exit(o, d)
    zstar = math.log(1.0 / th) / sig
    ok = hi - lo > zstar + 2 * osc.sample_step
    assert ok.sum() > 1500
    t = 1.0 / inv[ok].astype(np.float64)
    assert (t >= lo[ok] + zstar - 1e-5).all() and (t <= lo[ok] + zstar + osc.sample_step + 1e-5).all()
    assert (net[ok] == osc.group_map[0]).all()
    # rays that never reach T < th keep the caller's values
    never = hi - lo < zstar - 1e-3
    assert (inv[never] == np.float32(1e-8)).all() and (net[never] == -1).all()


def test_soft_hit_uses_last_tile_exit(oracle_mod):
    osc = S.make_octree_scene(6, tile_shape=(1, 1, 1), block_occupancy=0.0)   # empty tile
    o, d = _rays(4, 200)
    vis = np.zeros((d.shape[0], 1), np.int32)
    frame, inv, net, ns = oracle_mod.render_octree(osc, o, d, vis, soft_hit=True)
    lo, hi = _entry_exit(o, d)
    hit = hi > lo
    assert np.allclose(1.0 / inv[hit], hi[hit], rtol=1e-5)
    assert (net[hit] == osc.group_map[0]).all() and (ns == 0).all()
