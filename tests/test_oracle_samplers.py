"""Pins of the oracle's restatement of the reference's training samplers (training/src/Sample_kernel.cu)
against float64 brute force and library routines: the slab closed form for the tile chord
(uniform, bg), voxel membership of every sparse sample, the boundary samples around the flagged
runs, and for the reflection sampler the transmittance through flagged voxels with densities from
scipy's linear map_coordinates (the overlap trilinear of interpolation.h:43-134 with clamped edges)."""
import numpy as np
import pytest
from scipy import ndimage

NV, TS = 12, 1.0
CTR = np.array([0.5, 0.5, 0.5], np.float32)


def _rays(seed, n):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-1.0, 2.0, (n, 3)).astype(np.float32)
    tgt = rng.uniform(0.15, 0.85, (n, 3)).astype(np.float32)
    return o, (tgt - o).astype(np.float32)


def _slab64(o, d, lo, hi):
    o, d = o.astype(np.float64), d.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        t1, t2 = (lo - o) / d, (hi - o) / d
    tn = np.max(np.minimum(t1, t2), -1)
    tf = np.min(np.maximum(t1, t2), -1)
    return np.maximum(tn, 0.0), tf


def test_uniform_closed_form(oracle_mod):
    o, d = _rays(1, 500)
    z = np.zeros((500, 9), np.float32)
    oracle_mod.sample_uniform(o, d, CTR, TS, z)
    tn, tf = _slab64(o, d, CTR - 0.5, CTR + 0.5)
    ref = tn[:, None] + (tf - tn)[:, None] * np.arange(9) / 8.0
    assert np.abs(z - ref).max() < 2e-5


def test_bg_rule(oracle_mod):
    o, d = _rays(2, 400)
    rng = np.random.default_rng(3)
    bg = rng.uniform(0.0, 4.0, (400, 1)).astype(np.float32)
    bg[::7] = -1.0
    z = np.full((400, 6), -7.0, np.float32)
    oracle_mod.sample_bg(o, d, bg, CTR, TS, np.float32(0.5), z)
    _, tf = _slab64(o, d, CTR - 0.5, CTR + 0.5)
    for r in range(400):
        if bg[r, 0] == -1.0:
            assert (z[r] == -7.0).all()
            continue
        near = bg[r, 0] - 0.25
        if near <= tf[r]:
            near = tf[r] + 1e-5
        ref = near + 0.5 * np.arange(6) / 5.0
        assert np.abs(z[r] - ref).max() < 3e-5, r


def _flags(seed, p=0.15):
    return (np.random.default_rng(seed).random(NV ** 3) < p).astype(np.int16)


def _voxel_of(p):
    return np.floor((p - (CTR - 0.5)) / (TS / NV)).astype(int)


def test_sparse_samples_lie_in_flagged_voxels(oracle_mod):
    o, d = _rays(4, 300)
    fl = _flags(5)
    N = 40
    z = np.full((300, N), -7.0, np.float32)
    dist = np.full((300, N), -7.0, np.float32)
    oracle_mod.sample_sparse(o, d, fl, CTR, TS, NV, z, dist)
    vs = TS / NV
    # flagged voxel chords by brute force
    ijk = np.stack(np.meshgrid(*[np.arange(NV)] * 3, indexing="ij"), -1).reshape(-1, 3)
    lo = (CTR - 0.5) + ijk * vs
    checked = 0
    for r in range(300):
        tn, tf = _slab64(np.repeat(o[r:r + 1], ijk.shape[0], 0), np.repeat(d[r:r + 1], ijk.shape[0], 0), lo, lo + vs)
        hit = (tn < tf - 1e-6) & (fl == 1)
        if not hit.any():
            assert (z[r] == -7.0).all(), r
            continue
        checked += 1
        t0, t1 = tn[hit].min(), tf[hit].max()
        row = z[r]
        assert np.all(np.diff(row) >= -1e-6), r                     # non-decreasing depths
        assert np.all((row[:4] >= t0 - vs - 1e-5) & (row[:4] < t0 + 1e-5)), r   # 4 before the first run
        assert np.allclose(dist[r, :4], vs / 4, rtol=1e-4)
        inner = row[4:]
        inner = inner[(inner > t0 + 1e-5) & (inner < t1 - 1e-5)]
        for zz in inner:                                              # interior samples: inside flagged voxels
            p = o[r].astype(np.float64) + zz * d[r].astype(np.float64)
            v = _voxel_of(p)
            ok = False
            for dv in np.array(np.meshgrid([-1, 0, 1], [-1, 0, 1], [-1, 0, 1])).T.reshape(-1, 3):
                w = v + dv
                if np.all((w >= 0) & (w < NV)) and fl[(w[0] * NV + w[1]) * NV + w[2]] == 1:
                    b0 = (CTR - 0.5) + w * vs
                    if np.all(p >= b0 - 1e-4) and np.all(p <= b0 + vs + 1e-4):
                        ok = True
            assert ok, (r, zz)
    assert checked > 100


def test_sparse_short_rows_stay_in_bounds(oracle_mod):
    o, d = _rays(6, 200)
    fl = _flags(7, 0.3)
    for N in (1, 3, 5, 8):
        # 64 guard entries before and after the rows: the sampler must write only inside them
        zb = np.full(200 * N + 128, -7.0, np.float32)
        db = np.full(200 * N + 128, -7.0, np.float32)
        zz, dd = zb[64:64 + 200 * N].reshape(200, N), db[64:64 + 200 * N].reshape(200, N)
        oracle_mod.sample_sparse(o, d, fl, CTR, TS, NV, zz, dd)
        for b in (zb, db):
            assert (b[:64] == -7.0).all() and (b[-64:] == -7.0).all(), N
        assert np.all(np.isfinite(zz[zz != -7.0]))


def test_reflection_transmittance_threshold(oracle_mod):
    o, d = _rays(8, 300)
    fl = _flags(9, 0.25)
    rng = np.random.default_rng(10)
    vox = np.zeros((NV, NV, NV, 4), np.float32)
    vox[..., :3] = rng.random((NV, NV, NV, 3))
    vox[..., 3] = rng.uniform(0.0, 40.0, (NV, NV, NV))
    N, far, T_th = 16, 6.0, 0.3
    z = np.full((300, N), -7.0, np.float32)
    oracle_mod.sample_reflection(o, d, fl, vox, CTR, TS, far, NV, T_th, z)
    vs = TS / NV
    ijk = np.stack(np.meshgrid(*[np.arange(NV)] * 3, indexing="ij"), -1).reshape(-1, 3)
    lo = (CTR - 0.5) + ijk * vs
    decided = 0
    for r in range(300):
        tn, tf = _slab64(np.repeat(o[r:r + 1], ijk.shape[0], 0), np.repeat(d[r:r + 1], ijk.shape[0], 0), lo, lo + vs)
        sel = np.flatnonzero((tn < tf) & (fl == 1))
        sel = sel[np.argsort(tn[sel])]
        T, near, margin = 1.0, None, np.inf
        for k in sel:
            p = o[r].astype(np.float64) + tn[k] * d[r].astype(np.float64)
            f = (p - (CTR - 0.5)) / vs + 0.5                          # overlap index of the entry point
            sig = ndimage.map_coordinates(vox[..., 3].astype(np.float64), f.reshape(3, 1), order=1, mode="nearest")[0]
            T *= np.exp(-sig * (tf[k] - tn[k]))
            margin = min(margin, abs(T - T_th))
            if T <= T_th:
                near = tf[k]
                break
        if margin < 1e-3:
            continue                                                  # too close to the threshold to pin
        decided += 1
        if near is None:
            assert (z[r] == -7.0).all(), r
        else:
            ref = 1.0 / (np.arange(N) / (N - 1) * (1.0 / far - 1.0 / near) + 1.0 / near)
            assert np.abs(z[r] - ref).max() < 1e-4 * max(1.0, far), r
    assert decided > 200
