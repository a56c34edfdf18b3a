"""GPU training samplers (snisr_sample_*) against the FP32 oracle, bit-exact: sparse depths and
intervals, reflection inverse-depth samples, uniform chord samples and background samples, with rows
the reference leaves untouched (rays that miss the tile, no flagged voxel, bg = -1), axis-parallel
rays, short rows, and the CUDA_EXT call form (snisr.samplers)."""
import numpy as np
import pytest
import torch

from snisr import samplers as SM

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def _rays(seed, n, ctr, ts):
    rng = np.random.default_rng(seed)
    o = (ctr + rng.uniform(-1.5, 1.5, (n, 3)) * ts).astype(np.float32)
    tgt = (ctr + rng.uniform(-0.4, 0.4, (n, 3)) * ts).astype(np.float32)
    d = (tgt - o).astype(np.float32)
    d[:50, 1:] = 0.0              # axis-parallel
    d[50:80, 2] = 0.0
    d[80:100] *= -1.0             # pointing away: many miss the tile
    return o, d


@pytest.mark.parametrize("nv,N", [(16, 48), (32, 64), (12, 5), (7, 40), (64, 64), (81, 64)])
def test_sample_sparse_bit_exact(oracle_mod, nv, N):
    """nv 64 is the reference's NUM_VOXEL; 7 (343 voxels) builds the kernel's LDS occupancy bits with
    scalar loads, 81 (more voxels than the LDS bitmap holds) reads the int16 flags directly.  Flags
    other than 1 (0, -1, 2) are unoccupied."""
    ctr, ts = np.array([1.5, 0.5, 2.5], np.float32), np.float32(1.25)
    o, d = _rays(nv, 4000, ctr, ts)
    rng = np.random.default_rng(nv + 1)
    flags = np.where(rng.random((nv, nv, nv)) < 0.12, 1, rng.choice([-1, 0, 2], (nv, nv, nv))).astype(np.int16)
    z = np.full((4000, N), -7.0, np.float32)
    ds = np.full((4000, N), -7.0, np.float32)
    zg, dg = to_dev(z), to_dev(ds)
    SM.Sample_sparse(to_dev(o), to_dev(d), to_dev(flags), *ctr.tolist(), float(ts), nv, zg, dg)
    oracle_mod.sample_sparse(o, d, flags.reshape(-1), ctr, ts, nv, z, ds)
    assert (z[:, 0] != -7.0).sum() > 500
    assert np.array_equal(bits(zg.cpu().numpy()), bits(z))
    assert np.array_equal(bits(dg.cpu().numpy()), bits(ds))


def test_sample_reflection_bit_exact(oracle_mod):
    nv, N = 16, 16
    ctr, ts = np.array([0.5, 0.5, 0.5], np.float32), np.float32(1.0)
    o, d = _rays(3, 4000, ctr, ts)
    rng = np.random.default_rng(4)
    flags = (rng.random((nv, nv, nv)) < 0.25).astype(np.int16)
    vox = np.zeros((nv, nv, nv, 4), np.float32)
    vox[..., :3] = rng.random((nv, nv, nv, 3))
    vox[..., 3] = rng.uniform(0.0, 60.0, (nv, nv, nv))
    z = np.full((4000, N), -7.0, np.float32)
    zg = to_dev(z)
    SM.Sample_reflection(to_dev(o), to_dev(d), to_dev(flags), to_dev(vox), *ctr.tolist(), float(ts), 5.0, nv, 0.3, zg)
    oracle_mod.sample_reflection(o, d, flags.reshape(-1), vox, ctr, ts, 5.0, nv, 0.3, z)
    assert (z[:, 0] != -7.0).sum() > 500
    assert np.array_equal(bits(zg.cpu().numpy()), bits(z))


def test_sample_uniform_and_bg_bit_exact(oracle_mod):
    ctr, ts = np.array([2.0, -1.0, 0.5], np.float32), np.float32(2.0)
    o, d = _rays(5, 3000, ctr, ts)
    z = np.zeros((3000, 9), np.float32)
    zg = to_dev(z)
    SM.Sample_uniform(to_dev(o), to_dev(d), *ctr.tolist(), float(ts), zg)
    oracle_mod.sample_uniform(o, d, ctr, ts, z)
    assert np.array_equal(bits(zg.cpu().numpy()), bits(z))
    bg = np.random.default_rng(6).uniform(0.0, 6.0, (3000, 1)).astype(np.float32)
    bg[::5] = -1.0
    zb = np.full((3000, 7), -7.0, np.float32)
    zbg = to_dev(zb)
    SM.Sample_bg(to_dev(o), to_dev(d), to_dev(bg), to_dev(ctr), float(ts), 0.75, zbg)
    oracle_mod.sample_bg(o, d, bg, ctr, ts, np.float32(0.75), zb)
    assert np.array_equal(bits(zbg.cpu().numpy()), bits(zb))
    assert (zb[::5] == -7.0).all()


def test_sampler_argument_checks():
    o = torch.zeros((4, 3), device=DEV)
    z = torch.zeros((4, 8), device=DEV)
    with pytest.raises(ValueError):
        SM.Sample_uniform(o, o[:, :2].contiguous(), 0.0, 0.0, 0.0, 1.0, z)
    with pytest.raises(ValueError):
        SM.Sample_sparse(o, o, torch.zeros((4, 4, 4), dtype=torch.int16, device=DEV), 0, 0, 0, 1.0, 8, z, z.clone())
