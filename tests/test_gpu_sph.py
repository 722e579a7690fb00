"""SPH particles -> grid on the GPU (through the C ABI).

Bars: bit-identical to the oracle (the kernel sums over cells in z, y, x order
and over particles in stable cell order, as the oracle does); against the
compiled reference's own grid (golden_sph.npz) within the tolerance the
oracle itself meets there (rtol 2e-6, atol 1e-7), since the reference's
std::sort leaves the order of particles inside a cell unspecified.
"""
import pathlib

import numpy as np
import pytest

from mray_amd import synth
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def hip():
    from mray_amd import hip as H

    H.init(0)
    return H


@pytest.fixture(scope="module")
def O():
    return Oracle()


def run_gpu(hip, x, y, z, v, h, dmin, dmax, nx, ny, nz):
    import torch

    t = [torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda() for a in (x, y, z, v)]
    g = hip.particles_to_grid(*t, h, dmin, dmax, nx, ny, nz)
    torch.cuda.synchronize()
    return g.cpu().numpy()


def test_sph_golden(hip, O):
    g = dict(np.load(GOLD / "golden_sph.npz"))
    args = (g["sph_x"], g["sph_y"], g["sph_z"], g["sph_v"], float(g["sph_h"][0]), (0, 0, 0), (1, 1, 0.5), 16, 16, 8)
    grid = run_gpu(hip, *args)
    ref = O.particles_to_grid(*args)
    print(f"[sph] golden: nonzero {np.count_nonzero(grid)}/{grid.size}")
    assert np.array_equal(grid.view(np.uint32), ref.view(np.uint32))
    assert np.allclose(grid, g["sph_grid"], rtol=2e-6, atol=1e-7)


@pytest.mark.parametrize("n,h,res", [(200_000, 0.01, (64, 64, 32)), (50_000, 0.004, (96, 96, 48)),
                                     (1, 0.05, (8, 8, 4)), (3000, 0.2, (16, 12, 8))])
def test_sph_matches_oracle(hip, O, n, h, res):
    x, y, z, _ = synth.particle_mixture(n, seed=5)
    v = (0.5 + 0.5 * np.sin(37.0 * x) * np.cos(23.0 * y)).astype(np.float32)  # non-uniform values
    nx, ny, nz = res
    args = (x, y, z, v, h, (0, 0, 0), (1, 1, 0.5), nx, ny, nz)
    grid = run_gpu(hip, *args)
    ref = O.particles_to_grid(*args)
    diff = np.count_nonzero(grid.view(np.uint32) != ref.view(np.uint32))
    print(f"[sph] n={n} h={h} grid={res}: nonzero {np.count_nonzero(ref)}, differing voxels {diff}")
    assert diff == 0


def test_sph_domain_beyond_particles(hip, O):
    """Voxels far outside the particles' bounding box get 0."""
    x, y, z, v = synth.particle_mixture(5000, seed=9)
    args = (x * 0.3 + 0.2, y * 0.3 + 0.2, z * 0.3 + 0.1, v, 0.02, (-1, -1, -1), (2, 2, 2), 24, 20, 16)
    grid = run_gpu(hip, *args)
    ref = O.particles_to_grid(*args)
    assert np.array_equal(grid.view(np.uint32), ref.view(np.uint32))
    assert grid[0, 0, 0] == 0.0


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_particle_march_matches_oracle(hip, O, form):
    """generate on a particle set (mray_hip_volume_from_particles): every
    sample is Particles::sample, as the reference's generate on Particles
    does. The gather is the oracle's operation for operation (stable cell
    order), so sample counts and 16-bit bounds are bit-identical and the
    moments meet the bars of test_gpu_parity.py (plain <= 1e-5 normwise,
    mixed <= 1e-5 + the FP32 oracle's own distance from FP64); the golden
    march of the compiled reference (golden_sph.npz) within its sort-order
    tolerance."""
    import torch

    from mray_amd.types import GenParams
    from oracle_ffi import Particles

    g = dict(np.load(GOLD / "golden_sph.npz"))
    pts = [g["sph_x"], g["sph_y"], g["sph_z"], g["sph_v"]]
    h = float(g["sph_h"][0])
    dmin, dmax = (0.0, 0.0, 0.0), (1.0, 1.0, 0.5)
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=24.0, aspect=1.0)
    gp = GenParams(step_size=0.02, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    W = H = 14
    m = 12
    print(f"[particle march] form={form} n={len(pts[0])} W={W} H={H} m={m}", flush=True)
    pv = hip.ParticleVolume(*[torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda() for a in pts], h, dmin,
                            dmax)
    mo, b, ns = hip.generate(pv, cam, gp, W, H, m, form=form, want_samples=True)
    torch.cuda.synchronize()
    parts = Particles(O, *pts, h)
    mo_ref, b_ref, ns_ref = O.generate(parts, dmin, dmax, cam, W, H, m, gp)
    assert np.array_equal(ns.cpu().numpy(), ns_ref)
    assert np.array_equal(b.cpu().numpy().view(np.uint32), b_ref)
    a = mo.cpu().numpy().reshape(-1, m).astype(np.float64)
    r = mo_ref.reshape(-1, m).astype(np.float64)
    err = np.linalg.norm(a - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-3)
    if form == "plain":
        assert err.max() <= 1e-5
    else:
        mo64, _ = O.generate(parts, dmin, dmax, cam, W, H, m, gp, f64=True)
        r64 = mo64.reshape(-1, m)
        own = np.linalg.norm(r - r64, axis=1) / np.maximum(np.linalg.norm(r64, axis=1), 1e-3)
        assert np.all(err <= 1e-5 + own)
    gold = g["sph_march_moments"].reshape(-1, m).astype(np.float64)
    errg = np.linalg.norm(a - gold, axis=1) / np.maximum(np.linalg.norm(gold, axis=1), 1e-3)
    print(f"  vs oracle max {err.max():.2e}, vs compiled reference {errg.max():.2e}")
    assert errg.max() <= 2e-5
    pv.close()

