"""The particle workload end to end at reduced size (the shape of the 20 M
particle / 512 x 512 x 256 / 1024^2 / 40-moment benchmark): particles from the
32-Gaussian mixture in [0,1] x [0,1] x [0,0.5] gathered onto a grid, a march
of that grid in both forms, and the adjoint of the march with respect to the
grid values.

* gather: bit-identical to the oracle's Particles::sample at every voxel;
* generate on the gathered grid: 16-bit bounds and sample counts bit-exact;
  moments of both forms within 1e-5 normwise of the FP32 oracle plus the FP32
  oracle's own distance from the FP64 oracle (PLAIN takes the device's sincosf
  where the oracle takes the host's: on this grid's sharp occupancy edges
  their ulps reach 1.2e-5 at 40 moments);
* adjoint: dL/dgrid of generate (L = sum gbar * moments, no bound
  normalisation) against central differences of the FP64 oracle's generate
  (linear in the voxels: exact up to float storage), bar 2e-4 relative;
* the full chain generate -> prepare -> reconstruct through torch.autograd on
  the gathered grid: a short step against the gradient lowers an image loss.
"""
import numpy as np
import pytest

from mray_amd import synth, tf as tfmod
from mray_amd.types import GenParams, RecParams
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu

N, H_SPH = 60000, 0.02
NX, NY, NZ = 64, 64, 32
DMIN, DMAX = (0.0, 0.0, 0.0), (1.0, 1.0, 0.5)
W = H = 40
M = 40


@pytest.fixture(scope="module")
def gathered():
    import torch

    from mray_amd import hip

    hip.init(0)
    x, y, z, v = synth.particle_mixture(N, seed=5, n_gauss=32, box=(1.0, 1.0, 0.5))
    # values that vary across the box (uniform mass would gather to a 0 / 1
    # occupancy)
    v = (0.5 + 0.5 * np.cos(6.0 * x) * np.sin(5.0 * y)).astype(np.float32)
    print(f"[particles] {N} particles, h={H_SPH} -> {NX}x{NY}x{NZ}", flush=True)
    d = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (x, y, z, v)]
    grid = hip.particles_to_grid(*d, H_SPH, DMIN, DMAX, NX, NY, NZ)
    torch.cuda.synchronize()
    ref = Oracle().particles_to_grid(x, y, z, v, H_SPH, DMIN, DMAX, NX, NY, NZ)
    return hip, torch, grid, ref


def test_gather_matches_oracle(gathered):
    hip, torch, grid, ref = gathered
    g = grid.cpu().numpy()
    print(f"[particles] occupied voxels {np.mean(g > 0):.3f}", flush=True)
    assert np.mean(g > 0) > 0.05
    assert np.array_equal(g.view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("form", ["plain", "mixed"])
def test_march_of_gathered_grid(gathered, form):
    hip, torch, grid, ref = gathered
    O = Oracle()
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=25.0, znear=0.01, aspect=1.0)
    gp = GenParams(step_size=0.5 / NX, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0, bias=1e-4)
    mo_ref, b_ref, ns_ref = O.generate(ref, DMIN, DMAX, cam, W, H, M, gp)
    v = hip.Volume(grid, DMIN, DMAX)
    mo, b, ns = hip.generate(v, cam, gp, W, H, M, form=form, want_samples=True)
    torch.cuda.synchronize()
    assert np.array_equal(b.cpu().numpy().view(np.uint32), b_ref)
    assert np.array_equal(ns.cpu().numpy(), ns_ref)
    hit = ns_ref > 0
    a = mo.cpu().numpy().reshape(-1, M).astype(np.float64)
    r = mo_ref.reshape(-1, M).astype(np.float64)
    err = (np.linalg.norm(a - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-3))[hit]
    mo64, _ = O.generate(ref, DMIN, DMAX, cam, W, H, M, gp, f64=True)
    r64 = mo64.reshape(-1, M)
    own = (np.linalg.norm(r - r64, axis=1) / np.maximum(np.linalg.norm(r64, axis=1), 1e-3))[hit]
    print(f"[particles] {form} vs fp32 max {err.max():.2e}, fp32 own max {own.max():.2e}", flush=True)
    assert np.all(err <= 1e-5 + own)


def test_adjoint_to_grid_values(gathered):
    hip, torch, grid, ref = gathered
    O = Oracle()
    rng = np.random.default_rng(7)
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=25.0, znear=0.01, aspect=1.0)
    Wa = Ha = 16
    gp = GenParams(step_size=1.0 / NX, transform_to_bounds=0, data_min=-0.1, data_max=1.1, compact_image=0,
                   prediction_coding=0)
    gbar = rng.normal(size=(Ha, Wa, M))
    v = hip.Volume(grid, DMIN, DMAX)
    gv = hip.generate_backward(v, cam, gp, Wa, Ha, M, torch.from_numpy(gbar.astype(np.float32)).cuda())
    gv = gv.cpu().numpy().astype(np.float64)

    def loss(vv):
        mo, _ = O.generate(np.ascontiguousarray(vv, np.float32), DMIN, DMAX, cam, Wa, Ha, M, gp, f64=True)
        return float((gbar * mo).sum())

    for trial in range(2):
        d = rng.normal(size=ref.shape)
        eps = 1e-2
        fd = (loss(ref + eps * d) - loss(ref - eps * d)) / (2 * eps)
        an = float((gv * d).sum())
        print(f"[particles adjoint] fd {fd:.9e} adjoint {an:.9e}", flush=True)
        assert abs(fd - an) <= 2e-4 * np.abs(gv * d).sum()


def test_autograd_chain_on_gathered_grid(gathered):
    hip, torch, grid, ref = gathered
    from mray_amd import autograd as ag

    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=25.0, znear=0.01, aspect=1.0)
    gp = GenParams(step_size=1.0 / NX, transform_to_bounds=0, data_min=-0.1, data_max=1.1, compact_image=0,
                   prediction_coding=0)
    rp = RecParams(step_size=1.0 / NX, bias=1e-3, early_ray_termination=1.5, medium_emission=1.0)
    table = torch.from_numpy(tfmod.table_from_rgba8(synth.tf_ramps(0), rp.step_size * 8.0)).cuda()
    with torch.no_grad():
        target = ag.render(grid, DMIN, DMAX, cam, gp, 24, 24, M, table, rp)
    x = (grid * 0.7 + 0.1).clone().requires_grad_(True)
    img = ag.render(x, DMIN, DMAX, cam, gp, 24, 24, M, table, rp)
    loss = ((img - target) ** 2).sum()
    loss.backward()
    with torch.no_grad():
        x2 = x - (1e-3 / float(x.grad.abs().max())) * x.grad
        l2 = float(((ag.render(x2, DMIN, DMAX, cam, gp, 24, 24, M, table, rp) - target) ** 2).sum())
    print(f"[particles autograd] loss {float(loss):.6e} -> {l2:.6e}", flush=True)
    assert l2 < float(loss)
