"""A volume larger than 4 GiB: the mixed generate kernel then addresses
voxels with 64-bit offsets (up to 4 GiB it uses 32-bit byte offsets from the
base). 1025 x 1024 x 1024 f32 voxels (4.03 GiB), the last z slab lying past
the 4 GiB mark; 16 x 16 perspective rays cross the whole volume. The
reference renderer runs through the same voxels (bit-identical RGBA and
step counts).

Bars: sample counts and 16-bit bounds bit-exact against the FP32 oracle on
the same voxels (they exercise every address); moments against the FP64
oracle no further off than the FP32 reference itself, over the image (max
and median of the per-pixel normwise error, + 1e-5). These 2080-sample rays
through a slowly varying field have small moment norms, so both FP32
evaluations are ~1e-4 off in relative terms at worst (measured: this kernel
5.7e-4, the FP32 reference 7.4e-4)."""
import numpy as np
import pytest

from mray_amd import synth
from mray_amd.types import GenParams
from oracle_ffi import Oracle

pytestmark = pytest.mark.gpu


def normwise_rel(a, b, floor=1e-3):
    a = a.reshape(-1, a.shape[-1]).astype(np.float64)
    b = b.reshape(-1, b.shape[-1]).astype(np.float64)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), floor)


def test_generate_mixed_beyond_4gib():
    import torch
    from mray_amd import hip

    hip.init(0)
    nz, ny, nx = 1025, 1024, 1024
    print(f"[large volume] {nz}x{ny}x{nx} f32 = {nz * ny * nx * 4 / 2**30:.2f} GiB", flush=True)
    z = torch.linspace(0.0, 1.0, nz, device="cuda").view(nz, 1, 1)
    y = torch.linspace(0.0, 1.0, ny, device="cuda").view(1, ny, 1)
    x = torch.linspace(0.0, 1.0, nx, device="cuda").view(1, 1, nx)
    vox = (0.5 + 0.25 * torch.sin(7.0 * x) * torch.cos(5.0 * y) + 0.2 * (z - 0.5)).contiguous()
    # a distinct plateau in the slab past 4 GiB, so the far samples set the upper bounds
    vox[nz - 1] = 0.97
    dmin, dmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=12.0, znear=0.01, aspect=1.0)
    W = H = 16
    m = 24
    gp = GenParams(step_size=0.5 / 1024, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    v = hip.Volume(vox, dmin, dmax)
    mo, b, ns = hip.generate(v, cam, gp, W, H, m, form="mixed", want_samples=True)
    # the reference renderer through the same voxels (its 64-bit path too)
    from mray_amd import tf as tfmod
    from mray_amd.types import RefParams

    refp = RefParams(step_size=gp.step_size, medium_emission=1.0, early_ray_termination=0.99)
    table = tfmod.table_from_rgba8(synth.tf_ramps(0), 64 * gp.step_size)
    dvr, dvr_steps = hip.reference(v, cam, torch.from_numpy(table).cuda(), refp, W, H, want_steps=True)
    torch.cuda.synchronize()
    dvr = dvr.cpu().numpy()
    dvr_steps = dvr_steps.cpu().numpy()
    mo = mo.cpu().numpy()
    b = b.cpu().numpy().view(np.uint32)
    ns = ns.cpu().numpy()
    vol = vox.cpu().numpy()
    del v, vox
    O = Oracle()
    mo_ref, b_ref, ns_ref = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    assert np.array_equal(ns, ns_ref), "sample counts differ"
    assert np.array_equal(b, b_ref), "16-bit bounds differ"
    assert np.any((b_ref >> 16) >= int(0.97 * 65535)), "no ray reached the slab past 4 GiB"
    dvr_ref, dvr_ns = O.reference(vol, dmin, dmax, cam, W, H, table, refp)
    assert np.array_equal(dvr_steps, dvr_ns), "reference renderer step counts differ"
    assert np.array_equal(dvr.view(np.uint32), dvr_ref.view(np.uint32)), "reference renderer RGBA differs"
    mo64, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True)
    hit = ns_ref > 0
    err = normwise_rel(mo, mo64)[hit]
    own = normwise_rel(mo_ref, mo64)[hit]
    print(f"  vs FP64: this kernel max {err.max():.2e} median {np.median(err):.2e}; "
          f"FP32 reference max {own.max():.2e} median {np.median(own):.2e}", flush=True)
    assert err.max() <= own.max() + 1e-5
    assert np.median(err) <= np.median(own) + 1e-5
