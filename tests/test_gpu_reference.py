"""GPU reference direct volume renderer (mray_hip_reference) against the
compiled reference's output in tests/golden/golden_reference.npz and against
the oracle's restatement of reference_ray (src/raymarch_reference.h:93-162).

Bar: bit-identical float RGBA and step counts (the kernel follows the
reference CPU arithmetic operation for operation)."""
import pathlib

import numpy as np
import pytest

from mray_amd import synth
from mray_amd.types import RefParams
from oracle_ffi import Oracle
from test_oracle_golden import DVR_CASES

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def hip():
    from mray_amd import hip as h

    h.init(0)
    return h


def render(hip, vol, dmin, dmax, cam, W, H, table, p):
    import torch

    v = hip.Volume(torch.from_numpy(np.ascontiguousarray(vol)).cuda(), dmin, dmax)
    rgba, steps = hip.reference(v, cam, torch.from_numpy(table).cuda(), p, W, H, want_steps=True)
    torch.cuda.synchronize()
    return rgba.cpu().numpy(), steps.cpu().numpy()


@pytest.mark.parametrize("case", list(DVR_CASES))
def test_reference_matches_golden(hip, case):
    """The compiled reference's float RGBA (and RGBA8 after clamp and
    truncation) on its own fixture volumes, bit for bit."""
    g = dict(np.load(GOLD / "golden_reference.npz"))
    dmin, dmax, cam, W, H, pk, tf_step = DVR_CASES[case]
    O = Oracle()
    table = O.tf(g["tf_rgba8"], tf_step)
    print(f"[reference dvr] case={case} W={W} H={H}", flush=True)
    rgba, steps = render(hip, g[f"{case}_volume"], dmin, dmax, cam, W, H, table, RefParams(**pk))
    gold = g[f"{case}_rgba"]
    same = rgba.view(np.uint32) == gold.view(np.uint32)
    print(f"  bit-identical values {same.sum()}/{same.size}")
    assert same.all()
    q = (np.clip(rgba, 0, 1) * np.float32(255)).astype(np.uint8)
    assert np.array_equal(q, g[f"{case}_rgba8"])


@pytest.mark.parametrize("dtype", ["f32", "u8"])
def test_reference_matches_oracle(hip, dtype):
    """Larger volumes and images against the oracle: RGBA and the compositing
    step counts bit-identical (perspective camera partly outside the
    volume, early termination on part of the image)."""
    O = Oracle()
    if dtype == "f32":
        vol = synth.marschner_lobb(64)
        dmin, dmax = (-1, -1, -1), (1, 1, 1)
        cam = synth.perspective_camera(position=(0.6, 0.2, -2.4), fov_half_deg=26.0, aspect=56 / 44)
        p = RefParams(step_size=2.0 / 128, medium_emission=1.0, early_ray_termination=0.95)
        tf_step = 24 * 2.0 / 128
    else:
        vol = np.clip(synth.mixing_layer(48) * 255, 0, 255).astype(np.uint8)
        dmin, dmax = (0, 0, 0), (1, 1, 1)
        cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=56 / 44)
        p = RefParams(step_size=0.5 / 48, medium_emission=0.7, early_ray_termination=0.9, data_min=0.05,
                      data_max=0.95)
        tf_step = 400 * 0.5 / 48
    W, H = 56, 44
    table = O.tf(synth.tf_ramps(0), tf_step)
    print(f"[reference dvr] dtype={dtype} W={W} H={H}", flush=True)
    rgba, steps = render(hip, vol, dmin, dmax, cam, W, H, table, p)
    ref, ns = O.reference(vol, dmin, dmax, cam, W, H, table, p)
    same = rgba.view(np.uint32) == ref.view(np.uint32)
    print(f"  bit-identical values {same.sum()}/{same.size}, terminated early "
          f"{np.count_nonzero(ref[..., 3] == 1.0)}, missed {np.count_nonzero(ns == 0)}")
    assert np.count_nonzero(ref[..., 3] == 1.0) > 0
    assert same.all()
    assert np.array_equal(steps, ns)


def test_reference_rejects_single_scattering(hip):
    """Single scattering (the reference's ss.enable with sigma_s > 0) is not
    built: the call fails with a message instead of rendering without it."""
    import torch

    vol = synth.marschner_lobb(16)
    v = hip.Volume(torch.from_numpy(vol).cuda(), (-1, -1, -1), (1, 1, 1))
    table = torch.zeros((256, 4), dtype=torch.float32, device="cuda")
    p = RefParams(step_size=0.1, single_scattering=1, medium_scattering=0.5)
    with pytest.raises(hip.MrayError, match="single scattering"):
        hip.reference(v, synth.ortho_camera(), table, p, 8, 8)
