"""The FP64 adjoint of the reconstruct march in the oracle (the checker of the
GPU adjoint) against central finite differences of the FP64 forward march.

L = sum(g * rgba) for a fixed random g. Directional derivatives along random
directions of the multipliers and of the 256 x 4 table; steps small enough
(1e-8 relative) that no sample crosses a transfer-function cell boundary,
where the reference's lookup is piecewise. Early ray termination is off
(threshold above 1), since it makes the image discontinuous in tau."""
import numpy as np
import pytest

from mray_amd import synth, tf as tfmod
from mray_amd.types import GenParams, RecParams
from oracle_ffi import Oracle


@pytest.mark.parametrize("persp", [False, True])
def test_reconstruct_adjoint_matches_finite_differences(persp):
    O = Oracle()
    if persp:
        vol, dmin, dmax = synth.mixing_layer(24, width=0.05), (0, 0, 0), (1, 1, 1)
        W, H, m, step = 14, 12, 16, 0.5 / 24
        cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=W / H)
        factor = 40.0
    else:
        vol, dmin, dmax = synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1)
        W, H, m, step = 12, 10, 12, 2.0 / 64
        cam = synth.ortho_camera()
        factor = 4.0
    gp = GenParams(step_size=step, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    mo, b, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    rp = RecParams(step_size=step, bias=1e-3, early_ray_termination=1.5, medium_emission=1.0)
    lag = np.stack([O.prepare_std(m, mo.reshape(-1, m)[p], rp.bias) for p in range(W * H)]).astype(np.float32)
    table = tfmod.table_from_rgba8(synth.tf_ramps(0), step * factor)
    rng = np.random.default_rng(11)
    g = rng.normal(size=(H * W, 4))
    gl, gt = O.reconstruct_adjoint_f64(cam, dmin, dmax, W, H, m, lag.reshape(-1), table, rp, g, bounds=b)

    lag64, tf64 = lag.astype(np.float64).reshape(-1), table.astype(np.float64)

    def loss(l64, t64):
        return float((g.reshape(H, W, 4) * O.reconstruct_lagrange64_f64(cam, dmin, dmax, W, H, m, l64, t64, rp,
                                                                          bounds=b)).sum())

    for trial in range(3):
        d = rng.normal(size=lag64.shape) * np.maximum(np.abs(lag64), 1e-2)
        eps = 1e-8
        fd = (loss(lag64 + eps * d, tf64) - loss(lag64 - eps * d, tf64)) / (2 * eps)
        an = float(gl @ d)
        print(f"[adjoint] lambda: fd {fd:.9e} adjoint {an:.9e}")
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3) + 1e-9 * np.abs(gl).sum() * np.abs(d).max()

        dt = rng.normal(size=tf64.shape) * np.maximum(np.abs(tf64), 1e-3)
        fd = (loss(lag64, tf64 + eps * dt) - loss(lag64, tf64 - eps * dt)) / (2 * eps)
        an = float((gt * dt).sum())
        print(f"[adjoint] table: fd {fd:.9e} adjoint {an:.9e}")
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3)
