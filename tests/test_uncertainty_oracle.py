"""The oracle's error bounds and uncertainty renderings against the compiled
reference (golden_uncertainty.npz), without a GPU:

* the uncertainty transfer functions (box filter for the maximum error and
  the percentile, gaussian filter for the RMSE) bit for bit (every 17th row
  and column, and a SHA-256 of the whole 256 x 256 x RGBA table);
* reconstruct_error's per-pixel uint16 error bounds for the maximum, the
  RMSE and the P^2 90th percentile, bit for bit, on compact, prediction-coded
  and dense images, an orthographic f32 and a perspective u8 volume, and a
  step so coarse that rays hold fewer than five samples (the percentile
  estimator's start-up phase);
* the float RGBA of the uncertainty interpolation and of the convolution
  for each error type, bit for bit.
The error measures are also pinned by a closed form: a constant volume
against a constant reconstruction gives the same error at every sample, so
the maximum, the RMSE and the percentile estimate all equal it.
"""
import hashlib
import pathlib

import numpy as np
import pytest

from mray_amd import synth
from mray_amd.types import ErrParams, RecParams
from oracle_ffi import Oracle

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
CASES = ["ml_ortho", "mix_persp_u8", "ml_coded", "ml_dense_coarse"]
CASE_SPEC = {"ml_ortho": ("ml", 0, "ortho", 1.0 / 40), "mix_persp_u8": ("mix_u8", 0, "persp", 1.0 / 24),
             "ml_coded": ("ml", 1, "ortho", 1.0 / 40), "ml_dense_coarse": ("ml", 0, "ortho", 0.45)}


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD / "golden_uncertainty.npz")


def case_inputs(gold, name):
    vk, pc, ck, estep = CASE_SPEC[name]
    W, H, m = (int(v) for v in gold[f"{name}_shape"])
    if vk == "ml":
        vol, dmin, dmax = synth.marschner_lobb(32), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    else:
        vol = np.clip(synth.mixing_layer(32, width=0.05) * 255, 0, 255).astype(np.uint8)
        dmin, dmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    cam = synth.ortho_camera() if ck == "ortho" else synth.perspective_camera(
        position=(0.5, 0.5, -1.2), fov_half_deg=30.0, znear=0.01, aspect=W / H)
    index = gold[f"{name}_index"] if f"{name}_index" in gold.files else None
    cp = gold[f"{name}_coding_params"] if pc else None
    cpf = None if cp is None else (cp.astype(np.float32) / np.float32(65535.0)).astype(np.float32)
    return dict(vol=vol, dmin=dmin, dmax=dmax, cam=cam, W=W, H=H, m=m, index=index, pc=pc, cp=cp, cpf=cpf,
                estep=estep, data=gold[f"{name}_data"], bounds=gold[f"{name}_bounds"])


@pytest.mark.parametrize("et", [0, 1])
def test_uncertainty_tables_match_reference(gold, et):
    t = Oracle().tf_uncertainty(synth.tf_ramps(0), 0.1, et)
    assert np.array_equal(t[::17, ::17].view(np.uint32), gold[f"tf{et}_sub"].view(np.uint32))
    assert hashlib.sha256(t.tobytes()).digest() == gold[f"tf{et}_sha256"].tobytes()


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("et", [0, 1, 2])
def test_error_bounds_match_reference(gold, name, et):
    c = case_inputs(gold, name)
    ep = ErrParams(error_type=et, error_percentile=90 if et == 2 else 95, bias=1e-4, step_size=c["estep"])
    got = Oracle().error_bounds(c["vol"], c["dmin"], c["dmax"], c["cam"], c["W"], c["H"], c["m"], c["data"], ep,
                                index=c["index"], bounds=c["bounds"], prediction_coded=c["pc"],
                                coding_params=c["cpf"])
    assert np.array_equal(got, gold[f"{name}_err{et}"])


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mode,et,s", [(0, 2, 0.3), (1, 0, 0.0), (1, 1, 0.0), (1, 2, 0.0)])
def test_uncertainty_images_match_reference(gold, name, mode, et, s):
    c = case_inputs(gold, name)
    O = Oracle()
    rp = RecParams(step_size=1.0 / 40, bias=1e-4, early_ray_termination=0.99, medium_emission=1.0)
    tf = O.tf(synth.tf_ramps(0), rp.step_size * 4) if mode == 0 else O.tf_uncertainty(synth.tf_ramps(0),
                                                                                      rp.step_size * 4, et)
    got = O.uncertainty(c["cam"], c["dmin"], c["dmax"], c["W"], c["H"], c["m"], c["data"], gold[f"{name}_err{et}"],
                        tf, rp, mode, et, s, index=c["index"], bounds=c["bounds"], prediction_coded=c["pc"],
                        coding_params=c["cpf"])
    assert np.array_equal(got.view(np.uint32), gold[f"{name}_unc{mode}{et}"].view(np.uint32))


def test_error_measures_closed_forms():
    """A constant volume against a constant reconstruction: every sample's
    error is the same d, so the maximum, the RMSE and any percentile are d."""
    O = Oracle()
    vol = np.full((8, 8, 8), 0.25, np.float32)
    dmin, dmax = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    cam = synth.ortho_camera()
    W = H = 6
    # one moment of 0.5 per pixel: the bounded MESE of the zeroth moment
    # alone is the constant 0.5 (bounds [0, 1])
    index = np.arange(W * H + 1, dtype=np.uint32)
    for et in (0, 1, 2):
        ep = ErrParams(error_type=et, error_percentile=90, bias=1e-4, step_size=0.05)
        d = O.error_bounds(vol, dmin, dmax, cam, W, H, 1, np.full(W * H, 0.5, np.float32), ep, index=index)
        hit = d > 0
        assert hit.any()
        # |0.5 - 0.25| = 0.25 -> 0.25 * 65535 truncated
        assert np.all(d[hit] == np.uint16(np.float32(0.25) * np.float32(65535.0)))


@pytest.mark.parametrize("et", [0, 1, 2])
def test_host_library_tables_match_oracle(et):
    """read_from_disk_uncertainty of the C++ host library (the tables the
    product renders with) against the oracle's, bit for bit."""
    from mray_amd import moment_file, tf as tfmod

    if not moment_file.LIB_PATH.exists():
        pytest.skip("libmray_host.so not built")
    for seed, step in ((0, 0.1), (3, 0.0125)):
        rgba8 = synth.tf_ramps(seed)
        a = tfmod.uncertainty_table(rgba8, step, et)
        b = Oracle().tf_uncertainty(rgba8, step, 1 if et == 1 else 0)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
