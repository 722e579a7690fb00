"""Error bounds and uncertainty rendering on the GPU (mray_hip_error_bounds,
mray_hip_reconstruct_uncertainty), through the C ABI:

* the per-pixel uint16 error bounds (maximum, RMSE, P^2 90th percentile)
  bit-identical to the compiled reference's reconstruct_error
  (golden_uncertainty.npz) on compact, prediction-coded and dense images of an
  f32 and a u8 volume, including rays of fewer than five samples;
* the float RGBA of the interpolation and of the convolution (per error type)
  bit-identical to the reference's renderings, with the uncertainty tables of
  the C++ host library;
* larger images, a particle set and m = 100 against the oracle, bit for bit;
* the CLI's --error-bounds and --uncertainty-* modes end to end.
"""
import pathlib
import subprocess

import numpy as np
import pytest

from mray_amd import synth, tf as tfmod
from mray_amd.types import ErrParams, GenParams, RecParams
from oracle_ffi import Oracle, Particles
from test_uncertainty_oracle import CASES, case_inputs

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
ROOT = pathlib.Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def env():
    import torch

    from mray_amd import hip

    hip.init(0)
    return hip, torch, np.load(GOLD / "golden_uncertainty.npz")


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def plain_multipliers(hip, torch, c, bias=1e-4):
    data = dev(torch, c["data"])
    index = None if c["index"] is None else dev(torch, c["index"].view(np.int32))
    if c["pc"]:
        lag = hip.prepare_coded(data, c["W"], c["H"], c["m"], 0, dev(torch, c["cp"].view(np.int16)), index=index)
    else:
        lag = hip.prepare(data, c["W"], c["H"], c["m"], bias, form="plain", index=index)
    return lag, index


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("et", [0, 1, 2])
def test_error_bounds_match_reference(env, name, et):
    hip, torch, gold = env
    c = case_inputs(gold, name)
    print(f"[error-bounds] {name} type={et}", flush=True)
    lag, index = plain_multipliers(hip, torch, c)
    vol = hip.Volume(dev(torch, c["vol"]), c["dmin"], c["dmax"])
    ep = ErrParams(error_type=et, error_percentile=90 if et == 2 else 95, bias=1e-4, step_size=c["estep"])
    e = hip.error_bounds(vol, c["cam"], lag, c["W"], c["H"], c["m"], ep, index=index,
                         bounds=dev(torch, c["bounds"].view(np.int32)))
    torch.cuda.synchronize()
    assert np.array_equal(e.cpu().numpy().view(np.uint16), gold[f"{name}_err{et}"])


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("mode,et,s", [(0, 2, 0.3), (1, 0, 0.0), (1, 1, 0.0), (1, 2, 0.0)])
def test_uncertainty_images_match_reference(env, name, mode, et, s):
    hip, torch, gold = env
    c = case_inputs(gold, name)
    print(f"[uncertainty] {name} mode={mode} type={et}", flush=True)
    lag, index = plain_multipliers(hip, torch, c)
    rp = RecParams(step_size=1.0 / 40, bias=1e-4, early_ray_termination=0.99, medium_emission=1.0)
    step = rp.step_size * 4
    tf = tfmod.table_from_rgba8(synth.tf_ramps(0), step) if mode == 0 else \
        tfmod.uncertainty_table(synth.tf_ramps(0), step, et)
    rgba = hip.reconstruct_uncertainty(c["cam"], c["dmin"], c["dmax"], lag,
                                       dev(torch, gold[f"{name}_err{et}"].view(np.int16)), c["W"], c["H"], c["m"],
                                       dev(torch, tf), rp, mode, et, s, index=index,
                                       bounds=dev(torch, c["bounds"].view(np.int32)))
    torch.cuda.synchronize()
    assert np.array_equal(rgba.cpu().numpy().view(np.uint32), gold[f"{name}_unc{mode}{et}"].view(np.uint32))


@pytest.mark.parametrize("m,et", [(40, 2), (100, 1), (100, 0)])
def test_larger_images_match_oracle(env, m, et):
    hip, torch, _ = env
    O = Oracle()
    W, H = 96, 80
    vol, dmin, dmax = synth.marschner_lobb(48), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    cam = synth.perspective_camera(position=(0.2, -0.1, -2.6), fov_half_deg=24.0, znear=0.01, aspect=W / H)
    gp = GenParams(step_size=1.0 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0, bias=1e-4)
    print(f"[uncertainty] oracle {W}x{H} m={m} type={et}", flush=True)
    mo, b, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    ep = ErrParams(error_type=et, error_percentile=90, bias=1e-4, step_size=1.0 / 40)
    e_ref = O.error_bounds(vol, dmin, dmax, cam, W, H, m, mo.reshape(-1), ep, bounds=b)
    d_mo = dev(torch, mo)
    lag = hip.prepare(d_mo, W, H, m, 1e-4, form="plain")
    v = hip.Volume(dev(torch, vol), dmin, dmax)
    e = hip.error_bounds(v, cam, lag, W, H, m, ep, bounds=dev(torch, b.view(np.int32)))
    torch.cuda.synchronize()
    assert np.array_equal(e.cpu().numpy().view(np.uint16), e_ref)
    rp = RecParams(step_size=1.0 / 40, bias=1e-4, early_ray_termination=0.99, medium_emission=1.0)
    tf2 = tfmod.uncertainty_table(synth.tf_ramps(1), rp.step_size * 8, et)
    ref = O.uncertainty(cam, dmin, dmax, W, H, m, mo.reshape(-1), e_ref, tf2, rp, 1, et, 0.0, bounds=b)
    got = hip.reconstruct_uncertainty(cam, dmin, dmax, lag, e, W, H, m, dev(torch, tf2), rp, 1, et,
                                      bounds=dev(torch, b.view(np.int32)))
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_particle_error_bounds_match_oracle(env):
    hip, torch, _ = env
    O = Oracle()
    n = 4000
    x, y, z, val = synth.particle_mixture(n, seed=3, n_gauss=5, box=(1.0, 1.0, 1.0))
    h = 0.08
    dmin, dmax = (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    parts = Particles(O, x, y, z, val, h)
    W, H, m = 24, 20, 12
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.2), fov_half_deg=28.0, znear=0.01, aspect=W / H)
    gp = GenParams(step_size=1.0 / 40, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0, bias=1e-4)
    print(f"[error-bounds] particles n={n}", flush=True)
    mo, b, _ = O.generate(parts, dmin, dmax, cam, W, H, m, gp)
    ep = ErrParams(error_type=2, error_percentile=90, bias=1e-4, step_size=1.0 / 30)
    e_ref = O.error_bounds(parts, dmin, dmax, cam, W, H, m, mo.reshape(-1), ep, bounds=b)
    pv = hip.ParticleVolume(*(dev(torch, np.ascontiguousarray(a, np.float32)) for a in (x, y, z, val)), h, dmin, dmax)
    lag = hip.prepare(dev(torch, mo), W, H, m, 1e-4, form="plain")
    e = hip.error_bounds(pv, cam, lag, W, H, m, ep, bounds=dev(torch, b.view(np.int32)))
    torch.cuda.synchronize()
    assert np.array_equal(e.cpu().numpy().view(np.uint16), e_ref)


def test_cli_error_bounds_and_uncertainty(env, tmp_path):
    from test_gpu_cli import camera_file, run

    hip, torch, _ = env
    O = Oracle()
    vol, dmin, dmax = synth.marschner_lobb(40), (-1, -1, -1), (1, 1, 1)
    cam = synth.ortho_camera()
    W, H, m, step = 28, 24, 16, 2.0 / 80
    vol.tofile(tmp_path / "vol.raw")
    camera_file(tmp_path / "cam.txt", cam)
    rgba8 = synth.tf_ramps(0)
    rgba8.tofile(tmp_path / "tf.bin")
    gen = (f"width={W},height={H},num_moments={m},camera={tmp_path / 'cam.txt'},dims=40:40:40,"
           f"domain=-1:-1:-1:1:1:1,step_size={step!r},transform_to_bounds=1,transform_bound_eps=0.01,"
           f"bias=1e-4,form=plain,compact_image=1,prediction_coding=0,error_threshold=0.001")
    run("--generate", tmp_path / "mi.bin", tmp_path / "vol.raw", gen)
    from mi_file import read_moment_image

    before = read_moment_image(tmp_path / "mi.bin")
    run("--error-bounds", tmp_path / "mi.bin", tmp_path / "vol.raw",
        f"camera={tmp_path / 'cam.txt'},dims=40:40:40,domain=-1:-1:-1:1:1:1,error_type=2,error_percentile=90,"
        f"bias=1e-4,step_size=0.03")
    raw = (tmp_path / "mi.bin").read_bytes()
    assert raw[6] & 32  # the error-bounds flag
    npix = W * H
    off = 32 + npix + 4 * npix
    eb = np.frombuffer(raw, np.uint16, npix, off)
    ep = ErrParams(error_type=2, error_percentile=90, bias=1e-4, step_size=0.03)
    e_ref = O.error_bounds(vol, dmin, dmax, cam, W, H, m, before["data"], ep, index=before["index"],
                           bounds=before["bounds"])
    assert np.array_equal(eb, e_ref)
    for mode, flag in ((0, "--uncertainty-interpolate"), (1, "--uncertainty-convolution")):
        run(flag, tmp_path / "mi.bin", tmp_path / f"u{mode}.ppm",
            f"tf={tmp_path / 'tf.bin'},camera={tmp_path / 'cam.txt'},step_size={step!r},bias=1e-4,"
            f"early_ray_termination=0.99,medium_emission=1.0,preintegration_step_size_factor=4.0,error_type=2,"
            f"bound_interpolation=0.25,form=plain,float_rgba={tmp_path / f'u{mode}.f32'}")
        got = np.fromfile(tmp_path / f"u{mode}.f32", np.float32).reshape(H, W, 4)
        rp = RecParams(step_size=step, bias=1e-4, early_ray_termination=0.99, medium_emission=1.0)
        tf = O.tf(rgba8, step * 4.0) if mode == 0 else O.tf_uncertainty(rgba8, step * 4.0, 2)
        ref = O.uncertainty(cam, dmin, dmax, W, H, m, before["data"], e_ref, tf, rp, mode, 2, 0.25,
                            index=before["index"], bounds=before["bounds"])
        print(f"[cli] uncertainty mode {mode}: alpha mean {ref[..., 3].mean():.3f}", flush=True)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
