"""HIP kernels (through the C ABI) against the CPU oracle on the same inputs.

Bars (stated per test):
  * moments: per-pixel normwise relative error (norm floored at 1e-3, the
    moments are bound-normalised so m_0 is in [0, 1]) <= 1e-5 against the
    reference-faithful FP32 oracle for PLAIN. MIXED evaluates the same
    integral in another order, so it may differ from the FP32 oracle by the
    oracle's own rounding error: per pixel <= 1e-5 + |FP32 oracle - FP64
    oracle|. The 16-bit bounds and the per-ray sample counts are bit-exact;
  * Lagrange multipliers: PLAIN bit-identical to the FP32 oracle; MIXED
    checked through the densities they reconstruct, against the FP64 oracle;
  * float RGBA of the march on given multipliers: PLAIN bit-identical to the
    FP32 oracle; MIXED RMSE <= 1e-4 against the FP64 oracle; rays whose step
    count differs (early ray termination flips on rounding) are counted
    separately.
"""
import ctypes as C
import pathlib

import numpy as np
import pytest

from mray_amd import synth, tf as tfmod
from mray_amd.types import GenParams, RecParams
from oracle_ffi import Oracle, ptr

pytestmark = pytest.mark.gpu

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
_f = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def hip():
    import torch
    from mray_amd import hip as H

    H.init(0)
    return H


@pytest.fixture(scope="module")
def O():
    return Oracle()


def to_np(t):
    return t.detach().cpu().numpy()


def bounds_np(b):
    return to_np(b).view(np.uint32)


def normwise_rel(a, b, floor=1e-3):
    a = a.reshape(-1, a.shape[-1]).astype(np.float64)
    b = b.reshape(-1, b.shape[-1]).astype(np.float64)
    den = np.maximum(np.linalg.norm(b, axis=1), floor)
    return np.linalg.norm(a - b, axis=1) / den


CASES = {
    # name: volume factory, dmin, dmax, camera, W, H, m, gen kwargs
    "ml_ortho_m30": (lambda: synth.marschner_lobb(64), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 48, 40, 30,
                     dict(step_size=2.0 / 128, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                          prediction_coding=0, bias=5e-5)),
    "ml_persp_m40": (lambda: synth.marschner_lobb(48), (-1, -1, -1), (1, 1, 1),
                     synth.perspective_camera(position=(0.15, -0.1, -2.6), fov_half_deg=24.0, znear=0.01,
                                              aspect=44 / 36), 44, 36, 40,
                     dict(step_size=2.0 / 96, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                          prediction_coding=0)),
    # sub-voxel interface (tanh width 0.48 voxel): the MESE at m = 50 is ill-conditioned here
    "mix_persp_m50": (lambda: synth.mixing_layer(48), (0, 0, 0), (1, 1, 1),
                      synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=1.0), 40, 40, 50,
                      dict(step_size=0.5 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                           prediction_coding=0)),
    # resolved interface (tanh width 2.4 voxels)
    "mixw_persp_m50": (lambda: synth.mixing_layer(48, width=0.05), (0, 0, 0), (1, 1, 1),
                       synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=1.0), 40, 40, 50,
                       dict(step_size=0.5 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                            prediction_coding=0)),
    "ml_nobounds_m7": (lambda: synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1),
                       synth.ortho_camera(pos=(0.3, 0.0, -3.0), right=(1.3, 0, 0), up=(0, 1.3, 0)), 21, 13, 7,
                       dict(step_size=2.0 / 50, transform_to_bounds=0, compact_image=0, prediction_coding=0)),
    "u8_ortho_m16": (lambda: np.clip(synth.mixing_layer(32) * 255, 0, 255).astype(np.uint8), (0, 0, 0), (1, 1, 1),
                     synth.ortho_camera(pos=(0.5, 0.5, -1.0), right=(0.6, 0, 0), up=(0, 0.6, 0)), 24, 24, 16,
                     dict(step_size=1.0 / 64, transform_to_bounds=1, compact_image=0, prediction_coding=0)),
    # 64 moments: the mixed generate in two moment chunks
    "ml_ortho_m64": (lambda: synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 20, 16, 64,
                     dict(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                          prediction_coding=0, bias=5e-5)),
    # beyond one thread's registers: moment chunks over workgroups (generate),
    # 20-32 element slots per lane (prepare), four lanes per ray above 128
    # moments (reconstruct)
    "ml_ortho_m100": (lambda: synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 20, 16, 100,
                      dict(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                           prediction_coding=0, bias=5e-5)),
    "ml_persp_m200": (lambda: synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1),
                      synth.perspective_camera(position=(0.15, -0.1, -2.6), fov_half_deg=24.0, znear=0.01,
                                               aspect=22 / 18), 22, 18, 200,
                      dict(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                           prediction_coding=0)),
    "u8_ortho_m255": (lambda: np.clip(synth.mixing_layer(24, width=0.05) * 255, 0, 255).astype(np.uint8), (0, 0, 0),
                      (1, 1, 1), synth.ortho_camera(pos=(0.5, 0.5, -1.0), right=(0.6, 0, 0), up=(0, 0.6, 0)), 17, 15,
                      255, dict(step_size=1.0 / 48, transform_to_bounds=1, compact_image=0, prediction_coding=0)),
    # tile rows dealt out to the XCDs: 136 x 200 rays are 17 x 25 tiles, 107
    # workgroups of 4 tiles, 5 per tile row; two whole rounds of 8 rows' runs
    # of workgroups are remapped, the last 27 workgroups keep their index (the
    # smaller cases here fit no whole round)
    "mixw_persp_m24_rows": (lambda: synth.mixing_layer(48, width=0.05), (0, 0, 0), (1, 1, 1),
                            synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=136 / 200), 136, 200, 24,
                            dict(step_size=0.5 / 48, transform_to_bounds=1, transform_bound_eps=0.01,
                                 compact_image=0, prediction_coding=0)),
    "tiny_m1": (lambda: synth.marschner_lobb(16), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 9, 7, 1,
                dict(step_size=2.0 / 20, transform_to_bounds=1, compact_image=0, prediction_coding=0)),
    "tiny_m2": (lambda: synth.marschner_lobb(16), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 9, 7, 2,
                dict(step_size=2.0 / 20, transform_to_bounds=1, compact_image=0, prediction_coding=0)),
}

_cache = {}


def case_inputs(name):
    if name not in _cache:
        f, dmin, dmax, cam, W, H, m, gk = CASES[name]
        _cache[name] = (f(), dmin, dmax, cam, W, H, m, GenParams(**gk))
    return _cache[name]


def gpu_volume(hip, vol, dmin, dmax):
    import torch

    return hip.Volume(torch.from_numpy(vol).cuda(), dmin, dmax)


# opacity scale of the transfer function per case (the reference's
# preintegration_step_size_factor), chosen for a partly transparent image
TF_FACTOR = {"mixw_persp_m24_rows": 40.0, "ml_ortho_m64": 4.0, "ml_ortho_m30": 4.0, "ml_persp_m40": 4.0, "mix_persp_m50": 40.0, "mixw_persp_m50": 40.0,
             "ml_nobounds_m7": 4.0, "u8_ortho_m16": 20.0, "tiny_m1": 4.0, "tiny_m2": 4.0, "ml_ortho_m100": 4.0,
             "ml_persp_m200": 4.0, "u8_ortho_m255": 20.0}


# MESE bias (lerp toward the moments of a uniform density) per case: the
# sharp mixing-layer and uint8 volumes at 16-50 moments sit so close to the
# boundary of the moment space that at 5e-5 even float storage of exact
# moments moves the image by ~6e-4 RMSE; 1e-3 makes them well posed.
REC_BIAS = {"mixw_persp_m24_rows": 1e-3, "mix_persp_m50": 1e-3, "mixw_persp_m50": 1e-3, "u8_ortho_m16": 1e-3, "u8_ortho_m255": 1e-3,
            "ml_persp_m200": 1e-3}


def rec_params(gp, ert=0.99, name=None):
    return RecParams(step_size=gp.step_size, bias=REC_BIAS.get(name, 5e-5), early_ray_termination=ert,
                     medium_emission=1.0)


def tf_table(name, rp):
    return tfmod.table_from_rgba8(synth.tf_ramps(0), rp.step_size * TF_FACTOR[name])


@pytest.mark.parametrize("form", ["plain", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_generate_parity(hip, O, name, form):
    import torch

    vol, dmin, dmax, cam, W, H, m, gp = case_inputs(name)
    print(f"[generate] case={name} form={form} W={W} H={H} m={m}", flush=True)
    mo_ref, b_ref, ns_ref = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    v = gpu_volume(hip, vol, dmin, dmax)
    mo, b, ns = hip.generate(v, cam, gp, W, H, m, form=form, want_samples=True)
    torch.cuda.synchronize()
    mo = to_np(mo)
    assert np.array_equal(to_np(ns), ns_ref), "sample counts differ"
    if gp.transform_to_bounds:
        assert np.array_equal(bounds_np(b), b_ref), "16-bit bounds differ"
    hit = ns_ref > 0
    err = normwise_rel(mo, mo_ref)[hit]
    if form == "plain":
        assert err.max() <= 1e-5, f"max normwise rel err {err.max():.3e}"
    else:
        mo64, b64 = O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True)
        assert np.array_equal(b_ref, b64)
        own = normwise_rel(mo_ref, mo64)[hit]
        allow = 1e-5 + own
        print(f"  mixed vs fp32: max {err.max():.2e}, pixels over 1e-5: {np.count_nonzero(err > 1e-5)}, "
              f"fp32 own err max {own.max():.2e}")
        assert np.all(err <= allow), f"worst excess {np.max(err - allow):.3e}"
    missed = ~hit
    if missed.any() and not (gp.transform_to_bounds and gp.compact_image):
        assert np.all(mo.reshape(-1, m)[missed] == 0)


@pytest.mark.parametrize("name", ["ml_ortho_m30", "mix_persp_m50", "mixw_persp_m50"])
def test_generate_mixed_closer_to_f64_than_reference(hip, O, name):
    import torch

    vol, dmin, dmax, cam, W, H, m, gp = case_inputs(name)
    mo_ref, b_ref, ns_ref = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    mo64, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True)
    v = gpu_volume(hip, vol, dmin, dmax)
    mo, _, _ = hip.generate(v, cam, gp, W, H, m, form="mixed")
    torch.cuda.synchronize()
    hit = ns_ref > 0
    e_gpu = normwise_rel(to_np(mo), mo64)[hit]
    e_ref = normwise_rel(mo_ref, mo64)[hit]
    print(f"[f64] case={name} gpu-mixed max {e_gpu.max():.2e} median {np.median(e_gpu):.2e} | "
          f"fp32-oracle max {e_ref.max():.2e} median {np.median(e_ref):.2e}")
    assert e_gpu.max() <= 1e-5 + 1.5 * e_ref.max()
    assert np.median(e_gpu) <= 2.0 * np.median(e_ref) + 1e-7


def _b64(O, vol, dmin, dmax, cam, W, H, m, gp):
    return O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True)[1]


def densities(O, m, lag, phases):
    return np.float32([O.lib.orc_evaluate_mese(m, float(p), ptr(np.ascontiguousarray(lag, np.float32), _f))
                       for p in phases])


def densities64(O, m, lag, phases):
    lag = np.ascontiguousarray(lag, np.float64)
    return np.array([O.lib.orc_evaluate_mese_f64(m, float(p), ptr(lag, C.POINTER(C.c_double))) for p in phases])


@pytest.mark.parametrize("form", ["plain", "mixed"])
@pytest.mark.parametrize("name", ["ml_ortho_m30", "mix_persp_m50", "mixw_persp_m50", "tiny_m1", "tiny_m2",
                                  "ml_nobounds_m7", "ml_ortho_m100"])
def test_prepare_parity(hip, O, name, form):
    """Densities of the GPU multipliers against the oracle's, at 41 phases.

    MIXED: against the FP64 oracle, both evaluated in FP64. The kernel returns
    its FP64 multipliers rounded to float, so the bar is the rounding bound
    (|l_0| + 2 sum |l_k|) 2^-24 / pi of the series (the density is
    atan(series)/pi), plus 2e-6 for the FP64 arithmetic itself (a perturbation
    of the moments by m ulps of FP64 moves these densities by < 1e-10).
    PLAIN: the reference's FP32 operations with the host library's
    sinf / cosf restated, so the multipliers of every pixel must be
    bit-identical to the FP32 oracle's."""
    import torch

    vol, dmin, dmax, cam, W, H, m, gp = case_inputs(name)
    mo_ref, b_ref, ns_ref = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    bias = 5e-5
    lag = hip.prepare(torch.from_numpy(mo_ref).cuda(), W, H, m, bias, form=form)
    torch.cuda.synchronize()
    lag = to_np(lag).reshape(-1, m)
    if form == "plain":
        ref = np.stack([O.prepare_std(m, mo_ref.reshape(-1, m)[p], bias) for p in range(W * H)])
        same = (lag.view(np.uint32) == ref.view(np.uint32)).all(axis=1)
        print(f"[prepare] case={name} form=plain bit-identical pixels {same.sum()}/{W * H}")
        assert same.all()
        return
    phases = np.linspace(-np.pi, 0.0, 41)
    excess = []
    hit = np.nonzero(ns_ref > 0)[0]
    for p in hit[:: max(1, len(hit) // 96)]:
        mom = mo_ref.reshape(-1, m)[p]
        l64 = O.prepare_std_f64(m, mom.astype(np.float64), bias)
        r64 = densities64(O, m, l64, phases)
        bound = (abs(l64[0]) + 2.0 * np.abs(l64[1:]).sum()) * 2.0 ** -24 / np.pi * 1.01
        err = np.abs(densities64(O, m, lag[p], phases) - r64).max()
        excess.append(err - (bound + 2e-6))
    excess = np.array(excess) if excess else np.zeros(1)
    print(f"[prepare] case={name} form={form} worst excess over allowance {excess.max():.2e}")
    assert excess.max() <= 0.0


@pytest.mark.parametrize("m", [64, 72, 80, 88, 100, 128, 160, 200, 248, 255])
def test_prepare_mixed_large_m(hip, O, m):
    """MIXED preparation beyond the parity cases' 50 moments, up to 255 (16
    element slots per lane up to 128, then 20-32; at 10, 11, 13, 20 and 32
    slots, 80, 88, 100, 160 and 255 moments here, the per-pixel LDS arrays
    have no bank padding): densities against the FP64
    oracle at 41 phases with the float-storage bound of test_prepare_parity.
    Moments from the oracle's generate."""
    import torch

    vol = synth.marschner_lobb(32)
    dmin, dmax = (-1, -1, -1), (1, 1, 1)
    cam = synth.ortho_camera()
    W, H = 12, 10
    gp = GenParams(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                   prediction_coding=0)
    print(f"[prepare large m] m={m} W={W} H={H}", flush=True)
    mo_ref, _, ns_ref = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    bias = 5e-5
    lag = to_np(hip.prepare(torch.from_numpy(mo_ref).cuda(), W, H, m, bias, form="mixed")).reshape(-1, m)
    phases = np.linspace(-np.pi, 0.0, 41)
    worst = -1.0
    for p in np.nonzero(ns_ref > 0)[0][::4]:
        l64 = O.prepare_std_f64(m, mo_ref.reshape(-1, m)[p].astype(np.float64), bias)
        r64 = densities64(O, m, l64, phases)
        bound = (abs(l64[0]) + 2.0 * np.abs(l64[1:]).sum()) * 2.0 ** -24 / np.pi * 1.01
        err = np.abs(densities64(O, m, lag[p], phases) - r64).max()
        worst = max(worst, err - (bound + 2e-6))
    print(f"  worst excess over allowance {worst:.2e}")
    assert worst <= 0.0


@pytest.mark.parametrize("form", ["plain", "mixed"])
@pytest.mark.parametrize("name", ["ml_ortho_m30", "ml_persp_m40", "mix_persp_m50", "mixw_persp_m50", "u8_ortho_m16",
                                  "tiny_m2", "ml_ortho_m64", "ml_ortho_m100", "ml_persp_m200", "u8_ortho_m255",
                                  "mixw_persp_m24_rows"])
def test_reconstruct_parity(hip, O, name, form):
    """The march alone, on the FP32 oracle's multipliers: RMSE against the
    FP64 evaluation of the same multipliers <= 1e-4, or no worse than 1.5x the
    FP32 reference evaluation where that itself exceeds 1e-4 (large
    multipliers near the boundary of the moment space). PLAIN must also be
    bit-identical to the FP32 oracle."""
    import torch

    vol, dmin, dmax, cam, W, H, m, gp = case_inputs(name)
    mo_ref, b_ref, ns_ref = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    rp = rec_params(gp, name=name)
    lag_ref = np.stack([O.prepare_std(m, mo_ref.reshape(-1, m)[p], rp.bias) for p in range(W * H)])
    table = tf_table(name, rp)
    bnd = b_ref if gp.transform_to_bounds else None
    # the FP32 oracle prepares the moments itself (same FP32 arithmetic as lag_ref)
    rgba32, st32 = O.reconstruct(cam, dmin, dmax, W, H, m, mo_ref.reshape(-1), table, rp, bounds=bnd)
    rgba64, st64 = O.reconstruct_lagrange_f64(cam, dmin, dmax, W, H, m, lag_ref, table, rp, bounds=bnd)
    print(f"[reconstruct] case={name} form={form}", flush=True)
    lag_d = torch.from_numpy(np.ascontiguousarray(lag_ref, np.float32)).cuda()
    b_d = torch.from_numpy(bnd.view(np.int32)).cuda() if bnd is not None else None
    rgba, st = hip.reconstruct(cam, dmin, dmax, lag_d, W, H, m, torch.from_numpy(table), rp, form=form, bounds=b_d,
                               want_steps=True)
    torch.cuda.synchronize()
    rgba, st = to_np(rgba), to_np(st)
    same = (st == st64) & (st32 == st64)
    finite = np.isfinite(rgba64).all(axis=-1).reshape(-1) & same
    rms = lambda d: float(np.sqrt((d[finite] ** 2).mean()))
    r_gpu = rms((rgba - rgba64).reshape(-1, 4))
    r_32 = rms((rgba32 - rgba64).reshape(-1, 4))
    print(f"  vs fp64 evaluation: gpu RMSE {r_gpu:.2e}, fp32 oracle RMSE {r_32:.2e}, "
          f"rays with other step count {np.count_nonzero(~same)}/{W * H}")
    assert np.count_nonzero(~same) <= max(2, 0.01 * W * H)
    assert r_gpu <= 1e-4 or r_gpu <= 1.5 * r_32
    if form == "plain":
        # same FP32 operations as the reference (host sinf/cosf restated, exp
        # in double rounded to float): bit-identical rays and step counts
        ident = (rgba.view(np.uint32) == rgba32.view(np.uint32)).all(axis=-1).reshape(-1)
        print(f"  plain vs fp32 oracle: bit-identical rays {ident.sum()}/{W * H}, "
              f"max |diff| {np.abs(rgba - rgba32).max():.2e}")
        assert np.array_equal(st, st32)
        assert ident.all()


@pytest.mark.parametrize("name", ["ml_ortho_m30", "mixw_persp_m50", "ml_persp_m40", "mix_persp_m50", "ml_ortho_m100",
                                  "ml_persp_m200"])
def test_pipeline_mixed_vs_f64(hip, O, name):
    """Final float RGBA of the mixed pipeline against the FP64 oracle: RMSE <= 1e-4
    where the reference's own FP32 pipeline meets that; on the ill-conditioned
    sub-voxel case the mixed pipeline must be at least as accurate as it."""
    import torch

    vol, dmin, dmax, cam, W, H, m, gp = case_inputs(name)
    rp = rec_params(gp, name=name)
    mo64, b64 = O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True)
    mo32, b32, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    table = tf_table(name, rp)
    rgba64, st64 = O.reconstruct_f64(cam, dmin, dmax, W, H, m, mo64, table, rp, bounds=b64)
    rgba32, st32 = O.reconstruct(cam, dmin, dmax, W, H, m, mo32.reshape(-1), table, rp, bounds=b32)
    v = gpu_volume(hip, vol, dmin, dmax)
    mo, b, _ = hip.generate(v, cam, gp, W, H, m, form="mixed")
    lag = hip.prepare(mo, W, H, m, rp.bias, form="mixed")
    rgba, st = hip.reconstruct(cam, dmin, dmax, lag, W, H, m, torch.from_numpy(table), rp, form="mixed", bounds=b,
                               want_steps=True)
    torch.cuda.synchronize()
    rgba, st = to_np(rgba), to_np(st)
    same = (st == st64) & (st32 == st64) & (bounds_np(b) == b64)
    d = (rgba - rgba64).reshape(-1, 4)
    d32 = (rgba32 - rgba64).reshape(-1, 4)
    rmse = float(np.sqrt((d[same] ** 2).mean()))
    rmse32 = float(np.sqrt((d32[same] ** 2).mean()))
    print(f"[pipeline mixed vs f64] case={name} RMSE {rmse:.2e} max {np.abs(d[same]).max():.2e} | "
          f"fp32 reference RMSE {rmse32:.2e} | alpha mean {rgba64[..., 3].mean():.2f} | "
          f"separate rays {np.count_nonzero(~same)}")
    if rmse32 <= 1e-4:
        assert rmse <= 1e-4
    else:
        # the sharp mixing-layer cases at 50 moments: the FP32 reference
        # pipeline is measured 1.5e-3 (mixw) and 2.2e-3 (mix) RMSE from the
        # FP64 oracle on the CPU, whose own functions are pinned independently
        # (tests/test_oracle_f64.py); the relaxed bar applies only that far
        assert rmse32 <= 4e-3, f"FP32 oracle {rmse32:.2e} from the FP64 oracle, beyond its measured distance"
        assert rmse <= rmse32


@pytest.mark.parametrize("name", ["ml_ortho_m30", "mixw_persp_m50", "mixw_persp_m24_rows"])
def test_pipeline_plain_vs_f32(hip, O, name):
    import torch

    vol, dmin, dmax, cam, W, H, m, gp = case_inputs(name)
    rp = rec_params(gp, name=name)
    mo_ref, b_ref, _ = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    table = tf_table(name, rp)
    rgba_ref, st_ref = O.reconstruct(cam, dmin, dmax, W, H, m, mo_ref.reshape(-1), table, rp, bounds=b_ref)
    v = gpu_volume(hip, vol, dmin, dmax)
    mo, b, _ = hip.generate(v, cam, gp, W, H, m, form="plain")
    lag = hip.prepare(mo, W, H, m, rp.bias, form="plain")
    rgba, st = hip.reconstruct(cam, dmin, dmax, lag, W, H, m, torch.from_numpy(table), rp, form="plain", bounds=b,
                               want_steps=True)
    torch.cuda.synchronize()
    rgba, st = to_np(rgba), to_np(st)
    mo64, b64 = O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True)
    rgba64, st64 = O.reconstruct_f64(cam, dmin, dmax, W, H, m, mo64, table, rp, bounds=b64)
    same = (st == st_ref) & (st64 == st_ref)
    d = np.abs(rgba - rgba_ref).reshape(-1, 4)[same]
    own = float(np.sqrt((((rgba_ref - rgba64).reshape(-1, 4)[same]) ** 2).mean()))
    rms = float(np.sqrt((d ** 2).mean()))
    print(f"[pipeline plain vs f32] case={name} max {d.max():.2e} rms {rms:.2e} (fp32 reference vs fp64 {own:.2e})")
    # the plain form repeats the reference's FP32 arithmetic: it may differ from
    # the FP32 oracle by that arithmetic's own rounding noise
    assert rms <= 1e-4 + 2 * own


def test_golden_images_on_gpu(hip, O):
    import torch

    g = dict(np.load(GOLD / "golden_images.npz"))
    case = "ml_ortho"
    W, H, m = 24, 20, 12
    gp = GenParams(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                   prediction_coding=0, bias=5e-5)
    v = gpu_volume(hip, g[f"{case}_volume"], (-1, -1, -1), (1, 1, 1))
    for form in ("plain", "mixed"):
        mo, b, _ = hip.generate(v, synth.ortho_camera(), gp, W, H, m, form=form)
        torch.cuda.synchronize()
        assert np.array_equal(bounds_np(b), g[f"{case}_bounds"])
        err = normwise_rel(to_np(mo), g[f"{case}_data"].reshape(H, W, m))
        assert err.max() <= 1e-5


def test_error_paths(hip):
    import torch

    from mray_amd.hip import MrayError

    vol = torch.zeros((4, 4, 4), dtype=torch.float32, device="cuda")
    v = hip.Volume(vol, (0, 0, 0), (1, 1, 1))
    with pytest.raises(MrayError):
        hip.generate(v, synth.ortho_camera(), GenParams(step_size=0.0), 4, 4, 4)
    with pytest.raises(MrayError):
        hip.generate(v, synth.ortho_camera(), GenParams(step_size=0.1), 4, 4, 300)
    with pytest.raises(MrayError):
        hip.Volume(vol, (0, 0, 0), (0, 1, 1))
