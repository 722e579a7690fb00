"""The CPU oracle against the reference's golden vectors (bit-exact) and
against closed forms that pin what the golden vectors leave open.

Golden vectors come from the compiled reference (tests/golden/make_golden.py);
the live reference is also compared when oracle/_ref/libmrayref.so exists.
"""
import ctypes as C
import math
import pathlib

import numpy as np
import pytest

from mray_amd import synth, tf as tfmod
from mray_amd.types import GenParams, RecParams, RefParams
from oracle_ffi import Oracle, Reference, ptr, REF_SO

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
_f = C.POINTER(C.c_float)
_u32 = C.POINTER(C.c_uint32)


@pytest.fixture(scope="module")
def O():
    return Oracle()


@pytest.fixture(scope="module")
def mese():
    return dict(np.load(GOLD / "golden_mese.npz"))


@pytest.fixture(scope="module")
def coding():
    return dict(np.load(GOLD / "golden_coding.npz"))


@pytest.fixture(scope="module")
def images():
    return dict(np.load(GOLD / "golden_images.npz"))


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


# ---------------------------------------------------------------- MESE blocks
@pytest.mark.parametrize("name", ["trig4", "trig10", "trig28", "comp10"])
def test_mese_blocks_match_reference(O, mese, name):
    t = mese[f"{name}_in"]
    m = len(t)
    e = np.zeros(2 * m, np.float32)
    O.lib.orc_trig_to_exp(m, ptr(t, _f), ptr(e, _f))
    assert bits_equal(e, mese[f"{name}_exp"])
    back = np.zeros(2 * m, np.float32)
    O.lib.orc_exp_to_trig(m, ptr(e, _f), ptr(back, _f))
    assert bits_equal(back, mese[f"{name}_exp_to_trig"])
    lev = np.zeros(2 * m, np.float32)
    O.lib.orc_levinson(m, ptr(e, _f), ptr(lev, _f))
    assert bits_equal(lev, mese[f"{name}_levinson"])
    pm = np.zeros(m, np.float32)
    O.lib.orc_prepare_mese(m, ptr(t, _f), ptr(pm, _f))
    assert bits_equal(pm, mese[f"{name}_prepare"])
    ev = [O.lib.orc_evaluate_mese(m, float(p), ptr(pm, _f)) for p in mese[f"{name}_phases"]]
    assert bits_equal(np.float32(ev), mese[f"{name}_evaluate"])
    for bias in (1e-4, 5e-5):
        assert bits_equal(O.prepare_std(m, t, bias), mese[f"{name}_prepare_bias{bias:g}"])


def test_mese_complex_and_correlations(O, mese):
    ec = np.zeros(8, np.float32)
    O.lib.orc_trig_to_exp_c(4, ptr(mese["trigc4_in"], _f), ptr(ec, _f))
    assert bits_equal(ec, mese["trigc4_exp"])
    lc = np.zeros(8, np.float32)
    O.lib.orc_prepare_mese_c(4, ptr(mese["prepc4_in"], _f), ptr(lc, _f))
    assert bits_equal(lc, mese["prepc4_out"])
    s, r = mese["sig4"], mese["rhs4"]
    ac, cc, ic = np.zeros(8, np.float32), np.zeros(8, np.float32), np.zeros(4, np.float32)
    O.lib.orc_autocorrelation(4, ptr(s, _f), ptr(ac, _f))
    O.lib.orc_correlation(4, ptr(s, _f), ptr(r, _f), ptr(cc, _f))
    O.lib.orc_imag_correlation(4, ptr(s, _f), ptr(r, _f), ptr(ic, _f))
    assert bits_equal(ac, mese["autocorr4"]) and bits_equal(cc, mese["corr4"]) and bits_equal(ic, mese["imagcorr4"])
    out = np.float32([O.lib.orc_fast_atan(float(x)) for x in mese["atan_in"]])
    assert bits_equal(out, mese["atan_out"])


def test_fixed_size_mese_agrees_like_reference_tests(O, mese):
    # tests/test_mese.cpp:31-59: the dynamic MESE equals the unrolled m = 4, 10 versions
    p = np.float32(-np.pi + np.pi * 0.5)
    v = O.lib.orc_evaluate_mese(4, float(p), ptr(mese["mom4_fixed_prepare"], _f))
    assert abs(v - mese["mom4_fixed_eval"][0]) <= 4 * np.spacing(np.float32(v))
    v10 = O.lib.orc_evaluate_mese(10, float(p), ptr(mese["trig10_fixed_prepare"], _f))
    assert abs(v10 - mese["trig10_fixed_eval"][0]) <= 4 * np.spacing(np.float32(v10))
    e = np.zeros(8, np.float32)
    O.lib.orc_trig_to_exp(4, ptr(mese["trig4_in"], _f), ptr(e, _f))
    assert np.allclose(e, mese["trig4_fixed_exp"], rtol=1e-6, atol=1e-7)


# ------------------------------------------------------- closed-form pinning
def test_levinson_solves_toeplitz_system(O):
    # Levinson-Durbin on the Hermitian Toeplitz matrix of the exponential
    # moments, checked against a dense solve.
    rng = np.random.default_rng(0)
    m = 12
    t = np.zeros(m, np.float32)
    t[0] = 0.4
    t[1:] = (rng.standard_normal(m - 1) * 0.02).astype(np.float32)
    e = np.zeros(2 * m, np.float32)
    O.lib.orc_trig_to_exp(m, ptr(t, _f), ptr(e, _f))
    sol = np.zeros(2 * m, np.float32)
    O.lib.orc_levinson(m, ptr(e, _f), ptr(sol, _f))
    g = e[0::2].astype(np.float64) + 1j * e[1::2]
    a = sol[0::2].astype(np.float64) + 1j * sol[1::2]
    # Hermitian Toeplitz matrix with first column g and real diagonal Re(g_0)
    # (the recursion starts from 1 / Re(g_0)); it solves T a = e_0.
    g[0] = g[0].real
    T = np.array([[g[i - k] if i >= k else np.conj(g[k - i]) for k in range(m)] for i in range(m)])
    ref = np.linalg.solve(T, np.eye(m)[0].astype(complex))
    assert np.abs(a - ref).max() < 1e-5 * np.abs(ref).max()


def test_trig_to_exp_matches_series_exponential(O):
    # gamma(z) = gamma_0 * exp(2 pi sum_l t_l z^l): the recurrence is the power series of exp
    m = 10
    t = np.float32([0.3, 0.05, -0.02, 0.01, 0.004, -0.003, 0.002, 0.0, 0.001, -0.0005])
    e = np.zeros(2 * m, np.float32)
    O.lib.orc_trig_to_exp(m, ptr(t, _f), ptr(e, _f))
    g = e[0::2].astype(np.float64) + 1j * e[1::2]
    ph = math.pi * t[0] - math.pi / 2
    g0 = np.exp(1j * ph) / (4 * math.pi)
    # power series of exp(2 pi i sum_{l>=1} t_l z^l)
    coef = np.zeros(m, complex)
    coef[0] = 1.0
    s = 2j * math.pi * t.astype(np.float64)
    s[0] = 0
    for l in range(1, m):
        coef[l] = sum((l - j) * s[l - j] * coef[j] for j in range(l)) / l
    expect = g0 * coef
    expect[0] *= 2
    assert np.allclose(g, expect, rtol=1e-5, atol=1e-7)


def test_constant_density_gives_constant_moments(O):
    # constant rho: m0 = rho, mj = 0 (closed form of the integral)
    n = 16
    rho = np.float32(0.37)
    vol = np.full((n, n, n), rho, np.float32)
    gp = GenParams(step_size=1.0 / 40, compact_image=0, prediction_coding=0)
    mo, b, ns = O.generate(vol, (0, 0, 0), (1, 1, 1), synth.ortho_camera(pos=(0.5, 0.5, -1), right=(0.4, 0, 0),
                                                                           up=(0, 0.4, 0)), 6, 5, 8, gp)
    assert np.allclose(mo[..., 0], rho, atol=2e-6)
    assert np.abs(mo[..., 1:]).max() < 2e-6
    mo64, _ = O.generate(vol, (0, 0, 0), (1, 1, 1), synth.ortho_camera(pos=(0.5, 0.5, -1), right=(0.4, 0, 0),
                                                                         up=(0, 0.4, 0)), 6, 5, 8, gp, f64=True)
    # the first node sits at the reference's float phase p_1 - dphi, a few ulp from -pi
    assert np.allclose(mo64[..., 0], rho, atol=2e-6) and np.abs(mo64[..., 1:]).max() < 2e-6


def test_linear_ramp_closed_form(O):
    # f(phi) = a phi + b on [-pi, 0] => m_j = a (1 - (-1)^j) / (pi j^2), m_0 = b - a pi / 2
    n = 64
    z = synth.voxel_coords(n, 0.0, 1.0)
    vol = np.broadcast_to(z[:, None, None], (n, n, n)).astype(np.float32).copy()
    cam = synth.ortho_camera(pos=(0.5, 0.5, -1.0), right=(0.3, 0, 0), up=(0, 0.3, 0))
    gp = GenParams(step_size=1.0 / 128, compact_image=0, prediction_coding=0)
    mo64, _ = O.generate(vol, (0, 0, 0), (1, 1, 1), cam, 4, 4, 9, gp, f64=True)
    # the interpolant is exactly linear in z on [0, (n-1)/n] and clamped beyond: integrate that exactly
    mo = mo64[2, 2]
    zz = np.linspace(0.0, 1.0, 200001)
    f = np.minimum(zz, (n - 1) / n)
    phi = zz * math.pi - math.pi
    for j in range(9):
        ref = np.trapezoid(f * np.cos(j * phi), phi) / math.pi
        assert abs(mo[j] - ref) < 2e-6, (j, mo[j], ref)


def test_compositing_matches_direct_loop(O):
    # emission-absorption with constant density: L = e * c_rgb * sum_k T_k, T_k = exp(-k sigma)
    m = 6
    W = H = 2
    lag = np.zeros((H, W, m), np.float32)  # lagrange multipliers 0 => density 0.5 everywhere
    tf8 = np.zeros((256, 4), np.uint8)
    tf8[:, :3] = 200
    tf8[:, 3] = 40
    step = 1.0 / 64
    table = O.tf(tf8, step)
    rp = RecParams(step_size=step, early_ray_termination=2.0, medium_emission=1.0)
    cam = synth.ortho_camera(pos=(0.5, 0.5, -1.0), right=(0.3, 0, 0), up=(0, 0.3, 0))
    rgba, ns = O.reconstruct(cam, (0, 0, 0), (1, 1, 1), W, H, m, lag.reshape(-1), table, rp)
    lagp = np.zeros(m, np.float32)
    rho = O.lib.orc_evaluate_mese(m, -1.0, ptr(lagp, _f))
    assert abs(rho - 0.5) < 2e-5
    c = np.zeros(4, np.float32)
    O.lib.orc_tf_sample(ptr(table, _f), C.c_float(rho), C.c_float(rho), ptr(c, _f))
    T, L = 1.0, 0.0
    for _ in range(int(ns[0])):
        L += c[0] * T
        T = math.exp(-(_ + 1) * c[3])
    assert abs(rgba[0, 0, 0] - L) < 1e-5 and abs(rgba[0, 0, 3] - (1 - T)) < 1e-5


def test_tf_table_product_matches_oracle(O):
    tf8 = synth.tf_ramps(0)
    for step in (1.0 / 128, 0.005 * 50, 1.0):
        assert bits_equal(tfmod.table_from_rgba8(tf8, step), O.tf(tf8, step))


# -------------------------------------------------------- image-level golden
@pytest.mark.parametrize("case", ["ml_ortho", "ml_persp_plain", "u8_ortho"])
def test_generate_and_reconstruct_match_reference(O, images, case):
    params = {
        "ml_ortho": ((-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 24, 20, 12,
                     dict(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                          prediction_coding=0, bias=5e-5),
                     dict(step_size=2.0 / 64, bias=5e-5, early_ray_termination=0.99, medium_emission=1.0)),
        "ml_persp_plain": ((-1, -1, -1), (1, 1, 1),
                           synth.perspective_camera(position=(0.1, -0.05, -3.0), fov_half_deg=20.0, znear=0.01,
                                                    aspect=20 / 16), 20, 16, 10,
                           dict(step_size=2.0 / 48, compact_image=0, prediction_coding=0),
                           dict(step_size=2.0 / 48, early_ray_termination=0.95)),
        "u8_ortho": ((0, 0, 0), (1, 1, 1), synth.ortho_camera(pos=(0.5, 0.5, -1.0), right=(0.5, 0, 0),
                                                               up=(0, 0.5, 0)), 16, 16, 8,
                     dict(step_size=1.0 / 32, transform_to_bounds=1, compact_image=0, prediction_coding=0),
                     dict(step_size=1.0 / 32, early_ray_termination=0.9)),
    }[case]
    dmin, dmax, cam, W, H, m, gk, rk = params
    gp, rp = GenParams(**gk), RecParams(**rk)
    vol = images[f"{case}_volume"]
    mo, b, ns = O.generate(vol, dmin, dmax, cam, W, H, m, gp)
    assert bits_equal(mo.reshape(-1), images[f"{case}_data"])
    if gp.transform_to_bounds:
        assert np.array_equal(b, images[f"{case}_bounds"])
    table = O.tf(images["tf_rgba8"], rp.step_size)
    rgba, _ = O.reconstruct(cam, dmin, dmax, W, H, m, mo.reshape(-1), table, rp,
                            bounds=b if gp.transform_to_bounds else None)
    assert bits_equal(rgba, images[f"{case}_rgba"])
    q = (np.clip(rgba, 0, 1) * np.float32(255)).astype(np.uint8)
    assert np.array_equal(q, images[f"{case}_rgba8"])


def test_coded_pipeline_matches_reference(O, images):
    # generate with compaction + prediction coding, then reconstruct from the coded image
    case = "mix_coded"
    W, H, m = 16, 12, 16
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=16 / 12)
    gp = GenParams(step_size=1.0 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=1,
                   prediction_coding=1, error_threshold=0.01, bias=1e-4)
    rp = RecParams(step_size=1.0 / 48, bias=1e-4, early_ray_termination=0.98, medium_emission=1.0)
    vol = images[f"{case}_volume"]
    mo, b, ns = O.generate(vol, (0, 0, 0), (1, 1, 1), cam, W, H, m, gp)
    assert np.array_equal(b, images[f"{case}_bounds"])
    counts = O.num_moments(W, H, m, mo, b, gp.error_threshold)
    index, comp = O.compact(W, H, m, mo, counts)
    assert np.array_equal(index, images[f"{case}_index"])
    coded, q = O.prediction_encode(W, H, m, comp, index, gp.coding_warp, gp.bias)
    assert bits_equal(coded, images[f"{case}_data"])
    assert np.array_equal(q, images[f"{case}_coding_params"])
    table = O.tf(images["tf_rgba8"], rp.step_size)
    params_f = (q.astype(np.float32) / np.float32(65535.0)).astype(np.float32)
    rgba, _ = O.reconstruct(cam, (0, 0, 0), (1, 1, 1), W, H, m, coded, table, rp, index=index, bounds=b,
                            prediction_coded=1, coding_warp=0, coding_params=params_f)
    assert bits_equal(rgba, images[f"{case}_rgba"])


def test_coding_blocks_match_reference(O, coding):
    for name in ("pc5", "pc5b", "comp10", "trig28"):
        t = coding[f"{name}_in"]
        m = len(t)
        e = coding[f"{name}_exp"]
        code = np.zeros(2 * m, np.float32)
        O.lib.orc_encode(m, ptr(e, _f), ptr(code, _f))
        assert bits_equal(code, coding[f"{name}_code"])
        de, dp = np.zeros(2 * m, np.float32), np.zeros(2 * m, np.float32)
        O.lib.orc_decode(m, ptr(code, _f), ptr(de, _f), ptr(dp, _f))
        assert bits_equal(de, coding[f"{name}_dec_exp"]) and bits_equal(dp, coding[f"{name}_dec_poly"])
        tq = np.zeros(m, np.float32)
        O.lib.orc_transform_quantization_real(m, ptr(code, _f), C.c_float(t[0]), ptr(tq, _f))
        assert bits_equal(tq, coding[f"{name}_tq"])
        dq = np.zeros(2 * m, np.float32)
        O.lib.orc_transform_dequantization_real(m, ptr(tq, _f), ptr(dq, _f), 1, None)
        assert bits_equal(dq, coding[f"{name}_dq_none"])
        # decode(encode(x)) == x within 1e-6 (tests/test_prediction_coding.cpp:63-71)
        assert np.abs(de - e).max() <= 1e-6


def test_compaction_counts_match_reference(O, coding):
    for i, th in enumerate(coding["compaction_thr"]):
        got = [O.lib.orc_determine_best_number_moments(24, ptr(np.ascontiguousarray(mo), _f), C.c_float(th))
               for mo in coding["compaction_moments"]]
        assert np.array_equal(np.int32(got), coding["compaction_counts"][i])
    got10 = [O.lib.orc_determine_best_number_moments(10, ptr(coding["comp10_in"], _f), C.c_float(th))
             for th in coding["compaction_thr"]]
    assert np.array_equal(np.int32(got10), coding["comp10_counts"])
    ne = np.zeros(20, np.float32)
    O.lib.orc_evaluate_unknown_moments(4, ptr(coding["comp10_exp"], _f), 10, ptr(ne, _f))
    assert bits_equal(ne, coding["comp10_unknown_n4"])
    bm = np.zeros(20, np.float32)
    bm[0] = coding["comp10_in"][0]
    O.lib.orc_exp_to_bounded_moments(10, ptr(ne, _f), ptr(bm, _f), 1)
    assert bits_equal(bm, coding["comp10_bounded_n1"])


def test_compaction_offsets_like_reference_test(O):
    # tests/test_moment_image.cpp:31-63: counts x -> offsets 0,1,3,6 in row 0
    W, H, m = 8, 4, 10
    counts = np.tile(np.arange(8, dtype=np.uint32), H)
    dense = np.zeros((H, W, m), np.float32)
    for x in range(8):
        dense[:, x, :x] = x / 8.0
    index, comp = O.compact(W, H, m, dense, counts)
    assert list(index[1:5]) == [0, 1, 3, 6]
    assert len(comp) == counts.sum()


def test_coding_params_like_reference_test(O):
    # tests/test_coding_transform.cpp:32-61: values x/15 - 0.6 give params (0.6, 0.5)
    W, H, m = 16, 1, 10
    counts = np.full(W * H, 10, np.uint32)
    dense = np.zeros((H, W, m), np.float32)
    for x in range(W):
        dense[0, x, :] = np.float32(x) / np.float32(W - 1) - np.float32(0.6)
    index, comp = O.compact(W, H, m, dense, counts)
    params = np.zeros(2 * (m - 1), np.float32)
    O.lib.orc_find_coding_params(W, H, m, ptr(comp, _f), ptr(index, _u32), ptr(params, _f))
    assert np.allclose(params[0::2], 0.6, atol=1e-6) and np.allclose(params[1::2], 0.5, atol=1e-6)


# the reference direct volume renderer's cases in golden_reference.npz
# (tests/golden/make_golden.py, reference_dvr_vectors)
DVR_CASES = {
    "dvr_ml_ortho": ((-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 20, 16,
                     dict(step_size=2.0 / 48, medium_emission=1.0, early_ray_termination=0.99), 4 * 2.0 / 48),
    "dvr_u8_persp": ((0, 0, 0), (1, 1, 1), synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=16 / 14), 16,
                     14, dict(step_size=1.0 / 40, medium_emission=0.5, early_ray_termination=0.9, data_min=0.1,
                              data_max=0.9), 20.0 / 40),
    "dvr_ml_opaque": ((-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 12, 10,
                      dict(step_size=2.0 / 48, medium_emission=1.0, early_ray_termination=0.9), 60 * 2.0 / 48),
    "dvr_ml_miss": ((-1, -1, -1), (1, 1, 1),
                    synth.perspective_camera(position=(1.5, 0.2, -2.5), fov_half_deg=25.0, aspect=18 / 12), 18, 12,
                    dict(step_size=2.0 / 40, medium_emission=1.0, early_ray_termination=0.95), 4 * 2.0 / 40),
}


@pytest.mark.parametrize("case", list(DVR_CASES))
def test_reference_dvr_matches_reference(O, case):
    """The oracle's reference direct volume renderer reproduces the compiled
    reference's reference_ray bit for bit (float RGBA) and its
    generate_reference RGBA8 image (f32 and u8 volumes, rays that miss, early
    termination)."""
    g = dict(np.load(GOLD / "golden_reference.npz"))
    dmin, dmax, cam, W, H, pk, tf_step = DVR_CASES[case]
    table = O.tf(g["tf_rgba8"], tf_step)
    rgba, ns = O.reference(g[f"{case}_volume"], dmin, dmax, cam, W, H, table, RefParams(**pk))
    assert bits_equal(rgba, g[f"{case}_rgba"])
    q = (np.clip(rgba, 0, 1) * np.float32(255)).astype(np.uint8)
    assert np.array_equal(q, g[f"{case}_rgba8"])


QUANT_CASES = ["pc_default10", "pc_table", "pc_ragged12", "plain8", "plain13", "plain32"]


@pytest.mark.parametrize("case", QUANT_CASES)
def test_quantization_matches_reference(O, case):
    """Packed streams and dequantized values of moment_quantization
    (src/moment_quantization.cpp:244-360) bit for bit."""
    g = dict(np.load(GOLD / "golden_quant.npz"))
    index, data, pc, bits = g[f"{case}_index"], g[f"{case}_data"], int(g[f"{case}_pc"]), g[f"{case}_bits"]
    W, H = 19, 11
    stream = O.quantize(W, H, data, index, pc, bits)
    assert np.array_equal(stream, g[f"{case}_stream"])
    if f"{case}_decoded" in g:
        dec = O.dequantize(W, H, index, pc, bits, stream)
        assert bits_equal(dec, g[f"{case}_decoded"])


def test_quantization_packing_and_steps(O):
    """Independent of the reference: the stream is the concatenation of the
    words least significant bit first (numpy's little-endian bit order), and
    each value comes back within one quantization step of its clamp."""
    from oracle_ffi import quant_bits

    rng = np.random.default_rng(7)
    W, H = 9, 5
    counts = rng.integers(0, 12, W * H).astype(np.uint32)
    index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    n = int(index[-1])
    for pc, b in ((1, 10), (0, 7), (0, 32)):
        bits = quant_bits(256, pc, b)
        lo = np.where((np.arange(n) - np.repeat(index[:-1], counts) == 0) | (pc == 1), 0.0, -1 / math.pi)
        lo = lo.astype(np.float32)
        hi = np.where(lo == 0, 1.0, 1 / math.pi).astype(np.float32)
        data = (lo + (hi - lo) * rng.uniform(-0.1, 1.1, n)).astype(np.float32)
        stream = O.quantize(W, H, data, index, pc, bits)
        ls = np.concatenate([np.arange(c) for c in counts]).astype(np.int64)
        widths = bits[ls].astype(np.int64)
        dec = O.dequantize(W, H, index, pc, bits, stream)
        # re-pack the words the decoder returns, little-endian bit order
        words = []
        for v, w, l0, h0 in zip(dec, widths, lo, hi):
            if w == 32:
                words.append(int(np.float32(v).view(np.uint32)))
            else:
                words.append(int(round((float(v) - float(l0)) / (float(h0) - float(l0)) * ((1 << w) - 1))))
        bitlist = np.concatenate([[(q >> i) & 1 for i in range(w)] for q, w in zip(words, widths)]).astype(np.uint8)
        nb = len(bitlist) // 8
        assert np.array_equal(np.packbits(bitlist[:nb * 8], bitorder="little"), stream)
        clamped = np.clip(data, lo, hi)
        step = np.where(widths == 32, 0.0, (hi - lo) / ((2.0 ** np.minimum(widths, 31)) - 1))
        whole = np.arange(n) < n - (1 if len(bitlist) % 8 else 0)
        err = np.abs(dec - clamped)[whole]
        assert np.all(err <= step[whole] * (1 + 1e-3) + 1e-7)


CURVE_CASES = ["ml", "mix", "mix64"]


@pytest.mark.parametrize("case", CURVE_CASES)
def test_quantization_curve_matches_reference(O, case):
    """The quantization-curve search (create_quantization_table,
    src/measure_coding_error.cpp:374-437) picks the compiled reference's
    tables on coded images it generated."""
    g = dict(np.load(GOLD / "golden_curve.npz"))
    W, H, m = (int(v) for v in g[f"{case}_shape"])
    data, index = g[f"{case}_data"], g[f"{case}_index"]
    pf = (g[f"{case}_params"].astype(np.float32) / np.float32(65535.0)).astype(np.float32)
    for sb in (8, 10, 14):
        t = np.zeros(m, np.uint8)
        O.lib.orc_quantization_table(W, H, m, ptr(data, C.POINTER(C.c_float)), ptr(index, C.POINTER(C.c_uint32)), 0,
                                     ptr(pf, C.POINTER(C.c_float)), sb, ptr(t, C.POINTER(C.c_uint8)))
        assert np.array_equal(t, g[f"{case}_table{sb}"]), (sb, t.tolist(), g[f"{case}_table{sb}"].tolist())
        assert t[0] == 16 and t[1] == sb and np.all(np.diff(t[1:].astype(int)) <= 0) and t[-1] >= 2


def test_sph_gather_matches_reference():
    g = dict(np.load(GOLD / "golden_sph.npz"))
    O = Oracle()
    grid = O.particles_to_grid(g["sph_x"], g["sph_y"], g["sph_z"], g["sph_v"], float(g["sph_h"][0]),
                               (0, 0, 0), (1, 1, 0.5), 16, 16, 8)
    # the reference sorts particles with std::sort (unstable within a cell): summation order differs
    assert np.allclose(grid, g["sph_grid"], rtol=2e-6, atol=1e-7)
    assert np.count_nonzero(grid) > 0.05 * grid.size


@pytest.mark.skipif(not REF_SO.exists(), reason="compiled reference not built (build container only)")
def test_oracle_matches_live_reference_on_fresh_inputs():
    O, R = Oracle(), Reference()
    vol = synth.marschner_lobb(40)
    cam = synth.perspective_camera(position=(0.2, 0.1, -2.5), fov_half_deg=22.0, aspect=30 / 26)
    gp = GenParams(step_size=2.0 / 70, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                   prediction_coding=0, bias=5e-5)
    mo, b, _ = O.generate(vol, (-1, -1, -1), (1, 1, 1), cam, 30, 26, 20, gp)
    rd, rb, _, _ = R.generate(vol, (-1, -1, -1), (1, 1, 1), cam, 30, 26, 20, gp)
    assert bits_equal(mo.reshape(-1), rd) and np.array_equal(b, rb)


def test_particle_march_matches_reference():
    """generate on a particle set, Particles::sample at every ray sample
    (src/volume.cpp:83-94 inside generate_ray): the oracle's march over its
    particle set against the compiled reference's ::generate on a Particles
    (golden_sph.npz). The reference's std::sort leaves the order inside a cell
    unspecified, so sample sums may differ in the last bits: moments to 1e-5
    normwise, 16-bit bounds within one step."""
    g = dict(np.load(GOLD / "golden_sph.npz"))
    O = Oracle()
    from oracle_ffi import Particles

    parts = Particles(O, g["sph_x"], g["sph_y"], g["sph_z"], g["sph_v"], float(g["sph_h"][0]))
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=24.0, aspect=1.0)
    gp = GenParams(step_size=0.02, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    mo, b, ns = O.generate(parts, (0, 0, 0), (1, 1, 0.5), cam, 14, 14, 12, gp)
    ref = g["sph_march_moments"].astype(np.float64).reshape(-1, 12)
    err = np.linalg.norm(mo.reshape(-1, 12) - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), 1e-3)
    assert err.max() <= 1e-5
    lo, hi = b & 0xFFFF, b >> 16
    rlo, rhi = g["sph_march_bounds"] & 0xFFFF, g["sph_march_bounds"] >> 16
    assert np.abs(lo.astype(int) - rlo.astype(int)).max() <= 1 and np.abs(hi.astype(int) - rhi.astype(int)).max() <= 1
    assert np.count_nonzero(ns) > 0.5 * ns.size
