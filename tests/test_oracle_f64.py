"""Pins every FP64 function of the oracle (the ground truth of the MIXED
kernels) by checks that do not restate its formulas:

* against the compiled reference's FP32 outputs (golden_mese.npz,
  golden_images.npz) within FP32 rounding, on the reference's own MESE
  vectors (tests/test_mese.cpp) and image cases;
* against textbook routines: a dense Hermitian-Toeplitz solve for the
  Levinson recursion (src/mese/MESE_dynamic.h:100-135), the FFT of
  exp(2 pi i sum t_l z^l) for the exponential moments (:35-54), numpy's
  correlate for the correlations (:137-171), numpy's Chebyshev series for the
  Fourier evaluation (:213-226);
* against the defining property of the bounded MESE: the density
  atan(S(phi)) / pi + 1/2 built from the prepared multipliers has exactly the
  trigonometric moments it was prepared from (Peters et al. 2019, the
  moment-consistency of the bounded maximum-entropy spectral estimate), which
  a dropped term, a wrong sign, index or normalisation breaks;
* against closed forms of the march (constant density compositing).
"""
import ctypes as C
import math
import pathlib

import numpy as np
import pytest

from mray_amd import synth
from mray_amd.types import GenParams, RecParams
from oracle_ffi import Oracle, ptr

GOLD = pathlib.Path(__file__).resolve().parent / "golden"
_d = C.POINTER(C.c_double)
MESE_CASES = ["trig4", "trig10", "trig28", "comp10"]


@pytest.fixture(scope="module")
def O():
    return Oracle()


@pytest.fixture(scope="module")
def mese():
    return dict(np.load(GOLD / "golden_mese.npz"))


def cplx(a):
    return a[0::2] + 1j * a[1::2]


def exp_f64(O, t):
    m = len(t)
    e = np.zeros(2 * m)
    O.lib.orc_trig_to_exp_f64(m, ptr(np.ascontiguousarray(t, np.float64), _d), ptr(e, _d))
    return e


def levinson_f64(O, e):
    m = len(e) // 2
    out = np.zeros(2 * m)
    O.lib.orc_levinson_f64(m, ptr(np.ascontiguousarray(e, np.float64), _d), ptr(out, _d))
    return out


def prepare_f64(O, t):
    m = len(t)
    out = np.zeros(m)
    O.lib.orc_prepare_mese_f64(m, ptr(np.ascontiguousarray(t, np.float64), _d), ptr(out, _d))
    return out


def trig_moments(g, m, n=200001):
    """c_j = (1/pi) int_{-pi}^0 g(phi) cos(j phi) dphi: the moments generate
    writes (src/raymarch_generate.h:35-58 integrates f e^{-ij phi}; the real
    part over [-pi, 0], divided by pi). The trapezoid rule is spectrally
    accurate here because the even extension of g cos(j .) is periodic."""
    phi = np.linspace(-np.pi, 0.0, n)
    gv = g(phi)
    return np.array([np.trapezoid(gv * np.cos(j * phi), phi) / np.pi for j in range(m)]), phi


def density_exact_atan(lam, phi):
    S = np.polynomial.chebyshev.chebval(np.cos(phi), np.concatenate(([lam[0]], 2.0 * lam[1:])))
    return np.arctan(S) / np.pi + 0.5


SMOOTH = lambda phi: 0.5 + 0.3 * np.sin(3 * phi + 0.4) + 0.12 * np.cos(7 * phi) + 0.05 * np.sin(11 * phi)
SHARP = lambda phi: 0.02 + 0.96 * (0.5 + 0.5 * np.tanh((phi + 1.3) / 0.05))


# ------------------------------------------------ against the compiled reference
@pytest.mark.parametrize("name", MESE_CASES)
def test_f64_blocks_agree_with_reference_fp32(O, mese, name):
    """FP64 exponential moments, Levinson, multipliers and density against the
    compiled reference's FP32 values on its unit-test vectors: they differ by
    FP32 rounding only (measured 2e-8 .. 1e-6; a mistake shows at 1e-3)."""
    t = mese[f"{name}_in"].astype(np.float64)
    m = len(t)
    e = exp_f64(O, t)
    g32 = mese[f"{name}_exp"].astype(np.float64)
    assert np.abs(e - g32).max() <= 2e-7 * np.abs(g32).max()
    lev = levinson_f64(O, e)
    l32 = mese[f"{name}_levinson"].astype(np.float64)
    assert np.abs(lev - l32).max() <= 3e-6 * np.abs(l32).max()
    lam = prepare_f64(O, t)
    p32 = mese[f"{name}_prepare"].astype(np.float64)
    assert np.abs(lam - p32).max() <= 2e-6 * np.abs(p32).max()
    ev = np.array([O.lib.orc_evaluate_mese_f64(m, float(p), ptr(lam, _d)) for p in mese[f"{name}_phases"]])
    assert np.abs(ev - mese[f"{name}_evaluate"]).max() <= 1e-6
    for bias in (1e-4, 5e-5):
        got = O.prepare_std_f64(m, t, bias)
        ref = mese[f"{name}_prepare_bias{bias:g}"].astype(np.float64)
        assert np.abs(got - ref).max() <= 2e-6 * np.abs(ref).max()


def test_f64_fast_atan_matches_reference_polynomial(O, mese):
    """evaluate_mese_f64 with one multiplier is fast_atan(lambda_0) / pi + 1/2:
    the reference's polynomial (golden FP32 outputs of fast_atan) in double."""
    for x, y in zip(mese["atan_in"], mese["atan_out"]):
        lam = np.array([float(x)])
        v = O.lib.orc_evaluate_mese_f64(1, -1.0, ptr(lam, _d))
        assert abs((v - 0.5) * math.pi - float(y)) <= 4 * abs(np.spacing(np.float32(y))) + 1e-7


# --------------------------------------------------------- textbook routines
@pytest.mark.parametrize("m", [12, 30, 64, 100])
def test_levinson_f64_solves_toeplitz_system(O, m):
    """Levinson-Durbin against numpy's dense solve of T a = e_0, T the
    Hermitian Toeplitz matrix of the exponential moments (diagonal Re g_0)."""
    c, _ = trig_moments(SMOOTH, m)
    e = exp_f64(O, c)
    a = cplx(levinson_f64(O, e))
    g = cplx(e).copy()
    g[0] = g[0].real
    i, k = np.indices((m, m))
    T = np.where(i >= k, g[np.abs(i - k)], np.conj(g[np.abs(i - k)]))
    ref = np.linalg.solve(T, np.eye(m)[0].astype(complex))
    assert np.abs(a - ref).max() <= 1e-12 * np.abs(ref).max()


def test_levinson_f64_guard_zeroes_later_coefficients(O):
    """A first column whose reflection coefficient reaches |u| >= 1 (not
    positive definite) leaves the solution of the leading block: every later
    coefficient is forced to 0 (src/mese/MESE_dynamic.h:106-122)."""
    g = np.array([1.0, 0.0, 1.5, 0.3j, 0.1], complex)  # u_2 = -1.5 / 1 -> |u| >= 1
    e = np.zeros(10)
    e[0::2], e[1::2] = g.real, g.imag
    a = cplx(levinson_f64(O, e))
    # the 2x2 leading system gives a = (1, 0); the guard keeps it
    assert np.allclose(a, [1.0, 0, 0, 0, 0], atol=1e-15)


@pytest.mark.parametrize("m", [10, 50, 128])
def test_trig_to_exp_f64_is_series_exponential(O, m):
    """gamma(z) = gamma_0 exp(2 pi i sum_{l>=1} t_l z^l) with gamma_0 =
    e^{i(pi t_0 - pi/2)} / (4 pi): its Taylor coefficients from the FFT of the
    function on 4096 roots of unity (gamma_0 doubled, as the reference stores
    it)."""
    c, _ = trig_moments(SMOOTH, m)
    e = cplx(exp_f64(O, c))
    N = 4096
    z = np.exp(2j * np.pi * np.arange(N) / N)
    poly = np.polynomial.polynomial.polyval(z, np.concatenate(([0.0], c[1:])))
    G = np.exp(1j * (math.pi * c[0] - math.pi / 2)) / (4 * math.pi) * np.exp(2j * math.pi * poly)
    coef = np.fft.fft(G) / N
    expect = coef[:m].copy()
    expect[0] *= 2
    assert np.abs(e - expect).max() <= 1e-12 * np.abs(expect).max()


def test_correlations_f64_by_definition(O):
    """autocorrelation R_k = sum_j s_j conj(s_{j+k}) and imaginary correlation
    Im(sum_j l_{k+j} r_j) against numpy.correlate."""
    rng = np.random.default_rng(7)
    m = 17
    s = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    r = rng.standard_normal(m) + 1j * rng.standard_normal(m)
    il = lambda v: np.ascontiguousarray(np.column_stack((v.real, v.imag)).reshape(-1))
    R = np.zeros(2 * m)
    O.lib.orc_autocorrelation_f64(m, ptr(il(s), _d), ptr(R, _d))
    full = np.correlate(s, s, "full")  # full[m-1+k] = sum_n s[n+k] conj(s[n])
    assert np.allclose(cplx(R), np.conj(full[m - 1:]), rtol=0, atol=1e-12)
    ic = np.zeros(m)
    O.lib.orc_imag_correlation_f64(m, ptr(il(s), _d), ptr(il(r), _d), ptr(ic, _d))
    full2 = np.correlate(s, np.conj(r), "full")  # sum_n s[n+k] r[n]
    assert np.allclose(ic, full2[m - 1:].imag, rtol=0, atol=1e-12)


def test_evaluate_f64_series_is_chebyshev_sum(O):
    """The Fourier series lambda_0 + 2 sum lambda_l cos(l phi) against numpy's
    Chebyshev evaluation at cos(phi) (the fast_atan step is pinned above)."""
    rng = np.random.default_rng(3)
    m = 40
    lam = rng.standard_normal(m) * 0.01
    for phi in np.linspace(-math.pi, 0.0, 23):
        S = np.polynomial.chebyshev.chebval(math.cos(phi), np.concatenate(([lam[0]], 2 * lam[1:])))
        # S in (-1, 1) here, where fast_atan is its polynomial branch
        v = O.lib.orc_evaluate_mese_f64(m, float(phi), ptr(lam, _d))
        x = S
        P = (((0.020835 * x * x - 0.085133) * x * x + 0.180141) * x * x - 0.330299) * x * x + 0.999866
        assert abs(S) < 1.0 and abs((v - 0.5) * math.pi - x * P) <= 1e-6


# ---------------------------------------------- the bounded MESE's definition
@pytest.mark.parametrize("shape", ["smooth", "sharp"])
@pytest.mark.parametrize("m", [8, 30, 50, 100, 200])
@pytest.mark.parametrize("bias", [0.0, 1e-4])
def test_prepare_f64_reproduces_moments(O, shape, m, bias):
    """The multipliers of the bounded MESE define the density
    atan(lambda_0 + 2 sum lambda_l cos(l phi)) / pi + 1/2, whose trigonometric
    moments equal the (biased) moments it was prepared from. Measured to
    5e-16; a wrong sign, index or normalisation anywhere in the Herglotz
    transform, Levinson recursion or correlations misses by >= 1e-3."""
    g = SMOOTH if shape == "smooth" else SHARP
    c, phi = trig_moments(g, m)
    lam = O.prepare_std_f64(m, c, bias)
    cb = c * (1 - bias)
    cb[0] += 0.5 * bias
    rho = density_exact_atan(lam, phi)
    got = np.array([np.trapezoid(rho * np.cos(j * phi), phi) / np.pi for j in range(m)])
    assert np.abs(got - cb).max() <= 1e-12


def test_evaluate_f64_density_reproduces_moments(O):
    """Through orc_evaluate_mese_f64 itself (the reference's fast_atan, error
    <= 1.1e-5 in the angle) the moments are reproduced to 1.1e-5 / pi."""
    m = 30
    c, phi = trig_moments(SHARP, m, n=20001)
    lam = O.prepare_std_f64(m, c, 0.0)
    rho = np.array([O.lib.orc_evaluate_mese_f64(m, float(p), ptr(lam, _d)) for p in phi])
    got = np.array([np.trapezoid(rho * np.cos(j * phi), phi) / np.pi for j in range(m)])
    assert np.abs(got - c).max() <= 1.1e-5 / math.pi


# ------------------------------------------------------------ the FP64 march
def test_compositing_f64_closed_form(O):
    """Zero multipliers give density 1/2 at every sample; with a constant
    table entry (c, sigma) the march composites L = e c sum_{k<n} q^k,
    alpha = 1 - q^n, q = exp(-sigma): the geometric sums, in FP64."""
    m, W, H = 6, 2, 2
    tf8 = np.zeros((256, 4), np.uint8)
    tf8[:, :3] = 200
    tf8[:, 3] = 40
    step = 1.0 / 64
    table = O.tf(tf8, step)
    rp = RecParams(step_size=step, early_ray_termination=2.0, medium_emission=1.0)
    cam = synth.ortho_camera(pos=(0.5, 0.5, -1.0), right=(0.3, 0, 0), up=(0, 0.3, 0))
    lag = np.zeros(W * H * m)
    for rgba, ns in (O.reconstruct_lagrange_f64(cam, (0, 0, 0), (1, 1, 1), W, H, m, lag.astype(np.float32), table,
                                                rp),
                     (O.reconstruct_lagrange64_f64(cam, (0, 0, 0), (1, 1, 1), W, H, m, lag, table.astype(np.float64),
                                                   rp), None),
                     O.reconstruct_f64(cam, (0, 0, 0), (1, 1, 1), W, H, m,
                                       np.tile(np.float64([0.5] + [0.0] * (m - 1)), W * H), table, rp)):
        _, ns32 = O.reconstruct(cam, (0, 0, 0), (1, 1, 1), W, H, m, lag.astype(np.float32), table, rp)
        n = int(ns32[0])
        c = table.reshape(256, 4)[128].astype(np.float64)  # density 1/2 -> entry 128, weight of 129 is 0
        q = math.exp(-c[3])
        L = c[0] * (1 - q ** n) / (1 - q)
        assert abs(rgba[0, 0, 0] - L) <= 1e-12 * max(1.0, L) and abs(rgba[0, 0, 3] - (1 - q ** n)) <= 1e-12


IMAGE_CASES = {
    # the golden image cases of test_oracle_golden.py that are well conditioned
    "ml_ortho": ((-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 24, 20, 12,
                 dict(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                      prediction_coding=0, bias=5e-5),
                 dict(step_size=2.0 / 64, bias=5e-5, early_ray_termination=0.99, medium_emission=1.0)),
    "ml_persp_plain": ((-1, -1, -1), (1, 1, 1),
                       synth.perspective_camera(position=(0.1, -0.05, -3.0), fov_half_deg=20.0, znear=0.01,
                                                aspect=20 / 16), 20, 16, 10,
                       dict(step_size=2.0 / 48, compact_image=0, prediction_coding=0),
                       dict(step_size=2.0 / 48, early_ray_termination=0.95)),
}


@pytest.mark.parametrize("case", list(IMAGE_CASES))
def test_f64_images_agree_with_reference_fp32(O, case):
    """FP64 generate + prepare + march (orc_generate_image_f64,
    orc_reconstruct_image_f64) and the FP64 march on the reference's FP32
    multipliers (orc_reconstruct_lagrange_f64) against the compiled
    reference's float RGBA. Moments agree to 1.1e-6 (FP32 rounding); the
    images to the FP32 march's own evaluation error (measured RMS 5.8e-6 and
    2.0e-5, march alone 5.7e-6 and 2.8e-8); a wrong sample, table weight or
    compositing step misses by > 1e-3."""
    images = dict(np.load(GOLD / "golden_images.npz"))
    dmin, dmax, cam, W, H, m, gk, rk = IMAGE_CASES[case]
    gp, rp = GenParams(**gk), RecParams(**rk)
    vol = images[f"{case}_volume"]
    mo64, b = O.generate(vol, dmin, dmax, cam, W, H, m, gp, f64=True)
    mo32 = images[f"{case}_data"].reshape(H, W, m).astype(np.float64)
    rel = np.linalg.norm(mo64 - mo32, axis=2) / np.maximum(np.linalg.norm(mo32, axis=2), 1e-30)
    assert rel.max() <= 3e-6
    if gp.transform_to_bounds:
        assert np.array_equal(b, images[f"{case}_bounds"])
    bounds = b if gp.transform_to_bounds else None
    table = O.tf(images["tf_rgba8"], rp.step_size)
    gold = images[f"{case}_rgba"].astype(np.float64)
    rgba64, ns64 = O.reconstruct_f64(cam, dmin, dmax, W, H, m, mo64, table, rp, bounds=bounds)
    _, ns32 = O.reconstruct(cam, dmin, dmax, W, H, m, images[f"{case}_data"], table, rp, bounds=bounds)
    assert np.array_equal(ns64, ns32)
    d = rgba64 - gold
    assert np.sqrt((d ** 2).mean()) <= 5e-5 and np.abs(d).max() <= 2e-3
    lam = np.stack([O.prepare_std(m, v, rp.bias) for v in images[f"{case}_data"].reshape(-1, m)]).reshape(-1)
    r64, _ = O.reconstruct_lagrange_f64(cam, dmin, dmax, W, H, m, lam, table, rp, bounds=bounds)
    d = r64 - gold
    tol = 2e-5 if case == "ml_ortho" else 1e-6
    assert np.sqrt((d ** 2).mean()) <= tol
