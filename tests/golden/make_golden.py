"""Writes the golden vectors under tests/golden/ from the compiled reference.

Run in the build container (needs oracle/_ref/libmrayref.so, built by
`make -C oracle ref` from the reference sources):

    python tests/golden/make_golden.py            # every fixture
    python tests/golden/make_golden.py reference  # golden_reference.npz only

The moment-image files are written in a child process of this script: the
reference keeps its quantization table as process state.

Inputs are (a) the vectors of the reference's own unit tests
(tests/test_mese.cpp, tests/test_prediction_coding.cpp,
tests/test_moment_compaction.cpp) and (b) small synthetic workloads from
mray_amd.synth. Outputs are whatever the reference functions return. Only the
resulting arrays are committed (golden_*.npz); no reference code is stored.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import sys

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

from mray_amd import synth  # noqa: E402
from mray_amd.types import RefParams, GenParams, RecParams  # noqa: E402
from oracle_ffi import Reference, ptr  # noqa: E402

OUT = pathlib.Path(__file__).resolve().parent
_f = C.POINTER(C.c_float)

# vectors of the reference unit tests (tests/test_mese.cpp:33,47,61,97-103,155-156,187,205-206;
# tests/test_prediction_coding.cpp:32,74; tests/test_moment_compaction.cpp:32-33,85-104)
TRIG4 = [0.22008814, 0.0, 0.01400074, 0.00271224]
MOM4 = [0.5, 0.25, 0.25, 0.1]
TRIG10 = [0.22008814, 0.0, 0.01400074, 0.00271224, 0.00426201, 0.0, 0.0, 0.0, 0.0, 0.0]
TRIG28 = [5.2070445e-01, 7.4165657e-02, -1.3283056e-02, 1.2487045e-02, 1.3437086e-02, -9.5647079e-04,
          -1.2216626e-03, 2.1146263e-03, -5.4606516e-04, 1.0199437e-02, 1.1894707e-02, -1.9188672e-02,
          -5.5241492e-03, 2.8166445e-02, 6.7452774e-03, -1.9198583e-02, 4.3223342e-03, -6.1640824e-04,
          7.8988671e-03, 2.4080461e-02, -9.6553434e-03, -1.9514188e-02, 1.1532067e-02, -4.1229022e-03,
          2.3510400e-02, 1.7623754e-02, -1.4590423e-02, 1.1135901e-03]
TRIG_C4 = [(0.22008814, 0.0), (0.0, 0.0), (0.01400074, 0.5), (0.00271224, 0.1)]
SIG4 = [(2.0, 0.0), (4.0, 0.0), (8.0, 0.0), (16.0, 0.0)]
RHS4 = [(1.0, 0.0), (42.0, 0.0), (0.1, 0.0), (0.2, 0.0)]
PREP_C4 = [(0.22008814, 0.05), (-0.02, 0.01), (0.01400074, 0.04), (0.00271224, 0.1)]
PC5 = [0.22, 0.0, 0.014, 0.0027, 0.004]
PC5B = [0.22008814, 0.0, 0.01400074, 0.00271224, 0.00426201]
COMP10 = [1.9176584e-01, -3.7339486e-02, -1.4143054e-02, -1.6614912e-02, -2.2262430e-02, -1.2005451e-02,
          -7.5180288e-03, 2.5505440e-02, 2.8378680e-02, 1.6212923e-02]


def f32(v):
    return np.ascontiguousarray(np.asarray(v, dtype=np.float32))


def c64(pairs):
    return np.ascontiguousarray(np.asarray(pairs, dtype=np.float32).reshape(-1))


def mese_vectors(R: Reference) -> dict:
    L = R.lib
    out = {}
    for name, t in (("trig4", TRIG4), ("trig10", TRIG10), ("trig28", TRIG28), ("comp10", COMP10)):
        t = f32(t)
        m = len(t)
        e = np.zeros(2 * m, np.float32)
        L.ref_trig_to_exp(m, ptr(t, _f), ptr(e, _f))
        out[f"{name}_in"] = t
        out[f"{name}_exp"] = e
        back = np.zeros(2 * m, np.float32)
        L.ref_exp_to_trig(m, ptr(e, _f), ptr(back, _f))
        out[f"{name}_exp_to_trig"] = back
        lev = np.zeros(2 * m, np.float32)
        L.ref_levinson(m, ptr(e, _f), ptr(lev, _f))
        out[f"{name}_levinson"] = lev
        pm = np.zeros(m, np.float32)
        L.ref_prepare_mese(m, ptr(t, _f), ptr(pm, _f))
        out[f"{name}_prepare"] = pm
        phases = f32(np.linspace(-np.pi, 0.0, 33))
        out[f"{name}_phases"] = phases
        out[f"{name}_evaluate"] = f32([L.ref_evaluate_mese(m, float(p), ptr(pm, _f)) for p in phases])
        for bias in (1e-4, 5e-5):
            pb = np.zeros(m, np.float32)
            L.ref_prepare_dense(m, ptr(t, _f), C.c_float(bias), ptr(pb, _f))
            out[f"{name}_prepare_bias{bias:g}"] = pb
    # fixed-size versions the reference tests compare against (src/mese/MESE_4_10.h)
    m4 = f32(MOM4)
    pf = np.zeros(4, np.float32)
    L.ref_fixed_prepare4(ptr(m4, _f), ptr(pf, _f))
    out["mom4_in"] = m4
    out["mom4_fixed_prepare"] = pf
    p = np.float32(-np.pi + np.pi * 0.5)
    out["mom4_fixed_eval"] = f32([L.ref_fixed_evaluate4(float(p), ptr(pf, _f))])
    t10 = f32(TRIG10)
    pf10 = np.zeros(10, np.float32)
    L.ref_fixed_prepare10(ptr(t10, _f), ptr(pf10, _f))
    out["trig10_fixed_prepare"] = pf10
    out["trig10_fixed_eval"] = f32([L.ref_fixed_evaluate10(float(p), ptr(pf10, _f))])
    e4 = np.zeros(8, np.float32)
    L.ref_fixed_trig_to_exp4(ptr(f32(TRIG4), _f), ptr(e4, _f))
    out["trig4_fixed_exp"] = e4
    # complex-input overloads
    tc = c64(TRIG_C4)
    ec = np.zeros(8, np.float32)
    L.ref_trig_to_exp_c(4, ptr(tc, _f), ptr(ec, _f))
    out["trigc4_in"] = tc
    out["trigc4_exp"] = ec
    pc = c64(PREP_C4)
    bias = np.float32(1e-2)
    pcb = pc.reshape(4, 2).copy()
    pcb[0] = (1 - bias) * pcb[0] + bias * np.float32(0.5) * np.array([1, 0], np.float32)
    pcb[1:] = (1 - bias) * pcb[1:]
    pcb = np.ascontiguousarray(pcb.reshape(-1))
    lc = np.zeros(8, np.float32)
    L.ref_prepare_mese_c(4, ptr(pcb, _f), ptr(lc, _f))
    out["prepc4_in"] = pcb
    out["prepc4_out"] = lc
    s, r = c64(SIG4), c64(RHS4)
    ac = np.zeros(8, np.float32)
    cc = np.zeros(8, np.float32)
    ic = np.zeros(4, np.float32)
    L.ref_autocorrelation(4, ptr(s, _f), ptr(ac, _f))
    L.ref_correlation(4, ptr(s, _f), ptr(r, _f), ptr(cc, _f))
    L.ref_imag_correlation(4, ptr(s, _f), ptr(r, _f), ptr(ic, _f))
    out.update(sig4=s, rhs4=r, autocorr4=ac, corr4=cc, imagcorr4=ic)
    xs = f32(np.concatenate([np.linspace(-50, 50, 401), [0.0, 1.0, -1.0, 1e-8, 1e6]]))
    out["atan_in"] = xs
    out["atan_out"] = f32([L.ref_fast_atan(float(x)) for x in xs])
    return out


def coding_vectors(R: Reference) -> dict:
    L = R.lib
    out = {}
    for name, t in (("pc5", PC5), ("pc5b", PC5B), ("comp10", COMP10), ("trig28", TRIG28)):
        t = f32(t)
        m = len(t)
        e = np.zeros(2 * m, np.float32)
        L.ref_trig_to_exp(m, ptr(t, _f), ptr(e, _f))
        code = np.zeros(2 * m, np.float32)
        L.ref_encode(m, ptr(e, _f), ptr(code, _f))
        dec_e = np.zeros(2 * m, np.float32)
        dec_p = np.zeros(2 * m, np.float32)
        L.ref_decode(m, ptr(code, _f), ptr(dec_e, _f), ptr(dec_p, _f))
        tq = np.zeros(m, np.float32)
        L.ref_transform_quant(m, ptr(code, _f), C.c_float(t[0]), ptr(tq, _f))
        dq = np.zeros(2 * m, np.float32)
        L.ref_transform_dequant(m, ptr(tq, _f), ptr(dq, _f), 1, None)
        out.update({f"{name}_in": t, f"{name}_exp": e, f"{name}_code": code, f"{name}_dec_exp": dec_e,
                    f"{name}_dec_poly": dec_p, f"{name}_tq": tq, f"{name}_dq_none": dq})
    # compaction (determine_best_number_moments) on the reference vectors and synthetic moments
    thr = f32([0.0, 1e-4, 1e-3, 1e-2, 0.05, 0.2])
    rng_moments = []
    vol = synth.marschner_lobb(32)
    gp = GenParams(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                   prediction_coding=0)
    data, _, _, _ = R.generate(vol, (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 8, 6, 24, gp)
    rng_moments = data.reshape(48, 24)
    counts = np.zeros((len(thr), 48), np.int32)
    for i, th in enumerate(thr):
        for p in range(48):
            counts[i, p] = L.ref_best_num_moments(24, ptr(np.ascontiguousarray(rng_moments[p]), _f), C.c_float(th))
    c10 = [L.ref_best_num_moments(10, ptr(f32(COMP10), _f), C.c_float(th)) for th in thr]
    out.update(compaction_thr=thr, compaction_moments=f32(rng_moments), compaction_counts=counts,
               comp10_counts=np.asarray(c10, np.int32))
    ne = np.zeros(20, np.float32)
    e10 = out["comp10_exp"]
    L.ref_eval_unknown(4, ptr(e10, _f), 10, ptr(ne, _f))
    out["comp10_unknown_n4"] = ne
    bm = np.zeros(20, np.float32)
    bm[0] = COMP10[0]
    L.ref_exp_to_bounded(10, ptr(ne, _f), ptr(bm, _f), 1)
    out["comp10_bounded_n1"] = bm
    return out


def image_vectors(R: Reference) -> dict:
    out = {}
    tf8 = synth.tf_ramps(0)
    out["tf_rgba8"] = tf8
    cases = {
        # name: (volume, dmin, dmax, camera, W, H, m, gen kwargs, rec kwargs)
        "ml_ortho": (synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 24, 20, 12,
                     dict(step_size=2.0 / 64, transform_to_bounds=1, transform_bound_eps=0.005, compact_image=0,
                          prediction_coding=0, bias=5e-5),
                     dict(step_size=2.0 / 64, bias=5e-5, early_ray_termination=0.99, medium_emission=1.0)),
        "ml_persp_plain": (synth.marschner_lobb(32), (-1, -1, -1), (1, 1, 1),
                           synth.perspective_camera(position=(0.1, -0.05, -3.0), fov_half_deg=20.0, znear=0.01,
                                                    aspect=20 / 16),
                           20, 16, 10, dict(step_size=2.0 / 48, compact_image=0, prediction_coding=0),
                           dict(step_size=2.0 / 48, early_ray_termination=0.95)),
        "mix_coded": (synth.mixing_layer(24), (0, 0, 0), (1, 1, 1),
                      synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=16 / 12), 16, 12, 16,
                      dict(step_size=1.0 / 48, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=1,
                           prediction_coding=1, error_threshold=0.01, bias=1e-4),
                      dict(step_size=1.0 / 48, bias=1e-4, early_ray_termination=0.98, medium_emission=1.0)),
    }
    u8 = np.clip(synth.mixing_layer(24) * 255.0, 0, 255).astype(np.uint8)
    cases["u8_ortho"] = (u8, (0, 0, 0), (1, 1, 1),
                         synth.ortho_camera(pos=(0.5, 0.5, -1.0), right=(0.5, 0, 0), up=(0, 0.5, 0)), 16, 16, 8,
                         dict(step_size=1.0 / 32, transform_to_bounds=1, compact_image=0, prediction_coding=0),
                         dict(step_size=1.0 / 32, early_ray_termination=0.9))
    for name, (vol, dmin, dmax, cam, W, H, m, gk, rk) in cases.items():
        gp = GenParams(**gk)
        rp = RecParams(**rk)
        data, bounds, index, cp = R.generate(vol, dmin, dmax, cam, W, H, m, gp)
        tf_step = rp.step_size
        rgba, rgba8 = R.reconstruct(cam, dmin, dmax, W, H, m, data, tf8, tf_step, rp,
                                    index=index if gp.compact_image else None,
                                    bounds=bounds if gp.transform_to_bounds else None,
                                    prediction_coded=int(gp.prediction_coding), coding_warp=gp.coding_warp,
                                    coding_params=cp if gp.prediction_coding else None, want_rgba8=True)
        out[f"{name}_volume"] = vol
        out[f"{name}_data"] = data
        out[f"{name}_bounds"] = bounds
        out[f"{name}_index"] = index
        out[f"{name}_coding_params"] = cp
        out[f"{name}_rgba"] = rgba
        out[f"{name}_rgba8"] = rgba8
    return out


def reference_dvr_vectors(R: Reference) -> dict:
    """The reference's direct volume renderer (reference_ray, generate_reference):
    float RGBA per pixel and the RGBA8 image, on small f32 and u8 volumes."""
    out = {}
    tf8 = synth.tf_ramps(0)
    out["tf_rgba8"] = tf8
    u8 = np.clip(synth.mixing_layer(20) * 255.0, 0, 255).astype(np.uint8)
    cases = {
        # name: (volume, dmin, dmax, camera, W, H, reference parameters, tf step)
        "dvr_ml_ortho": (synth.marschner_lobb(24), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 20, 16,
                         dict(step_size=2.0 / 48, medium_emission=1.0, early_ray_termination=0.99), 4 * 2.0 / 48),
        "dvr_u8_persp": (u8, (0, 0, 0), (1, 1, 1),
                         synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=16 / 14), 16, 14,
                         dict(step_size=1.0 / 40, medium_emission=0.5, early_ray_termination=0.9, data_min=0.1,
                              data_max=0.9), 20.0 / 40),
        # opaque enough for early ray termination (alpha set to 1)
        "dvr_ml_opaque": (synth.marschner_lobb(24), (-1, -1, -1), (1, 1, 1), synth.ortho_camera(), 12, 10,
                          dict(step_size=2.0 / 48, medium_emission=1.0, early_ray_termination=0.9), 60 * 2.0 / 48),
        "dvr_ml_miss": (synth.marschner_lobb(24), (-1, -1, -1), (1, 1, 1),
                        synth.perspective_camera(position=(1.5, 0.2, -2.5), fov_half_deg=25.0, aspect=18 / 12),
                        18, 12, dict(step_size=2.0 / 40, medium_emission=1.0, early_ray_termination=0.95),
                        4 * 2.0 / 40),
    }
    for name, (vol, dmin, dmax, cam, W, H, pk, tf_step) in cases.items():
        p = RefParams(**pk)
        rgba, rgba8 = R.reference(vol, dmin, dmax, cam, W, H, tf8, tf_step, p, want_rgba8=True)
        out[f"{name}_volume"] = vol
        out[f"{name}_rgba"] = rgba
        out[f"{name}_rgba8"] = rgba8
    return out


def sph_vectors(R: Reference) -> dict:
    x, y, z, v = synth.particle_mixture(3000, seed=5, n_gauss=6)
    v = (0.5 + 0.5 * synth.hash_uniform(7, 1, np.arange(3000))).astype(np.float32)
    h = np.float32(0.05)
    dmin, dmax = (0.0, 0.0, 0.0), (1.0, 1.0, 0.5)
    n = (16, 16, 8)
    pts = []
    for iz in range(n[2]):
        for iy in range(n[1]):
            for ix in range(n[0]):
                pts.append([np.float32(dmin[0]) + np.float32(ix) / np.float32(n[0]) * np.float32(dmax[0] - dmin[0]),
                            np.float32(dmin[1]) + np.float32(iy) / np.float32(n[1]) * np.float32(dmax[1] - dmin[1]),
                            np.float32(dmin[2]) + np.float32(iz) / np.float32(n[2]) * np.float32(dmax[2] - dmin[2])])
    pts = f32(pts)
    res = np.zeros(len(pts), np.float32)
    bmin = (C.c_float * 3)(*dmin)
    bmax = (C.c_float * 3)(*dmax)
    R.lib.ref_particles_sample(ptr(x, _f), ptr(y, _f), ptr(z, _f), ptr(v, _f), C.c_uint32(len(x)), C.c_float(h),
                               bmin, bmax, ptr(pts, _f), C.c_uint32(len(pts)), ptr(res, _f))
    # ::generate on the particle set itself: Particles::sample at every ray sample
    cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), fov_half_deg=24.0, aspect=1.0)
    gp = GenParams(step_size=0.02, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=0,
                   prediction_coding=0)
    mo, b = R.generate_particles(x, y, z, v, float(h), dmin, dmax, cam, 14, 14, 12, gp)
    return dict(sph_x=x, sph_y=y, sph_z=z, sph_v=v, sph_h=f32([h]), sph_grid=res.reshape(n[2], n[1], n[0]),
                sph_march_moments=mo, sph_march_bounds=b)


def quant_vectors(R: Reference) -> dict:
    """moment_quantization's packed streams (quantize_prediction_coding with the
    default and a loaded table, quantize at several widths) of compact images
    with empty pixels, out-of-range and signed-zero values, and their
    dequantized values where the stream holds every word whole."""
    from oracle_ffi import quant_bits

    rng = np.random.default_rng(2024)
    table = [16, 14, 12, 12, 11, 10, 10, 9, 9, 8, 8, 8, 7, 7, 7, 6, 6, 6, 5, 5, 5, 5, 4, 4, 4, 4, 3]
    cases = {"pc_default10": (1, quant_bits(256, 1, 10), True), "pc_table": (1, quant_bits(256, 1, table=table), True),
             "pc_ragged12": (1, quant_bits(256, 1, 12), False), "plain8": (0, quant_bits(256, 0, 8), True),
             "plain13": (0, quant_bits(256, 0, 13), True), "plain32": (0, quant_bits(256, 0, 32), True)}
    out = {}
    for name, (pc, bits, whole) in cases.items():
        W, H = 19, 11
        counts = rng.integers(0, 40, W * H).astype(np.uint32)
        counts[rng.random(W * H) < 0.2] = 0
        cum = np.concatenate([[0], np.cumsum(bits.astype(np.int64))])
        # the last pixel's count sets whether the stream ends on a byte boundary
        for c in range(1, 60):
            counts[-1] = c
            if (int(cum[counts].sum()) % 8 == 0) == whole:
                break
        index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
        n = int(index[-1])
        if pc:
            data = rng.uniform(-0.05, 1.05, n).astype(np.float32)
        else:
            data = rng.uniform(-0.34, 0.34, n).astype(np.float32)
        data[::13] = 0.0
        data[::17] = -0.0
        stream = R.quantize(W, H, data, index, pc, bits)
        out.update({f"{name}_index": index, f"{name}_data": data, f"{name}_pc": np.int32(pc),
                    f"{name}_bits": bits, f"{name}_stream": stream})
        if whole:
            out[f"{name}_decoded"] = R.dequantize(W, H, index, pc, bits, stream)
    return out


CURVE_CASES = {"ml": ("ml", 32, 32, 24, 1e-3), "mix": ("mix", 40, 40, 30, 1e-2), "mix64": ("mix", 64, 64, 40, 1e-3)}


def curve_vectors(R: Reference) -> dict:
    """create_quantization_table (the quantization-curve search) of coded
    compact images the reference generated, at start widths 8, 10 and 14."""
    import ctypes as C

    from oracle_ffi import ptr, _f, _u8, _u16, _u32

    out = {}
    for name, (kind, W, H, m, thr) in CURVE_CASES.items():
        if kind == "ml":
            vol, dmin, dmax, cam = synth.marschner_lobb(40), (-1, -1, -1), (1, 1, 1), synth.ortho_camera()
        else:
            vol, dmin, dmax = synth.mixing_layer(40, width=0.05), (0, 0, 0), (1, 1, 1)
            cam = synth.perspective_camera(position=(0.5, 0.5, -1.0), aspect=1.0)
        gp = GenParams(step_size=1.0 / 40, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=1,
                       prediction_coding=1, error_threshold=thr, bias=1e-4)
        data, _, index, cp = R.generate(vol, dmin, dmax, cam, W, H, m, gp)
        data = np.ascontiguousarray(data, np.float32)
        index = np.ascontiguousarray(index, np.uint32)
        cp = np.ascontiguousarray(cp, np.uint16)
        out.update({f"{name}_data": data, f"{name}_index": index, f"{name}_params": cp,
                    f"{name}_shape": np.array([W, H, m], np.int32)})
        for sb in (8, 10, 14):
            t = np.zeros(m, np.uint8)
            R.lib.ref_quantization_table(W, H, m, ptr(data, _f), ptr(index, _u32), 0, ptr(cp, _u16), sb, ptr(t, _u8))
            out[f"{name}_table{sb}"] = t
    return out


def _entropy_image(rng, W, H, m, pc, empty=0.2):
    counts = rng.integers(1, m + 1, W * H).astype(np.uint32)
    counts[rng.random(W * H) < empty] = 0
    index = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32)
    n = int(index[-1])
    if pc:
        data = rng.beta(2.0, 2.0, n).astype(np.float32)
    else:
        data = rng.normal(0.0, 0.08, n).astype(np.float32)
        data[index[:-1][counts > 0]] = rng.random(int(np.count_nonzero(counts)))
    return index, data


UNC_CASES = {
    # name: (volume, W, H, m, compact, prediction_coded, camera, gen step, error step)
    "ml_ortho": ("ml", 24, 20, 16, 1, 0, "ortho", 1.0 / 32, 1.0 / 40),
    "mix_persp_u8": ("mix_u8", 26, 22, 12, 1, 0, "persp", 1.0 / 32, 1.0 / 24),
    "ml_coded": ("ml", 20, 18, 14, 1, 1, "ortho", 1.0 / 32, 1.0 / 40),
    "ml_dense_coarse": ("ml", 16, 14, 10, 0, 0, "ortho", 1.0 / 32, 0.45),
}


def uncertainty_volume(kind):
    if kind == "ml":
        return synth.marschner_lobb(32), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    vol = np.clip(synth.mixing_layer(32, width=0.05) * 255, 0, 255).astype(np.uint8)
    return vol, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)


def uncertainty_camera(kind, W, H):
    if kind == "ortho":
        return synth.ortho_camera()
    return synth.perspective_camera(position=(0.5, 0.5, -1.2), fov_half_deg=30.0, znear=0.01, aspect=W / H)


def uncertainty_vectors(R: Reference) -> dict:
    """Error bounds (reconstruct_error: max, RMSE and the P^2 percentile),
    the uncertainty transfer functions (box and gaussian filters; every 17th
    row and column plus a SHA-256 of each full table) and the uncertainty
    renderings (interpolation, convolution per error type) of the reference."""
    import hashlib

    from mray_amd.types import ErrParams

    out = {}
    rgba8 = synth.tf_ramps(0)
    for et in (0, 1):
        t = R.tf_uncertainty(rgba8, 0.1, et)
        out[f"tf{et}_sub"] = t[::17, ::17].copy()
        out[f"tf{et}_sha256"] = np.frombuffer(hashlib.sha256(t.tobytes()).digest(), np.uint8)
    for name, (vk, W, H, m, compact, pc, ck, gstep, estep) in UNC_CASES.items():
        vol, dmin, dmax = uncertainty_volume(vk)
        cam = uncertainty_camera(ck, W, H)
        gp = GenParams(step_size=gstep, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=compact,
                       prediction_coding=pc, error_threshold=2e-3, bias=1e-4)
        data, bounds, index, cp = R.generate(vol, dmin, dmax, cam, W, H, m, gp)
        data = np.ascontiguousarray(data, np.float32)
        idx = index if compact else None
        cpq = cp if pc else None
        out.update({f"{name}_data": data, f"{name}_bounds": bounds, f"{name}_shape": np.array([W, H, m], np.int32)})
        if compact:
            out[f"{name}_index"] = index
        if pc:
            out[f"{name}_coding_params"] = cp
        for et in (0, 1, 2):
            ep = ErrParams(error_type=et, error_percentile=90 if et == 2 else 95, bias=1e-4, step_size=estep)
            out[f"{name}_err{et}"] = R.error_bounds(vol, dmin, dmax, cam, W, H, m, data, ep, index=idx, bounds=bounds,
                                                   prediction_coded=pc, coding_params=cpq)
        rp = RecParams(step_size=1.0 / 40, bias=1e-4, early_ray_termination=0.99, medium_emission=1.0)
        for mode, et, s in ((0, 2, 0.3), (1, 0, 0.0), (1, 1, 0.0), (1, 2, 0.0)):
            eb = out[f"{name}_err{et}"]
            rgba, rgba8_img = R.uncertainty(cam, dmin, dmax, W, H, m, data, eb, rgba8, rp.step_size * 4, rp, mode, et, s,
                                            index=idx, bounds=bounds, prediction_coded=pc, coding_params=cpq,
                                            want_rgba8=True)
            out[f"{name}_unc{mode}{et}"] = rgba
            out[f"{name}_unc{mode}{et}_rgba8"] = rgba8_img
    return out


def halving_data(W, H, levels, level):
    """The 2-moment image of the halving case: moment 0 = (p % 97) / 97,
    moment 1 = levels[level[p]]."""
    data = np.empty(2 * W * H, np.float32)
    data[0::2] = (np.arange(W * H) % 97) / np.float32(97.0)
    data[1::2] = levels[level]
    return data


def entropy_vectors(R: Reference) -> dict:
    """entropy_encode / entropy_decode (the FastAC adaptive arithmetic coder
    per moment) of compact images: the reference's section bytes and the
    values it decodes from them."""
    rng = np.random.default_rng(77)
    out = {}
    table12 = np.array([16, 11, 10, 10, 9, 8, 8, 7, 6, 6, 5, 4], np.uint8)
    cases = {"plain8": (16, 12, 10, 0, np.full(10, 8, np.uint8)),
             "pc_table": (23, 17, 12, 1, table12),
             "w14": (20, 20, 6, 0, np.full(6, 14, np.uint8)),
             "w1": (30, 7, 5, 0, np.full(5, 1, np.uint8)),
             "pc_w32_zeroth": (9, 9, 4, 1, np.array([32, 8, 8, 8], np.uint8))}
    for name, (W, H, m, pc, bits) in cases.items():
        index, data = _entropy_image(rng, W, H, m, pc)
        sec = R.entropy_encode(W, H, m, data, index, pc, bits)
        out.update({f"{name}_shape": np.array([W, H, m, pc], np.int32), f"{name}_index": index,
                    f"{name}_data": data, f"{name}_bits": bits, f"{name}_section": sec,
                    f"{name}_decoded": R.entropy_decode(W, H, m, sec, index, pc, bits)})
    # a stream long enough for the model's count halving (total > 2^17):
    # 380 x 380 pixels of 2 moments (the index is 2 p, not stored), moment 1
    # drawn from 24 levels with a skewed distribution (a compact fixture)
    W = H = 380
    index = (np.arange(W * H + 1) * 2).astype(np.uint32)
    levels = np.linspace(-0.3, 0.3, 24).astype(np.float32)
    p = np.exp(-0.5 * ((np.arange(24) - 11.5) / 2.0) ** 2)
    level = rng.choice(24, W * H, p=p / p.sum()).astype(np.uint8)
    data = halving_data(W, H, levels, level)
    bits = np.array([16, 7], np.uint8)
    sec = R.entropy_encode(W, H, 2, data, index, 0, bits)
    dec = R.entropy_decode(W, H, 2, sec, index, 0, bits)
    import hashlib

    # the section (mostly the 16-bit moment-0 words) is kept as its SHA-256;
    # the decoded moment 1 as the set of its values per level
    out.update({"halving_shape": np.array([W, H, 2, 0], np.int32), "halving_levels": levels,
                "halving_level": level, "halving_bits": bits, "halving_section_len": np.int64(len(sec)),
                "halving_section_sha256": np.frombuffer(hashlib.sha256(sec.tobytes()).digest(), np.uint8),
                "halving_stream1_tail": sec[-64:],
                "halving_decoded_levels": np.array([dec[1::2][level == k][0] if np.any(level == k) else np.nan
                                                    for k in range(24)], np.float32)})
    return out


# the moment-image files, in writing order: the quantization table is set
# before the "tab" cases and, as in the reference, stays set after them
FILE_CASES = [("c_f32", "A", 0, 32, 0), ("c_q10", "A", 0, 10, 0), ("c_lz4_f32", "A", 1, 32, 0),
              ("c_lz4_q10", "A", 1, 10, 0), ("c_ent8", "A", 0, 8, 1), ("c_lz4_ent10", "A", 1, 10, 1),
              ("c_eb_lz4", "E", 1, 32, 0), ("d_f32", "D", 0, 32, 0), ("d_q10_raw", "D", 0, 10, 0),
              ("d_lz4_f32", "D", 1, 32, 0), ("d_lz4_q12", "D", 1, 12, 0), ("d_lz4_q16", "D", 1, 16, 0),
              ("pc_q10", "P", 0, 10, 0), ("pc_lz4_q10", "P", 1, 10, 0), ("pc_tab_q10", "P", 0, 10, 0),
              ("pc_tab_lz4_ent10", "P", 1, 10, 1), ("c_tab_lz4_ent10", "A", 1, 10, 1)]
# the reference's reader cannot read compressed quantized dense images: it
# hands the quantization byte in front of the LZ4 block to LZ4
FILE_UNREADABLE = ("d_lz4_q12", "d_lz4_q16")


def _file_images(R: Reference) -> dict:
    vol, dmin, dmax = synth.marschner_lobb(32), (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    cam = synth.ortho_camera()
    imgs = {}
    for key, (W, H, m, compact, pc) in {"A": (24, 20, 20, 1, 0), "D": (16, 12, 12, 0, 0),
                                        "P": (24, 20, 20, 1, 1)}.items():
        gp = GenParams(step_size=1.0 / 32, transform_to_bounds=1, transform_bound_eps=0.01, compact_image=compact,
                       prediction_coding=pc, error_threshold=2e-3, bias=1e-4)
        data, bounds, index, cp = R.generate(vol, dmin, dmax, cam, W, H, m, gp)
        imgs[key] = dict(W=W, H=H, m=m, dmin=dmin, dmax=dmax, prediction_coded=pc, warp=0,
                         data=np.ascontiguousarray(data, np.float32), bounds=bounds,
                         index=index if compact else None, coding_params=cp if pc else None)
    e = dict(imgs["A"])
    e["error_bounds"] = np.random.default_rng(5).integers(0, 65536, e["W"] * e["H"]).astype(np.uint16)
    imgs["E"] = e
    return imgs


def _files_worker(outfile):
    """Writes every file case in a fresh process (the reference's table is
    process state) and saves inputs, file bytes and what the reference reads."""
    import tempfile

    from oracle_ffi import ptr, _f, _u8, _u16, _u32

    R = Reference()
    imgs = _file_images(R)
    out = {}
    for key, img in imgs.items():
        for k in ("data", "bounds", "index", "coding_params", "error_bounds"):
            if img.get(k) is not None:
                out[f"img{key}_{k}"] = img[k]
        out[f"img{key}_shape"] = np.array([img["W"], img["H"], img["m"], img["prediction_coded"]], np.int32)
    P = imgs["P"]
    table = None
    with tempfile.TemporaryDirectory() as td:
        for name, key, compress, bits, entropy in FILE_CASES:
            if name.startswith("pc_tab") and table is None:
                # the quantization curve of image P at 10 bits sets the table
                table = np.zeros(P["m"], np.uint8)
                R.lib.ref_quantization_table(P["W"], P["H"], P["m"], ptr(P["data"], _f), ptr(P["index"], _u32), 0,
                                             ptr(P["coding_params"], _u16), 10, ptr(table, _u8))
                out["table"] = table
            path = pathlib.Path(td) / f"{name}.bin"
            R.write_moment_image(path, imgs[key], compress, bits, entropy)
            out[f"{name}_file"] = np.frombuffer(path.read_bytes(), np.uint8).copy()
            if name not in FILE_UNREADABLE:
                out[f"{name}_read"] = R.load_moment_image(path, 0)["data"]
    np.savez_compressed(outfile, **out)


def file_vectors(R: Reference) -> dict:
    import subprocess
    import tempfile

    with tempfile.TemporaryDirectory() as td:
        f = pathlib.Path(td) / "files.npz"
        subprocess.run([sys.executable, __file__, "_files_worker", str(f)], check=True)
        with np.load(f) as z:
            out = {k: z[k] for k in z.files}
    out["order"] = np.array([c[0] for c in FILE_CASES])
    return out


def main():
    # python make_golden.py [name ...]: only the named fixtures (default all)
    R = Reference()
    if sys.argv[1:2] == ["_files_worker"]:
        _files_worker(sys.argv[2])
        return
    makers = {"mese": mese_vectors, "coding": coding_vectors, "images": image_vectors, "sph": sph_vectors,
              "reference": reference_dvr_vectors, "quant": quant_vectors,
              "curve": curve_vectors, "entropy": entropy_vectors, "files": file_vectors,
              "uncertainty": uncertainty_vectors}
    for name in (sys.argv[1:] or list(makers)):
        np.savez_compressed(OUT / f"golden_{name}.npz", **makers[name](R))
    for f in sorted(OUT.glob("golden_*.npz")):
        print(f.name, f.stat().st_size)


if __name__ == "__main__":
    main()
