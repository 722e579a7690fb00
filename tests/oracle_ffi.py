"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, where it was
built, the compiled reference (oracle/_ref/libmrayref.so). Test-side only."""
from __future__ import annotations

import ctypes as C
import os
import pathlib

import numpy as np

from mray_amd.types import Camera, ErrParams, GenParams, RecParams, RefParams

ROOT = pathlib.Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libmrayref.so"

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_u32 = C.POINTER(C.c_uint32)
_i32 = C.POINTER(C.c_int32)
_u16 = C.POINTER(C.c_uint16)
_u8 = C.POINTER(C.c_uint8)


class OrcVolume(C.Structure):
    _fields_ = [("data", C.c_void_p), ("dtype", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
                ("nz", C.c_int32), ("dmin", C.c_float * 3), ("dmax", C.c_float * 3)]


def ptr(a, t):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(t)


def _load(path):
    return C.CDLL(str(path)) if path.exists() else None


def oracle():
    lib = _load(ORACLE_SO)
    if lib is None:
        raise RuntimeError(f"{ORACLE_SO} missing: run `make -C oracle`")
    return lib


def reference():
    return _load(REF_SO)


def nthreads():
    return int(os.environ.get("MRAY_ORACLE_THREADS", "0"))


def quant_bits(m, prediction_coded, bits=10, table=None):
    """The per-moment word widths the reference quantizes with: the table when
    one is given, else get_prediction_coding_quantization_table's default
    (16 bits for moment 0, then `bits`) for prediction-coded images and
    `bits` for every moment otherwise (src/moment_quantization.cpp:186-203)."""
    t = np.full(256, bits, np.uint8)
    if table is not None:
        t[:len(table)] = table
    elif prediction_coded:
        t[0] = 16
    return t


def entropy_capacity(index, m):
    """An upper bound of an entropy-coded section's length: the layout's
    headers, 2 bytes per moment 0 and 4 bytes (+16 per stream) per symbol."""
    npix = len(index) - 1
    return 4 + 2 * npix + (m - 1) * 20 + 4 * int(index[-1]) + 64


def stream_bytes(index, bits):
    """floor(total bits / 8): the length of the reference's packed stream."""
    counts = np.diff(index.astype(np.int64))
    cum = np.concatenate([[0], np.cumsum(bits.astype(np.int64))])
    return int(cum[counts].sum()) // 8


class Particles:
    """A particle set of the oracle (orc_sph_create): Particles::create_grid
    of the reference; pass it to Oracle.generate in place of a volume."""

    def __init__(self, oracle, x, y, z, v, h):
        self._lib = oracle.lib
        self._arrays = [np.ascontiguousarray(a, np.float32) for a in (x, y, z, v)]
        fp = C.POINTER(C.c_float)
        self._lib.orc_sph_create.restype = C.c_void_p
        self._lib.orc_sph_create.argtypes = [fp, fp, fp, fp, C.c_uint32, C.c_float]
        self._lib.orc_sph_destroy.argtypes = [C.c_void_p]
        self.handle = self._lib.orc_sph_create(*[a.ctypes.data_as(fp) for a in self._arrays], len(self._arrays[0]),
                                               C.c_float(h))

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.orc_sph_destroy(self.handle)
            self.handle = None


class Oracle:
    """numpy-level wrappers over liboracle.so."""

    def __init__(self):
        self.lib = oracle()
        L = self.lib
        L.orc_fast_atan.restype = C.c_float
        L.orc_fast_atan.argtypes = [C.c_float]
        L.orc_evaluate_mese.restype = C.c_float
        L.orc_evaluate_mese.argtypes = [C.c_int, C.c_float, _f]
        L.orc_evaluate_mese_f64.restype = C.c_double
        L.orc_evaluate_mese_f64.argtypes = [C.c_int, C.c_double, _d]
        L.orc_determine_best_number_moments.restype = C.c_int
        L.orc_determine_best_number_moments.argtypes = [C.c_int, _f, C.c_float]
        L.orc_rmse.restype = C.c_float
        L.orc_compact.restype = C.c_uint32
        L.orc_volume_sample.restype = C.c_float
        L.orc_cubic_spline.restype = C.c_float
        L.orc_cubic_spline.argtypes = [C.c_float]
        L.orc_generate_ray.restype = C.c_uint32

    @staticmethod
    def volume(vol, dmin, dmax):
        v = OrcVolume()
        if isinstance(vol, Particles):
            # a particle set sampled by Particles::sample at every ray sample
            v.data = vol.handle
            v.dtype = 2
            v.nz = v.ny = v.nx = 1
            for i in range(3):
                v.dmin[i] = dmin[i]
                v.dmax[i] = dmax[i]
            v._keep = vol
            return v
        v.data = vol.ctypes.data
        v.dtype = 1 if vol.dtype == np.uint8 else 0
        v.nz, v.ny, v.nx = vol.shape
        for i in range(3):
            v.dmin[i] = dmin[i]
            v.dmax[i] = dmax[i]
        v._keep = vol
        return v

    def generate(self, vol, dmin, dmax, cam: dict, W, H, m, gp: GenParams, row0=0, row_step=1, f64=False):
        v = self.volume(vol, dmin, dmax)
        c = Camera.from_dict(cam)
        bounds = np.zeros(W * H, np.uint32)
        if f64:
            mo = np.zeros((H, W, m), np.float64)
            self.lib.orc_generate_image_f64(C.byref(v), C.byref(c), W, H, m, C.byref(gp), ptr(mo, _d),
                                            ptr(bounds, _u32), row0, row_step, nthreads())
            return mo, bounds
        mo = np.zeros((H, W, m), np.float32)
        ns = np.zeros(W * H, np.int32)
        self.lib.orc_generate_image(C.byref(v), C.byref(c), W, H, m, C.byref(gp), ptr(mo, _f), ptr(bounds, _u32),
                                    ptr(ns, _i32), row0, row_step, nthreads())
        return mo, bounds, ns

    def tf(self, rgba8: np.ndarray, step: float) -> np.ndarray:
        out = np.zeros((256, 4), np.float32)
        self.lib.orc_tf_from_rgba8(ptr(np.ascontiguousarray(rgba8, np.uint8), _u8), C.c_float(step), ptr(out, _f))
        return out

    def reconstruct(self, cam, dmin, dmax, W, H, m, data, tf, rp: RecParams, index=None, bounds=None,
                    prediction_coded=0, coding_warp=0, coding_params=None, row0=0, row_step=1):
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float32)
        ns = np.zeros(W * H, np.int32)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.orc_reconstruct_image(C.byref(c), bmin, bmax, W, H, m, ptr(data, _f), ptr(index, _u32),
                                       ptr(bounds, _u32), prediction_coded, coding_warp, ptr(coding_params, _f),
                                       ptr(tf, _f), C.byref(rp), ptr(rgba, _f), ptr(ns, _i32), row0, row_step,
                                       nthreads())
        return rgba, ns

    def reference(self, vol, dmin, dmax, cam, W, H, tf, p: RefParams, row0=0, row_step=1):
        """Reference direct volume rendering (orc_reference_image): float RGBA and steps per ray."""
        v = self.volume(vol, dmin, dmax)
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float32)
        ns = np.zeros(W * H, np.int32)
        self.lib.orc_reference_image(C.byref(v), C.byref(c), W, H, ptr(tf, _f), C.byref(p), ptr(rgba, _f),
                                     ptr(ns, _i32), row0, row_step, nthreads())
        return rgba, ns

    def reconstruct_f64(self, cam, dmin, dmax, W, H, m, moments64, tf, rp: RecParams, bounds=None, row0=0,
                        row_step=1):
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float64)
        ns = np.zeros(W * H, np.int32)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.orc_reconstruct_rows_f64(C.byref(c), bmin, bmax, W, H, m, ptr(moments64, _d), ptr(bounds, _u32),
                                          ptr(tf, _f), C.byref(rp), ptr(rgba, _d), ptr(ns, _i32), row0, row_step,
                                          nthreads())
        return rgba, ns

    def reconstruct_lagrange_f64(self, cam, dmin, dmax, W, H, m, lag32, tf, rp: RecParams, bounds=None):
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float64)
        ns = np.zeros(W * H, np.int32)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.orc_reconstruct_lagrange_f64(C.byref(c), bmin, bmax, W, H, m,
                                              ptr(np.ascontiguousarray(lag32, np.float32), _f), ptr(bounds, _u32),
                                              ptr(tf, _f), C.byref(rp), ptr(rgba, _d), ptr(ns, _i32), nthreads())
        return rgba, ns

    def reconstruct_lagrange64_f64(self, cam, dmin, dmax, W, H, m, lag64, tf64, rp: RecParams, bounds=None):
        """FP64 march on FP64 multipliers and an FP64 table (finite-difference reference)."""
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float64)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.orc_reconstruct_lagrange64_f64(C.byref(c), bmin, bmax, W, H, m,
                                                ptr(np.ascontiguousarray(lag64, np.float64), _d), ptr(bounds, _u32),
                                                ptr(np.ascontiguousarray(tf64, np.float64), _d), C.byref(rp),
                                                ptr(rgba, _d), None, nthreads())
        return rgba

    def reconstruct_adjoint_f64(self, cam, dmin, dmax, W, H, m, lag32, tf, rp: RecParams, g_rgba, bounds=None):
        """(dL/dlambda [H*W*m], dL/dtable [256, 4]) of L = sum(g_rgba * rgba) in FP64."""
        c = Camera.from_dict(cam)
        gl = np.zeros(W * H * m, np.float64)
        gt = np.zeros((256, 4), np.float64)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.orc_reconstruct_adjoint_f64(C.byref(c), bmin, bmax, W, H, m,
                                             ptr(np.ascontiguousarray(lag32, np.float32), _f), ptr(bounds, _u32),
                                             ptr(np.ascontiguousarray(tf, np.float32), _f), C.byref(rp),
                                             ptr(np.ascontiguousarray(g_rgba, np.float64), _d), ptr(gl, _d),
                                             ptr(gt, _d), nthreads())
        return gl, gt

    def prepare_std(self, m, moments, bias):
        out = np.zeros(m, np.float32)
        self.lib.orc_prepare_moments_std(m, ptr(np.ascontiguousarray(moments, np.float32), _f), C.c_float(bias),
                                         ptr(out, _f))
        return out

    def prepare_std_f64(self, m, moments, bias):
        out = np.zeros(m, np.float64)
        self.lib.orc_prepare_moments_std_f64(m, ptr(np.ascontiguousarray(moments, np.float64), _d),
                                             C.c_double(bias), ptr(out, _d))
        return out

    def num_moments(self, W, H, m, dense, bounds, thr):
        counts = np.zeros(W * H, np.uint32)
        self.lib.orc_compute_num_moments_image(W, H, m, ptr(dense, _f), ptr(bounds, _u32), C.c_float(thr),
                                               ptr(counts, _u32), nthreads())
        return counts

    def compact(self, W, H, m, dense, counts):
        index = np.zeros(W * H + 1, np.uint32)
        total = int(counts.astype(np.int64).sum())
        out = np.zeros(max(total, 1), np.float32)
        self.lib.orc_compact(W, H, m, ptr(dense, _f), ptr(counts, _u32), ptr(index, _u32), ptr(out, _f))
        return index, out[:total]

    def prediction_encode(self, W, H, m, data, index, warp, bias):
        d = np.ascontiguousarray(data, np.float32).copy()
        q = np.zeros(max(2 * (m - 1), 1), np.uint16)
        self.lib.orc_prediction_encode_image(W, H, m, ptr(d, _f), ptr(index, _u32), warp, C.c_float(bias),
                                             ptr(q, _u16), nthreads())
        return d, q

    def prepare_pred(self, nm, moments, warp, params_f):
        """prepare_moments_from_pred_coding on one pixel's nm coded moments."""
        out = np.zeros(max(nm, 1), np.float32)
        pf = np.ascontiguousarray(params_f if params_f is not None else np.zeros(max(2 * (nm - 1), 1)), np.float32)
        self.lib.orc_prepare_moments_pred(nm, ptr(np.ascontiguousarray(moments, np.float32), _f), warp, ptr(pf, _f),
                                          ptr(out, _f))
        return out[:nm]

    def quantize(self, W, H, data, index, prediction_coded, bits):
        out = np.zeros(stream_bytes(index, bits) + 8, np.uint8)
        self.lib.orc_quantize_image.restype = C.c_size_t
        n = self.lib.orc_quantize_image(W, H, ptr(np.ascontiguousarray(data, np.float32), _f), ptr(index, _u32),
                                        prediction_coded, ptr(bits, _u8), ptr(out, _u8))
        return out[:n]

    def dequantize(self, W, H, index, prediction_coded, bits, stream):
        out = np.zeros(max(int(index[-1]), 1), np.float32)
        st = np.ascontiguousarray(stream, np.uint8)
        self.lib.orc_dequantize_image(W, H, ptr(index, _u32), prediction_coded, ptr(bits, _u8), ptr(st, _u8),
                                      C.c_size_t(len(st)), ptr(out, _f))
        return out[:int(index[-1])]

    def tf_uncertainty(self, rgba8, step, error_type):
        """The 2D uncertainty table (256, 256, 4): box filter for the max error
        and the percentile, gaussian for the RMSE."""
        out = np.zeros((256, 256, 4), np.float32)
        fn = self.lib.orc_tf_gaussian_filter if error_type == 1 else self.lib.orc_tf_box_filter
        fn(ptr(np.ascontiguousarray(rgba8, np.uint8), _u8), C.c_float(step), ptr(out, _f))
        return out

    def error_bounds(self, vol, dmin, dmax, cam, W, H, m, data, ep: ErrParams, index=None, bounds=None,
                     prediction_coded=0, coding_warp=0, coding_params=None):
        v = self.volume(vol, dmin, dmax)
        c = Camera.from_dict(cam)
        out = np.zeros(W * H, np.uint16)
        self.lib.orc_error_bounds(C.byref(v), C.byref(c), W, H, m, ptr(np.ascontiguousarray(data, np.float32), _f),
                                  ptr(index, _u32), ptr(bounds, _u32), prediction_coded, coding_warp,
                                  ptr(coding_params, _f), C.byref(ep), ptr(out, _u16), nthreads())
        return out

    def uncertainty(self, cam, dmin, dmax, W, H, m, data, error_bounds, tf, rp: RecParams, mode, error_type=1,
                    bound_interpolation=0.5, index=None, bounds=None, prediction_coded=0, coding_warp=0,
                    coding_params=None):
        """Float RGBA of the uncertainty renderings (mode 0 interpolation with
        the preintegrated table tf (256, 4); mode 1 convolution with a 2D
        table tf (256, 256, 4))."""
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float32)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.orc_uncertainty_image(C.byref(c), bmin, bmax, W, H, m,
                                       ptr(np.ascontiguousarray(data, np.float32), _f), ptr(index, _u32),
                                       ptr(bounds, _u32), ptr(np.ascontiguousarray(error_bounds, np.uint16), _u16),
                                       prediction_coded, coding_warp, ptr(coding_params, _f),
                                       ptr(np.ascontiguousarray(tf, np.float32), _f), C.byref(rp), mode, error_type,
                                       C.c_float(bound_interpolation), ptr(rgba, _f), nthreads())
        return rgba

    def entropy_encode(self, W, H, m, data, index, prediction_coded, bits):
        """The entropy-coded moment section (bytes) of a compact image."""
        cap = entropy_capacity(index, m)
        out = np.zeros(cap, np.uint8)
        self.lib.orc_entropy_encode.restype = C.c_long
        n = self.lib.orc_entropy_encode(W, H, m, ptr(np.ascontiguousarray(data, np.float32), _f), ptr(index, _u32),
                                        prediction_coded, ptr(np.ascontiguousarray(bits, np.uint8), _u8),
                                        ptr(out, _u8), C.c_size_t(cap))
        assert 0 <= n <= cap, n
        return out[:n]

    def entropy_decode(self, W, H, m, section, index, prediction_coded, bits):
        out = np.zeros(max(int(index[-1]), 1), np.float32)
        sec = np.ascontiguousarray(section, np.uint8)
        rc = self.lib.orc_entropy_decode(W, H, m, ptr(sec, _u8), C.c_size_t(len(sec)), ptr(index, _u32),
                                         prediction_coded, ptr(np.ascontiguousarray(bits, np.uint8), _u8),
                                         ptr(out, _f))
        assert rc == 0
        return out[:int(index[-1])]

    def particles_to_grid(self, x, y, z, v, h, dmin, dmax, nx, ny, nz):
        g = np.zeros((nz, ny, nx), np.float32)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.orc_particles_to_grid(ptr(x, _f), ptr(y, _f), ptr(z, _f), ptr(v, _f), C.c_uint32(len(x)),
                                       C.c_float(h), bmin, bmax, nx, ny, nz, ptr(g, _f), nthreads())
        return g


class Reference:
    """numpy-level wrappers over the compiled reference (build container only)."""

    def __init__(self):
        self.lib = reference()
        if self.lib is None:
            raise RuntimeError("oracle/_ref/libmrayref.so not built")
        L = self.lib
        L.ref_generate.restype = C.c_long
        L.ref_evaluate_mese.restype = C.c_float
        L.ref_evaluate_mese.argtypes = [C.c_int, C.c_float, _f]
        L.ref_fast_atan.restype = C.c_float
        L.ref_fast_atan.argtypes = [C.c_float]
        L.ref_best_num_moments.restype = C.c_int
        L.ref_best_num_moments.argtypes = [C.c_int, _f, C.c_float]
        L.ref_fixed_evaluate4.restype = C.c_float
        L.ref_fixed_evaluate4.argtypes = [C.c_float, _f]
        L.ref_fixed_evaluate10.restype = C.c_float
        L.ref_fixed_evaluate10.argtypes = [C.c_float, _f]

    def generate(self, vol, dmin, dmax, cam, W, H, m, gp: GenParams):
        c = Camera.from_dict(cam)
        data = np.zeros(W * H * m, np.float32)
        bounds = np.zeros(W * H, np.uint32)
        index = np.zeros(W * H + 1, np.uint32)
        cp = np.zeros(max(2 * (m - 1), 1), np.uint16)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        nz, ny, nx = vol.shape
        n = self.lib.ref_generate(vol.ctypes.data_as(C.c_void_p), 1 if vol.dtype == np.uint8 else 0, nx, ny, nz,
                                  bmin, bmax, C.byref(c), W, H, m, C.byref(gp), ptr(data, _f), ptr(bounds, _u32),
                                  ptr(index, _u32), ptr(cp, _u16))
        return data[:n], bounds, index, cp

    def generate_particles(self, x, y, z, v, h, dmin, dmax, cam, W, H, m, gp: GenParams):
        """::generate on a Particles (per-sample Particles::sample): dense moments, bounds."""
        c = Camera.from_dict(cam)
        data = np.zeros(W * H * m, np.float32)
        bounds = np.zeros(W * H, np.uint32)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        arr = [np.ascontiguousarray(a, np.float32) for a in (x, y, z, v)]
        self.lib.ref_generate_particles(*[ptr(a, _f) for a in arr], C.c_uint32(len(arr[0])), C.c_float(h), bmin,
                                        bmax, C.byref(c), W, H, m, C.byref(gp), ptr(data, _f), ptr(bounds, _u32))
        return data.reshape(H, W, m), bounds

    def quantize(self, W, H, data, index, prediction_coded, bits):
        out = np.zeros(stream_bytes(index, bits) + 8, np.uint8)
        self.lib.ref_quantize.restype = C.c_long
        n = self.lib.ref_quantize(W, H, ptr(np.ascontiguousarray(data, np.float32), _f), ptr(index, _u32),
                                  prediction_coded, ptr(bits, _u8), ptr(out, _u8))
        return out[:n]

    def dequantize(self, W, H, index, prediction_coded, bits, stream):
        out = np.zeros(max(int(index[-1]), 1), np.float32)
        st = np.ascontiguousarray(stream, np.uint8)
        self.lib.ref_dequantize(W, H, ptr(index, _u32), prediction_coded, ptr(bits, _u8), ptr(st, _u8),
                                C.c_long(len(st)), ptr(out, _f))
        return out[:int(index[-1])]

    def tf_uncertainty(self, rgba8, step, error_type):
        out = np.zeros((256, 256, 4), np.float32)
        self.lib.ref_tf_uncertainty(ptr(np.ascontiguousarray(rgba8, np.uint8), _u8), C.c_float(step), error_type,
                                    ptr(out, _f))
        return out

    def error_bounds(self, vol, dmin, dmax, cam, W, H, m, data, ep: ErrParams, index=None, bounds=None,
                     prediction_coded=0, coding_warp=0, coding_params=None):
        c = Camera.from_dict(cam)
        out = np.zeros(W * H, np.uint16)
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        nz, ny, nx = vol.shape
        data = np.ascontiguousarray(data, np.float32)
        self.lib.ref_error_bounds(vol.ctypes.data_as(C.c_void_p), 1 if vol.dtype == np.uint8 else 0, nx, ny, nz,
                                  bmin, bmax, C.byref(c), W, H, m, ptr(data, _f), C.c_long(len(data)),
                                  ptr(index, _u32), ptr(bounds, _u32), prediction_coded, coding_warp,
                                  ptr(coding_params, _u16), C.byref(ep), ptr(out, _u16))
        return out

    def uncertainty(self, cam, dmin, dmax, W, H, m, data, error_bounds, tf_rgba8, tf_step, rp: RecParams, mode,
                    error_type=1, bound_interpolation=0.5, index=None, bounds=None, prediction_coded=0,
                    coding_warp=0, coding_params=None, want_rgba8=False):
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float32)
        rgba8 = np.zeros((H, W, 4), np.uint8) if want_rgba8 else None
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        data = np.ascontiguousarray(data, np.float32)
        self.lib.ref_uncertainty(C.byref(c), bmin, bmax, W, H, m, ptr(data, _f), C.c_long(len(data)),
                                 ptr(index, _u32), ptr(bounds, _u32),
                                 ptr(np.ascontiguousarray(error_bounds, np.uint16), _u16), prediction_coded,
                                 coding_warp, ptr(coding_params, _u16),
                                 ptr(np.ascontiguousarray(tf_rgba8, np.uint8), _u8), C.c_float(tf_step), C.byref(rp),
                                 mode, error_type, C.c_float(bound_interpolation), ptr(rgba, _f), ptr(rgba8, _u8))
        return rgba, rgba8

    def entropy_encode(self, W, H, m, data, index, prediction_coded, bits):
        cap = entropy_capacity(index, m)
        out = np.zeros(cap, np.uint8)
        self.lib.ref_entropy_encode.restype = C.c_long
        n = self.lib.ref_entropy_encode(W, H, m, ptr(np.ascontiguousarray(data, np.float32), _f), ptr(index, _u32),
                                        prediction_coded, ptr(np.ascontiguousarray(bits, np.uint8), _u8),
                                        ptr(out, _u8), C.c_long(cap))
        assert 0 <= n <= cap, n
        return out[:n]

    def entropy_decode(self, W, H, m, section, index, prediction_coded, bits):
        out = np.zeros(max(int(index[-1]), 1), np.float32)
        # the reference's decoder reads two bytes past a stream; give it zeros
        sec = np.concatenate([np.ascontiguousarray(section, np.uint8), np.zeros(16, np.uint8)])
        self.lib.ref_entropy_decode(W, H, m, ptr(sec, _u8), C.c_long(len(section)), ptr(index, _u32),
                                    prediction_coded, ptr(np.ascontiguousarray(bits, np.uint8), _u8), ptr(out, _f))
        return out[:int(index[-1])]

    def write_moment_image(self, path, img, compress, bits, entropy, table=None):
        """write_moment_image of a MomentFile-like image (see moment_file())."""
        bmin = (C.c_float * 3)(*img["dmin"])
        bmax = (C.c_float * 3)(*img["dmax"])
        data = np.ascontiguousarray(img["data"], np.float32)
        opt = lambda k, t: None if img.get(k) is None else np.ascontiguousarray(img[k], t)
        index, bounds, eb, cp = opt("index", np.uint32), opt("bounds", np.uint32), opt("error_bounds", np.uint16), \
            opt("coding_params", np.uint16)
        tb = None if table is None else np.ascontiguousarray(table, np.uint8)
        self.lib.ref_write_moment_image(str(path).encode(), img["W"], img["H"], img["m"], int(index is not None),
                                        int(img.get("prediction_coded", 0)), int(img.get("warp", 0)), bmin, bmax,
                                        ptr(index, _u32), ptr(data, _f), C.c_long(len(data)), ptr(bounds, _u32),
                                        ptr(eb, _u16), ptr(cp, _u16), ptr(tb, _u8), int(compress), int(bits),
                                        int(entropy))

    def load_moment_image(self, path, n_data_max):
        hdr = np.zeros(10, np.int64)
        self.lib.ref_load_moment_image(str(path).encode(), ptr(hdr, C.POINTER(C.c_long)), None, None, None, None)
        W, H, m = int(hdr[0]), int(hdr[1]), int(hdr[2])
        index = np.zeros(W * H + 1, np.uint32) if hdr[4] else None
        data = np.zeros(max(int(hdr[9]), 1), np.float32)
        bounds = np.zeros(W * H, np.uint32) if hdr[6] else None
        cp = np.zeros(max(2 * (m - 1), 1), np.uint16) if hdr[8] else None
        self.lib.ref_load_moment_image(str(path).encode(), ptr(hdr, C.POINTER(C.c_long)), ptr(index, _u32),
                                       ptr(data, _f), ptr(bounds, _u32), ptr(cp, _u16))
        return dict(W=W, H=H, m=m, warp=int(hdr[3]), compact=bool(hdr[4]), prediction_coded=bool(hdr[5]),
                    index=index, data=data[:int(hdr[9])], bounds=bounds, coding_params=cp)

    def reference(self, vol, dmin, dmax, cam, W, H, tf_rgba8, tf_step, p: RefParams, want_rgba8=False):
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float32)
        rgba8 = np.zeros((H, W, 4), np.uint8) if want_rgba8 else None
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        nz, ny, nx = vol.shape
        self.lib.ref_reference_render(vol.ctypes.data_as(C.c_void_p), 1 if vol.dtype == np.uint8 else 0, nx, ny, nz,
                                      bmin, bmax, C.byref(c), W, H,
                                      ptr(np.ascontiguousarray(tf_rgba8, np.uint8), _u8), C.c_float(tf_step),
                                      C.byref(p), ptr(rgba, _f), ptr(rgba8, _u8))
        return rgba, rgba8

    def reconstruct(self, cam, dmin, dmax, W, H, m, data, tf_rgba8, tf_step, rp: RecParams, index=None,
                    bounds=None, prediction_coded=0, coding_warp=0, coding_params=None, want_rgba8=False):
        c = Camera.from_dict(cam)
        rgba = np.zeros((H, W, 4), np.float32)
        rgba8 = np.zeros((H, W, 4), np.uint8) if want_rgba8 else None
        bmin = (C.c_float * 3)(*dmin)
        bmax = (C.c_float * 3)(*dmax)
        self.lib.ref_reconstruct(C.byref(c), bmin, bmax, W, H, m, ptr(data, _f), C.c_long(len(data)),
                                 ptr(index, _u32), ptr(bounds, _u32), prediction_coded, coding_warp,
                                 ptr(coding_params, _u16), ptr(np.ascontiguousarray(tf_rgba8, np.uint8), _u8),
                                 C.c_float(tf_step), C.byref(rp), ptr(rgba, _f), ptr(rgba8, _u8))
        return rgba, rgba8
