"""Moment-image files through the C++ host library (libmray_host.so, C entry
points in include/mray_host.h): write_moment_image / load_moment_image of
src/moment_image_io.cpp, byte-compatible with the reference in every variant
(LZ4 sections, quantized and entropy-coded moments). Quantization and
entropy coding run on the GPU; files of float moments need none.

An image is a dict: W, H, m, dmin, dmax, data (float32), and optionally index
(compact images: W*H+1 uint32 offsets), bounds (W*H uint32, lower | upper <<
16), error_bounds (W*H uint16), coding_params (2(m-1) uint16),
prediction_coded, warp.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

from .hip import MrayError

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libmray_host.so"

_f3 = C.c_float * 3
_lib = None


def load_library() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise MrayError(f"{LIB_PATH} is missing: build it with `make -C mray_amd`")
        lib = C.CDLL(str(LIB_PATH))
        lib.mray_host_last_error.restype = C.c_char_p
        lib.mray_host_write_moment_image.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _f3, _f3,
                                                     C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p,
                                                     C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int]
        lib.mray_host_load_moment_image.argtypes = [C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                    C.c_void_p, C.c_void_p]
        lib.mray_host_quantization_table.argtypes = [C.c_void_p]
        lib.mray_host_uncertainty_tf.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_void_p]
        _lib = lib
    return _lib


def _check(rc, what):
    if rc != 0:
        raise MrayError(f"{what} failed ({rc}): {load_library().mray_host_last_error().decode()}")


def _arr(img, key, dtype):
    v = img.get(key)
    return None if v is None else np.ascontiguousarray(v, dtype)


def _p(a):
    return None if a is None else a.ctypes.data


def write_moment_image(path, img: dict, compress: bool = False, bits: int = 32, entropy_coding: bool = False,
                       table=None) -> None:
    """write_moment_image of img; table (m widths) first sets the
    prediction-coding quantization table, which stays set in this process."""
    data = np.ascontiguousarray(img["data"], np.float32)
    index, bounds = _arr(img, "index", np.uint32), _arr(img, "bounds", np.uint32)
    eb, cp = _arr(img, "error_bounds", np.uint16), _arr(img, "coding_params", np.uint16)
    tb = None if table is None else np.ascontiguousarray(table, np.uint8)
    _check(load_library().mray_host_write_moment_image(
        str(path).encode(), img["W"], img["H"], img["m"], int(img.get("prediction_coded", 0)), int(img.get("warp", 0)),
        _f3(*img["dmin"]), _f3(*img["dmax"]), _p(index), _p(data), data.size, _p(bounds), _p(eb), _p(cp), _p(tb),
        int(bool(compress)), int(bits), int(bool(entropy_coding))), "mray_host_write_moment_image")


def load_moment_image(path) -> dict:
    """load_moment_image: the image dict (index None for dense images)."""
    lib = load_library()
    hdr = np.zeros(10, np.int64)
    _check(lib.mray_host_load_moment_image(str(path).encode(), _p(hdr), None, None, None, None, None),
           "mray_host_load_moment_image")
    W, H, m = int(hdr[0]), int(hdr[1]), int(hdr[2])
    index = np.zeros(W * H + 1, np.uint32) if hdr[4] else None
    data = np.zeros(max(int(hdr[9]), 1), np.float32)
    bounds = np.zeros(W * H, np.uint32) if hdr[6] else None
    eb = np.zeros(W * H, np.uint16) if hdr[7] else None
    cp = np.zeros(max(2 * (m - 1), 1), np.uint16) if hdr[8] else None
    _check(lib.mray_host_load_moment_image(str(path).encode(), _p(hdr), _p(index), _p(data), _p(bounds), _p(eb),
                                           _p(cp)), "mray_host_load_moment_image")
    return dict(W=W, H=H, m=m, warp=int(hdr[3]), prediction_coded=bool(hdr[5]), index=index,
                data=data[:int(hdr[9])], bounds=bounds, error_bounds=eb,
                coding_params=None if cp is None else cp[:2 * (m - 1)])


def quantization_table():
    """The prediction-coding quantization table (255 widths) when set, else None."""
    t = np.zeros(255, np.uint8)
    return t if load_library().mray_host_quantization_table(_p(t)) else None
