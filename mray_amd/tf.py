"""Transfer-function tables.

read_tf + preintegrate of the reference (src/transfer_function.cpp:46-119):
each of the 256 RGBA8 entries becomes (r*a*s, g*a*s, b*a*s, a*s) with
r,g,b = (byte/255)^2.2, a = byte/255 and s = step_size *
preintegration_step_size_factor (src/main.cpp:92-105). The reference stores
this row-constant 256x256 table; the kernels take the 256 distinct rows.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import pathlib

import numpy as np

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.powf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]


def _powf_table() -> np.ndarray:
    """C powf(i / 255.f, 2.2f) for the 256 byte values (std::pow on floats)."""
    return np.float32([_libm.powf(float(np.float32(i) / np.float32(255.0)), float(np.float32(2.2)))
                       for i in range(256)])


_GAMMA = _powf_table()


def table_from_rgba8(rgba8: np.ndarray, step_size: float) -> np.ndarray:
    """(256, 4) float32 table, float32 arithmetic in the reference's order."""
    b = np.asarray(rgba8, dtype=np.uint8).reshape(256, 4)
    r, g, bb = _GAMMA[b[:, 0]], _GAMMA[b[:, 1]], _GAMMA[b[:, 2]]
    a = (b[:, 3].astype(np.float32) / np.float32(255.0)) * np.float32(step_size)
    a = a.astype(np.float32)
    return np.ascontiguousarray(np.stack([r * a, g * a, bb * a, a], axis=1).astype(np.float32))


def read_tf_file(path) -> np.ndarray:
    """256 x RGBA8 bytes of a .bin transfer function (src/transfer_function.cpp:99-119)."""
    data = pathlib.Path(path).read_bytes()
    if len(data) < 1024:
        raise ValueError(f"{path}: transfer function needs 1024 bytes, got {len(data)}")
    return np.frombuffer(data[:1024], dtype=np.uint8).reshape(256, 4).copy()


def uncertainty_table(rgba8: np.ndarray, step_size: float, error_type: int) -> np.ndarray:
    """The 256 x 256 x RGBA uncertainty transfer function of the C++ host
    library (mray_host_uncertainty_tf, read_from_disk_uncertainty of
    src/transfer_function.cpp:233-248): box filter for the maximum error and
    the percentile (error_type 0, 2), gaussian filter for the RMSE (1)."""
    from . import moment_file

    out = np.zeros((256, 256, 4), np.float32)
    rgba8 = np.ascontiguousarray(rgba8, np.uint8)
    if rgba8.size != 1024:
        raise ValueError("rgba8 must hold 256 RGBA entries")
    moment_file._check(moment_file.load_library().mray_host_uncertainty_tf(rgba8.ctypes.data, float(step_size),
                                                                          int(error_type), out.ctypes.data),
                       "mray_host_uncertainty_tf")
    return out
