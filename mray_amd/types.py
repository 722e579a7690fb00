"""ctypes mirrors of include/mray_types.h and the reference parameter defaults.

GenerationParameters (src/parameters.h:49-70) and ReconstructionParameters
(src/parameters.h:231-247) keep the reference's field names and defaults.
"""
from __future__ import annotations

import ctypes as C

MRAY_OK = 0
MRAY_DTYPE_F32 = 0
MRAY_DTYPE_U8 = 1
MRAY_FORM_PLAIN = 0
MRAY_FORM_MIXED = 1
CODING_WARP_DEFAULT_TRANSFORMED = 0
CODING_WARP_NONE = 1


class Camera(C.Structure):
    _fields_ = [
        ("orthographic", C.c_int32),
        ("transform", C.c_float * 16),
        ("position", C.c_float * 3),
        ("right", C.c_float * 3),
        ("up", C.c_float * 3),
        ("dir", C.c_float * 3),
        ("scale", C.c_float),
        ("aspect", C.c_float),
        ("znear", C.c_float),
    ]

    @classmethod
    def from_dict(cls, d: dict) -> "Camera":
        c = cls()
        c.orthographic = int(d["orthographic"])
        for i, v in enumerate(d["transform"]):
            c.transform[i] = float(v)
        for name in ("position", "right", "up", "dir"):
            arr = getattr(c, name)
            for i, v in enumerate(d[name]):
                arr[i] = float(v)
        c.scale = float(d["scale"])
        c.aspect = float(d["aspect"])
        c.znear = float(d["znear"])
        return c


class GenParams(C.Structure):
    _fields_ = [
        ("step_size", C.c_float),
        ("data_min", C.c_float),
        ("data_max", C.c_float),
        ("error_threshold", C.c_float),
        ("bias", C.c_float),
        ("compact_image", C.c_int32),
        ("prediction_coding", C.c_int32),
        ("coding_warp", C.c_int32),
        ("transform_to_bounds", C.c_int32),
        ("transform_bound_eps", C.c_float),
    ]

    def __init__(self, **kw):
        defaults = dict(step_size=1.0, data_min=0.0, data_max=1.0, error_threshold=1.0, bias=1e-4,
                        compact_image=1, prediction_coding=1, coding_warp=0, transform_to_bounds=0,
                        transform_bound_eps=0.01)
        defaults.update(kw)
        super().__init__(**defaults)


class ErrParams(C.Structure):
    """mray_error_params (ErrorReconstructionParameters without the file options)."""
    _fields_ = [
        ("error_type", C.c_int32),
        ("error_percentile", C.c_int32),
        ("bias", C.c_float),
        ("data_min", C.c_float),
        ("data_max", C.c_float),
        ("step_size", C.c_float),
    ]

    def __init__(self, **kw):
        defaults = dict(error_type=2, error_percentile=90, bias=1e-6, data_min=0.0, data_max=1.0, step_size=0.01)
        defaults.update(kw)
        super().__init__(**defaults)


class RecParams(C.Structure):
    _fields_ = [
        ("bias", C.c_float),
        ("medium_emission", C.c_float),
        ("medium_scattering", C.c_float),
        ("early_ray_termination", C.c_float),
        ("data_min", C.c_float),
        ("data_max", C.c_float),
        ("step_size", C.c_float),
        ("use_truncated_fourier", C.c_int32),
    ]

    def __init__(self, **kw):
        defaults = dict(bias=1e-6, medium_emission=0.1, medium_scattering=0.1, early_ray_termination=0.95,
                        data_min=0.0, data_max=1.0, step_size=0.01, use_truncated_fourier=0)
        defaults.update(kw)
        super().__init__(**defaults)


class RefParams(C.Structure):
    """ReferenceParameters of the reference direct volume renderer
    (src/parameters.h:339-350) without single scattering; mirrors
    mray_ref_params."""
    _fields_ = [
        ("step_size", C.c_float),
        ("medium_emission", C.c_float),
        ("medium_scattering", C.c_float),
        ("early_ray_termination", C.c_float),
        ("data_min", C.c_float),
        ("data_max", C.c_float),
        ("single_scattering", C.c_int32),
    ]

    def __init__(self, **kw):
        defaults = dict(step_size=1.0, medium_emission=0.1, medium_scattering=0.1, early_ray_termination=0.95,
                        data_min=0.0, data_max=1.0, single_scattering=0)
        defaults.update(kw)
        super().__init__(**defaults)
