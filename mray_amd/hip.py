"""ctypes binding of libmray_hip.so (include/mray_hip.h) over torch device tensors.

torch is used for device memory and streams only; all compute runs in the HIP
kernels of mray_amd/csrc. There is no CPU fallback: a missing library raises.
Every call is issued on torch's current stream unless a stream is given.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np
import torch

from .types import Camera, ErrParams, GenParams, RecParams, RefParams, MRAY_DTYPE_F32, MRAY_DTYPE_U8, MRAY_FORM_MIXED, MRAY_FORM_PLAIN

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libmray_hip.so"

_vp = C.c_void_p
_f3 = C.c_float * 3

_SIGNATURES = {
    "mray_hip_init": ([C.c_int], C.c_int),
    "mray_hip_last_error": ([], C.c_char_p),
    "mray_hip_volume_create": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, _f3, _f3, C.POINTER(_vp)], C.c_int),
    "mray_hip_volume_wrap": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, _f3, _f3, C.POINTER(_vp)], C.c_int),
    "mray_hip_volume_destroy": ([_vp], C.c_int),
    "mray_hip_volume_from_particles": ([_vp, _vp, _vp, _vp, C.c_uint32, C.c_float, _f3, _f3, C.POINTER(_vp)], C.c_int),
    "mray_hip_volume_data": ([_vp], _vp),
    "mray_hip_generate": ([_vp, C.POINTER(Camera), C.POINTER(GenParams), C.c_int, C.c_int, C.c_int, C.c_int, _vp,
                           _vp, _vp, _vp, _vp], C.c_int),
    "mray_hip_prepare": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, _vp, _vp], C.c_int),
    "mray_hip_error_bounds": ([_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp], C.c_int),
    "mray_hip_reconstruct_uncertainty": ([_vp, _f3, _f3, _vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp,
                                          C.c_int, C.c_int, C.c_float, _vp, _vp], C.c_int),
    "mray_hip_prepare_f64": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp], C.c_int),
    "mray_hip_reconstruct_f64": ([C.POINTER(Camera), _f3, _f3, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                  C.POINTER(RecParams), _vp, _vp, _vp], C.c_int),
    "mray_hip_reconstruct": ([C.POINTER(Camera), _f3, _f3, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                              C.POINTER(RecParams), C.c_int, _vp, _vp, _vp], C.c_int),
    "mray_hip_reference": ([_vp, C.POINTER(Camera), _vp, C.POINTER(RefParams), C.c_int, C.c_int, _vp, _vp, _vp],
                           C.c_int),
    "mray_hip_reconstruct_backward": ([C.POINTER(Camera), _f3, _f3, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                       C.POINTER(RecParams), _vp, _vp, _vp, _vp, _vp], C.c_int),
    "mray_hip_generate_backward": ([_vp, C.POINTER(Camera), C.POINTER(GenParams), C.c_int, C.c_int, C.c_int, _vp, _vp,
                                    _vp, _vp], C.c_int),
    "mray_hip_prepare_backward": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp, _vp], C.c_int),
    "mray_hip_compute_num_moments": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp], C.c_int),
    "mray_hip_moment_index": ([_vp, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "mray_hip_moment_index_noalloc": ([_vp, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "mray_hip_clamp_num_moments": ([_vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp], C.c_int),
    "mray_hip_compact": ([_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "mray_hip_prediction_encode": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, _vp, _vp], C.c_int),
    "mray_hip_prepare_coded": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp], C.c_int),
    "mray_hip_quantize": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int, _vp, C.c_size_t,
                           C.POINTER(C.c_size_t), _vp], C.c_int),
    "mray_hip_quantization_table": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, C.c_int,
                                     C.POINTER(C.c_uint8), _vp], C.c_int),
    "mray_hip_dequantize": ([_vp, C.c_size_t, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), C.c_int, _vp,
                             _vp], C.c_int),
    "mray_hip_quantize_dense": ([_vp, C.c_size_t, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "mray_hip_dequantize_dense": ([_vp, C.c_size_t, C.c_int, C.c_int, _vp, _vp], C.c_int),
    "mray_hip_entropy_encode": ([_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8), _vp,
                                 C.c_size_t, C.POINTER(C.c_size_t), _vp], C.c_int),
    "mray_hip_entropy_decode": ([_vp, C.c_size_t, _vp, C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_uint8),
                                 _vp, _vp], C.c_int),
    "mray_hip_particles_to_grid": ([_vp, _vp, _vp, _vp, C.c_uint32, C.c_float, _f3, _f3, C.c_int, C.c_int, C.c_int,
                                    _vp, _vp], C.c_int),
    "mray_hip_synth_mixing_layer": ([_vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                     C.POINTER(C.c_double), C.c_uint64, _vp], C.c_int),
    "mray_hip_synth_marschner_lobb": ([_vp, C.c_int, _vp], C.c_int),
    "mray_hip_synth_noise_blobs": ([_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                    C.POINTER(C.c_float), C.c_int, C.c_uint64, _vp], C.c_int),
}

_lib = None
_initialised_device = None


class MrayError(RuntimeError):
    pass


def load_library() -> C.CDLL:
    """Loads libmray_hip.so and declares the C ABI; raises if it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise MrayError(f"{LIB_PATH} is missing: build it with `make -C mray_amd` (no CPU fallback exists)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (argtypes, restype) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = restype
        _lib = lib
    return _lib


def _check(rc: int, what: str):
    if rc != 0:
        msg = load_library().mray_hip_last_error().decode(errors="replace")
        raise MrayError(f"{what} failed ({rc}): {msg}")


def init(device: int = 0):
    global _initialised_device
    lib = load_library()
    if _initialised_device != device:
        torch.cuda.set_device(device)
        _check(lib.mray_hip_init(device), "mray_hip_init")
        _initialised_device = device


def _ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise MrayError("expected a device tensor")
    if not t.is_contiguous():
        raise MrayError("expected a contiguous tensor")
    if t.numel() == 0:
        # an empty compact image: torch reports a null data pointer for a
        # zero-length view; its storage (allocated non-empty by _like /
        # compact) is a valid address that no kernel dereferences
        return _vp(t.untyped_storage().data_ptr())
    return _vp(t.data_ptr())


def _like(t: torch.Tensor) -> torch.Tensor:
    """empty_like that never hands out a null pointer (empty compact images)."""
    return torch.empty(max(t.numel(), 1), dtype=t.dtype, device=t.device)[: t.numel()].view(t.shape)


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return _vp(s.cuda_stream)


def form_id(form) -> int:
    if isinstance(form, str):
        return {"plain": MRAY_FORM_PLAIN, "mixed": MRAY_FORM_MIXED}[form]
    return int(form)


class Volume:
    """Device-resident RegularGrid (F32 or U8, layout [nz][ny][nx])."""

    def __init__(self, voxels: torch.Tensor, dmin, dmax):
        init(voxels.device.index or 0)
        if voxels.dtype not in (torch.float32, torch.uint8) or voxels.dim() != 3:
            raise MrayError("voxels must be a 3-D float32 or uint8 tensor")
        self.voxels = voxels.contiguous()
        self.dmin = tuple(float(v) for v in dmin)
        self.dmax = tuple(float(v) for v in dmax)
        nz, ny, nx = self.voxels.shape
        dtype = MRAY_DTYPE_U8 if voxels.dtype == torch.uint8 else MRAY_DTYPE_F32
        h = _vp()
        _check(load_library().mray_hip_volume_wrap(_ptr(self.voxels), dtype, nx, ny, nz, _f3(*self.dmin),
                                                   _f3(*self.dmax), C.byref(h)), "mray_hip_volume_wrap")
        self.handle = h

    def close(self):
        if self.handle:
            load_library().mray_hip_volume_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ParticleVolume(Volume):
    """A particle set marched directly (mray_hip_volume_from_particles): every
    generate sample is the reference's Particles::sample of the particles
    (device arrays x, y, z, values) with smoothing length h, over the domain
    [dmin, dmax]."""

    def __init__(self, x: torch.Tensor, y: torch.Tensor, z: torch.Tensor, v: torch.Tensor, h: float, dmin, dmax):
        init(x.device.index or 0)
        self.voxels = None
        self._parts = [t.contiguous() for t in (x, y, z, v)]
        if len({t.numel() for t in self._parts}) != 1 or any(t.dtype != torch.float32 for t in self._parts):
            raise MrayError("particle arrays must be float32 and of equal length")
        self.dmin = tuple(float(c) for c in dmin)
        self.dmax = tuple(float(c) for c in dmax)
        hnd = _vp()
        _check(load_library().mray_hip_volume_from_particles(*[_ptr(t) for t in self._parts], self._parts[0].numel(),
                                                             C.c_float(h), _f3(*self.dmin), _f3(*self.dmax),
                                                             C.byref(hnd)), "mray_hip_volume_from_particles")
        self.handle = hnd


def generate(vol: Volume, cam: dict, gp: GenParams, W: int, H: int, m: int, form="mixed", moments=None,
             bounds=None, samples=None, want_samples=False, ray_range=None, stream=None):
    """Dense moments (H, W, m), packed bounds (H*W,) uint32 (or None), samples per ray (or None).
    ray_range: optional (H*W, 2) float32 tensor that receives each ray's normalisation (lo, hi)."""
    dev = vol.voxels.device if vol.voxels is not None else vol._parts[0].device
    if moments is None:
        moments = torch.empty((H, W, m), dtype=torch.float32, device=dev)
    if bounds is None and gp.transform_to_bounds:
        bounds = torch.empty(H * W, dtype=torch.int32, device=dev)
    if samples is None and want_samples:
        samples = torch.empty(H * W, dtype=torch.int32, device=dev)
    c = Camera.from_dict(cam)
    _check(load_library().mray_hip_generate(vol.handle, C.byref(c), C.byref(gp), form_id(form), W, H, m,
                                            _ptr(moments), _ptr(bounds), _ptr(samples), _ptr(ray_range),
                                            _stream(stream)),
           "mray_hip_generate")
    return moments, bounds, samples


def prepare(moments: torch.Tensor, W: int, H: int, m: int, bias: float, form="mixed", index=None, out=None,
            stream=None):
    """Lagrange multipliers, layout of the moments: float for "plain" and
    "mixed", FP64 for "mixed64" (mray_hip_prepare_f64)."""
    if form == "mixed64":
        if out is None:
            out = torch.empty(max(moments.numel(), 1), dtype=torch.float64,
                              device=moments.device)[: moments.numel()].view(moments.shape)
        _check(load_library().mray_hip_prepare_f64(_ptr(moments), _ptr(index), W, H, m, C.c_float(bias), _ptr(out),
                                                   _stream(stream)), "mray_hip_prepare_f64")
        return out
    if out is None:
        out = _like(moments)
    _check(load_library().mray_hip_prepare(_ptr(moments), _ptr(index), W, H, m, C.c_float(bias), form_id(form),
                                           _ptr(out), _stream(stream)), "mray_hip_prepare")
    return out


def reference(vol: Volume, cam: dict, tf: torch.Tensor, p: RefParams, W: int, H: int, rgba=None, steps=None,
              want_steps=False, stream=None):
    """The reference direct volume renderer (mray_hip_reference): float RGBA
    (H, W, 4) of the volume marched through the table, and optionally the
    compositing steps per ray."""
    dev = vol.voxels.device
    if rgba is None:
        rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    if steps is None and want_steps:
        steps = torch.empty(H * W, dtype=torch.int32, device=dev)
    c = Camera.from_dict(cam)
    tf = tf.to(device=dev, dtype=torch.float32).contiguous()
    if tf.numel() != 1024:
        raise MrayError("transfer function table must have 256 RGBA entries")
    _check(load_library().mray_hip_reference(vol.handle, C.byref(c), _ptr(tf), C.byref(p), W, H, _ptr(rgba),
                                             _ptr(steps), _stream(stream)), "mray_hip_reference")
    return rgba, steps


def reconstruct(cam: dict, dmin, dmax, lagrange: torch.Tensor, W: int, H: int, m: int, tf: torch.Tensor,
                rp: RecParams, form="mixed", index=None, bounds=None, rgba=None, steps=None, want_steps=False,
                stream=None):
    """Float RGBA (H, W, 4) and optionally the compositing steps per ray."""
    dev = lagrange.device
    if rgba is None:
        rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    if steps is None and want_steps:
        steps = torch.empty(H * W, dtype=torch.int32, device=dev)
    c = Camera.from_dict(cam)
    tf = tf.to(device=dev, dtype=torch.float32).contiguous()
    if tf.numel() != 1024:
        raise MrayError("transfer function table must have 256 RGBA entries")
    if lagrange.dtype == torch.float64:
        # FP64 multipliers of prepare(..., form="mixed64"): the mixed march
        _check(load_library().mray_hip_reconstruct_f64(C.byref(c), _f3(*dmin), _f3(*dmax), _ptr(lagrange),
                                                       _ptr(index), _ptr(bounds), W, H, m, _ptr(tf), C.byref(rp),
                                                       _ptr(rgba), _ptr(steps), _stream(stream)),
               "mray_hip_reconstruct_f64")
        return rgba, steps
    _check(load_library().mray_hip_reconstruct(C.byref(c), _f3(*dmin), _f3(*dmax), _ptr(lagrange), _ptr(index),
                                               _ptr(bounds), W, H, m, _ptr(tf), C.byref(rp), form_id(form),
                                               _ptr(rgba), _ptr(steps), _stream(stream)), "mray_hip_reconstruct")
    return rgba, steps


def error_bounds(vol: "Volume", cam: dict, lagrange: torch.Tensor, W: int, H: int, m: int, ep: ErrParams,
                 index=None, bounds=None, out=None, stream=None) -> torch.Tensor:
    """Per-pixel uint16 error bounds (as an int16 tensor of W*H) of the
    reconstruction from PLAIN multipliers against the volume (mray_hip_error_bounds)."""
    if out is None:
        out = torch.empty(W * H, dtype=torch.int16, device=lagrange.device)
    c = Camera.from_dict(cam)
    _check(load_library().mray_hip_error_bounds(vol.handle, C.byref(c), _ptr(lagrange), _ptr(index), _ptr(bounds), W,
                                                H, m, C.byref(ep), _ptr(out), _stream(stream)),
           "mray_hip_error_bounds")
    return out


def reconstruct_uncertainty(cam: dict, dmin, dmax, lagrange: torch.Tensor, error_bounds: torch.Tensor, W: int, H: int,
                            m: int, tf: torch.Tensor, rp: RecParams, mode: int, error_type: int = 1,
                            bound_interpolation: float = 0.5, index=None, bounds=None, rgba=None,
                            stream=None) -> torch.Tensor:
    """Float RGBA (H, W, 4) of the uncertainty renderings
    (mray_hip_reconstruct_uncertainty): mode 0 interpolation with the
    preintegrated table (256, 4), mode 1 convolution with a 2D table (256, 256, 4)."""
    dev = lagrange.device
    if rgba is None:
        rgba = torch.empty((H, W, 4), dtype=torch.float32, device=dev)
    tf = tf.to(device=dev, dtype=torch.float32).contiguous()
    if tf.numel() != (1024 if mode == 0 else 256 * 256 * 4):
        raise MrayError("mode 0 takes the 256-entry table, mode 1 the 256 x 256 table")
    c = Camera.from_dict(cam)
    _check(load_library().mray_hip_reconstruct_uncertainty(C.byref(c), _f3(*dmin), _f3(*dmax), _ptr(lagrange),
                                                           _ptr(index), _ptr(bounds), _ptr(error_bounds), W, H, m,
                                                           _ptr(tf), C.byref(rp), mode, error_type,
                                                           bound_interpolation, _ptr(rgba), _stream(stream)),
           "mray_hip_reconstruct_uncertainty")
    return rgba


def reconstruct_backward(cam: dict, dmin, dmax, lagrange: torch.Tensor, W: int, H: int, m: int, tf: torch.Tensor,
                         rp: RecParams, rgba: torch.Tensor, grad_rgba: torch.Tensor, index=None, bounds=None,
                         want_lagrange=True, want_tf=True, stream=None):
    """(dL/dlagrange or None, dL/dtable (256, 4) or None) for L = sum(grad_rgba * rgba)."""
    dev = lagrange.device
    g_lag = _like(lagrange) if want_lagrange else None
    g_tf = torch.zeros((256, 4), dtype=torch.float32, device=dev) if want_tf else None
    c = Camera.from_dict(cam)
    tf = tf.to(device=dev, dtype=torch.float32).contiguous()
    # contiguous copies are bound to names so they outlive the call
    rgba_c = rgba.contiguous()
    grad_c = grad_rgba.contiguous()
    _check(load_library().mray_hip_reconstruct_backward(C.byref(c), _f3(*dmin), _f3(*dmax), _ptr(lagrange),
                                                        _ptr(index), _ptr(bounds), W, H, m, _ptr(tf), C.byref(rp),
                                                        _ptr(rgba_c), _ptr(grad_c),
                                                        _ptr(g_lag), _ptr(g_tf), _stream(stream)),
           "mray_hip_reconstruct_backward")
    return g_lag, g_tf


def generate_backward(vol: Volume, cam: dict, gp: GenParams, W: int, H: int, m: int, grad_moments: torch.Tensor,
                      out=None, ray_range=None, stream=None) -> torch.Tensor:
    """dL/dvoxels (shape of the volume) for a given dL/dmoments of generate; ray_range is
    the forward's per-ray normalisation (generate(..., ray_range=...)) or None to recompute it."""
    if out is None:
        out = torch.zeros_like(vol.voxels, dtype=torch.float32)
    c = Camera.from_dict(cam)
    grad_c = grad_moments.contiguous()
    _check(load_library().mray_hip_generate_backward(vol.handle, C.byref(c), C.byref(gp), W, H, m,
                                                     _ptr(grad_c), _ptr(ray_range), _ptr(out),
                                                     _stream(stream)),
           "mray_hip_generate_backward")
    return out


def prepare_backward(moments: torch.Tensor, W: int, H: int, m: int, bias: float, grad_lagrange: torch.Tensor,
                     index=None, stream=None) -> torch.Tensor:
    """dL/dmoments for a given dL/dlagrange of prepare (MIXED mathematics)."""
    out = _like(moments)
    grad_c = grad_lagrange.contiguous()
    _check(load_library().mray_hip_prepare_backward(_ptr(moments), _ptr(index), W, H, m, C.c_float(bias),
                                                    _ptr(grad_c), _ptr(out), _stream(stream)),
           "mray_hip_prepare_backward")
    return out


def compute_num_moments(moments: torch.Tensor, W: int, H: int, m: int, error_threshold: float, bounds=None,
                        out=None, stream=None) -> torch.Tensor:
    """Moments to keep per pixel, (H*W,) int32 holding uint32 counts
    (compute_num_moments, src/moment_compaction.h:147-174)."""
    if out is None:
        out = torch.empty(H * W, dtype=torch.int32, device=moments.device)
    _check(load_library().mray_hip_compute_num_moments(_ptr(moments), _ptr(bounds), W, H, m,
                                                       C.c_float(error_threshold), _ptr(out), _stream(stream)),
           "mray_hip_compute_num_moments")
    return out


def moment_index(counts: torch.Tensor, W: int, H: int, out=None, stream=None) -> torch.Tensor:
    """Exclusive scan of the counts: (H*W+1,) int32 holding uint32 offsets, last = total."""
    if out is None:
        out = torch.empty(H * W + 1, dtype=torch.int32, device=counts.device)
    _check(load_library().mray_hip_moment_index(_ptr(counts), W, H, _ptr(out), _stream(stream)),
           "mray_hip_moment_index")
    return out


def moment_index_noalloc(counts: torch.Tensor, W: int, H: int, out: torch.Tensor, stream=None) -> torch.Tensor:
    """moment_index into a caller-provided (H*W+1,) int32 tensor, with no device
    allocation by the library (mray_hip_moment_index_noalloc); W*H up to 8192^2."""
    _check(load_library().mray_hip_moment_index_noalloc(_ptr(counts), W, H, _ptr(out), _stream(stream)),
           "mray_hip_moment_index_noalloc")
    return out


def compact(moments: torch.Tensor, index: torch.Tensor, W: int, H: int, m: int, stream=None) -> torch.Tensor:
    """Compact copy of a dense moment image (reads the total from index[-1])."""
    s = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        total = int(index[-1].to(torch.int64).item()) & 0xFFFFFFFF
    out = torch.empty(max(total, 1), dtype=torch.float32, device=moments.device)
    _check(load_library().mray_hip_compact(_ptr(moments), _ptr(index), W, H, m, _ptr(out), _stream(stream)),
           "mray_hip_compact")
    return out[:total]


def quantization_table(m: int, prediction_coded: bool, bits: int = 10, table=None):
    """Word widths per moment (uint8, m entries): the given table (the
    reference's loaded quantization_table), else the reference's default for
    prediction-coded images (16 bits for moment 0, then `bits`;
    get_prediction_coding_quantization_table, src/moment_quantization.cpp:186-203)
    or `bits` for every moment (quantize(mi, bits))."""
    t = (C.c_uint8 * m)(*([bits] * m))
    if table is not None:
        for i, b in enumerate(table[:m]):
            t[i] = int(b)
    elif prediction_coded:
        t[0] = 16
    return t


def _table(bits):
    return bits if isinstance(bits, C.Array) else (C.c_uint8 * len(bits))(*[int(b) for b in bits])


def quantize(data: torch.Tensor, index: torch.Tensor, W: int, H: int, prediction_coded: bool, bits,
             stream=None) -> torch.Tensor:
    """The quantized bit stream of a compact image (uint8 tensor): words of
    bits[l] bits per moment, least significant bit first (mray_hip_quantize)."""
    t = _table(bits)
    n = C.c_size_t(0)
    s = stream if stream is not None else torch.cuda.current_stream()
    with torch.cuda.stream(s):
        total = int(index[-1].to(torch.int64).item()) & 0xFFFFFFFF
    # every word is at most max(bits) wide: one call with that capacity
    cap = (total * max(t) + 7) // 8
    out = torch.empty(max(cap, 1), dtype=torch.uint8, device=index.device)
    _check(load_library().mray_hip_quantize(_ptr(data), _ptr(index), W, H, int(bool(prediction_coded)), t, len(t),
                                            _ptr(out), out.numel(), C.byref(n), _stream(stream)), "mray_hip_quantize")
    return out[:n.value]


def dequantize(stream_bytes: torch.Tensor, index: torch.Tensor, W: int, H: int, prediction_coded: bool, bits,
               stream=None) -> torch.Tensor:
    """The compact image a quantized stream holds (mray_hip_dequantize)."""
    t = _table(bits)
    total = int(index[-1].to(torch.int64).item()) & 0xFFFFFFFF
    out = torch.empty(max(total, 1), dtype=torch.float32, device=index.device)
    sb = stream_bytes if stream_bytes.numel() > 0 else None
    _check(load_library().mray_hip_dequantize(_ptr(sb), stream_bytes.numel(), _ptr(index), W, H,
                                              int(bool(prediction_coded)), t, len(t), _ptr(out), _stream(stream)),
           "mray_hip_dequantize")
    return out[:total]


def quantize_dense(data: torch.Tensor, m: int, bits: int, stream=None) -> torch.Tensor:
    """The byte-aligned quantized values of a dense image (mray_hip_quantize_dense):
    (bits + 7) // 8 little-endian bytes per value."""
    n = data.numel()
    out = torch.empty(max(n * ((bits + 7) // 8), 1), dtype=torch.uint8, device=data.device)
    _check(load_library().mray_hip_quantize_dense(_ptr(data), n, m, bits, _ptr(out), _stream(stream)),
           "mray_hip_quantize_dense")
    return out[:n * ((bits + 7) // 8)]


def dequantize_dense(qs: torch.Tensor, n: int, m: int, bits: int, stream=None) -> torch.Tensor:
    """The values byte-aligned quantized data holds (mray_hip_dequantize_dense)."""
    out = torch.empty(max(n, 1), dtype=torch.float32, device=qs.device)
    _check(load_library().mray_hip_dequantize_dense(_ptr(qs), n, m, bits, _ptr(out), _stream(stream)),
           "mray_hip_dequantize_dense")
    return out[:n]


def entropy_encode(data: torch.Tensor, index: torch.Tensor, W: int, H: int, m: int, prediction_coded: bool, bits,
                   stream=None):
    """The entropy-coded moment section of a compact image as a host numpy
    uint8 array (mray_hip_entropy_encode): moment-0 words, then one adaptive
    arithmetic code per moment of bits[l] bits."""
    t = _table(bits)
    if len(t) < m:
        raise ValueError("entropy_encode needs m widths")
    total = int(index[-1].to(torch.int64).item()) & 0xFFFFFFFF
    # at most 4 bytes per symbol, 2 per moment 0, and the headers
    cap = 4 + 2 * W * H + 20 * m + 4 * total + 64
    out = np.zeros(cap, np.uint8)
    n = C.c_size_t(0)
    _check(load_library().mray_hip_entropy_encode(_ptr(data), _ptr(index), W, H, m, int(bool(prediction_coded)), t,
                                                  out.ctypes.data, cap, C.byref(n), _stream(stream)),
           "mray_hip_entropy_encode")
    return out[:n.value].copy()


def entropy_decode(section, index: torch.Tensor, W: int, H: int, m: int, prediction_coded: bool, bits,
                   stream=None) -> torch.Tensor:
    """The compact image an entropy-coded section (host bytes) holds (mray_hip_entropy_decode)."""
    t = _table(bits)
    if len(t) < m:
        raise ValueError("entropy_decode needs m widths")
    sec = np.ascontiguousarray(section, np.uint8)
    total = int(index[-1].to(torch.int64).item()) & 0xFFFFFFFF
    out = torch.empty(max(total, 1), dtype=torch.float32, device=index.device)
    _check(load_library().mray_hip_entropy_decode(sec.ctypes.data, sec.size, _ptr(index), W, H, m,
                                                  int(bool(prediction_coded)), t, _ptr(out), _stream(stream)),
           "mray_hip_entropy_decode")
    return out[:total]


def curve_table(data: torch.Tensor, index: torch.Tensor, W: int, H: int, m: int, coding_params, start_bits: int,
                coding_warp: int = 0, stream=None):
    """The quantization table the reference's curve search picks for a
    prediction-coded compact image (mray_hip_quantization_table); m widths."""
    t = (C.c_uint8 * m)()
    _check(load_library().mray_hip_quantization_table(_ptr(data), _ptr(index), W, H, m, coding_warp,
                                                      _ptr(coding_params), start_bits, t, _stream(stream)),
           "mray_hip_quantization_table")
    return list(t)


def clamp_num_moments(counts: torch.Tensor, W: int, H: int, lo: int, hi: int, stream=None) -> torch.Tensor:
    """Counts of non-empty pixels clamped to [lo, hi] in place (mray_hip_clamp_num_moments)."""
    _check(load_library().mray_hip_clamp_num_moments(_ptr(counts), W, H, lo, hi, _stream(stream)),
           "mray_hip_clamp_num_moments")
    return counts


def compact_image(moments: torch.Tensor, W: int, H: int, m: int, error_threshold: float, bounds=None, stream=None,
                  count_range=None):
    """The compaction of generate with compact_image set: (compact data, index).
    count_range (lo, hi): the moment counts of non-empty pixels clamped to it."""
    counts = compute_num_moments(moments, W, H, m, error_threshold, bounds=bounds, stream=stream)
    if count_range is not None:
        clamp_num_moments(counts, W, H, count_range[0], count_range[1], stream=stream)
    index = moment_index(counts, W, H, stream=stream)
    return compact(moments, index, W, H, m, stream=stream), index


def prediction_encode(data: torch.Tensor, index: torch.Tensor, W: int, H: int, m: int, coding_warp: int = 0,
                      bias: float = 1e-4, stream=None):
    """In-place prediction coding of a compact image; returns the quantized
    coding parameters, (2*(m-1),) int16 holding uint16 values (None for the
    warp without parameters)."""
    params = None
    if coding_warp == 0 and m > 1:
        params = torch.zeros(2 * (m - 1), dtype=torch.int16, device=data.device)
    _check(load_library().mray_hip_prediction_encode(_ptr(data), _ptr(index), W, H, m, coding_warp, C.c_float(bias),
                                                     _ptr(params), _stream(stream)), "mray_hip_prediction_encode")
    return params


def prepare_coded(moments: torch.Tensor, W: int, H: int, m: int, coding_warp: int = 0, coding_params=None,
                  index=None, out=None, stream=None) -> torch.Tensor:
    """Lagrange multipliers of a prediction-coded image (same layout as the input)."""
    if out is None:
        out = _like(moments)
    _check(load_library().mray_hip_prepare_coded(_ptr(moments), _ptr(index), W, H, m, coding_warp,
                                                 _ptr(coding_params), _ptr(out), _stream(stream)),
           "mray_hip_prepare_coded")
    return out


def particles_to_grid(x: torch.Tensor, y: torch.Tensor, z: torch.Tensor, v: torch.Tensor, h: float, dmin, dmax,
                      nx: int, ny: int, nz: int, out=None, stream=None) -> torch.Tensor:
    """SPH particles sampled at the voxel positions of an (nz, ny, nx) grid."""
    if out is None:
        out = torch.empty((nz, ny, nx), dtype=torch.float32, device=x.device)
    _check(load_library().mray_hip_particles_to_grid(_ptr(x), _ptr(y), _ptr(z), _ptr(v), x.numel(), C.c_float(h),
                                                     _f3(*dmin), _f3(*dmax), nx, ny, nz, _ptr(out),
                                                     _stream(stream)), "mray_hip_particles_to_grid")
    return out


def synth_mixing_layer(n: int, seed: int = 1, device="cuda", stream=None) -> torch.Tensor:
    from .synth import mixing_layer_modes

    init(torch.device(device).index or 0)
    amp, theta, phi = mixing_layer_modes(seed)
    out = torch.empty((n, n, n), dtype=torch.float32, device=device)
    arr = C.c_double * 16
    _check(load_library().mray_hip_synth_mixing_layer(_ptr(out), n, arr(*amp), arr(*theta), arr(*phi),
                                                      C.c_uint64(seed), _stream(stream)),
           "mray_hip_synth_mixing_layer")
    return out


def synth_marschner_lobb(n: int, device="cuda", stream=None) -> torch.Tensor:
    init(torch.device(device).index or 0)
    out = torch.empty((n, n, n), dtype=torch.float32, device=device)
    _check(load_library().mray_hip_synth_marschner_lobb(_ptr(out), n, _stream(stream)),
           "mray_hip_synth_marschner_lobb")
    return out


def synth_noise_blobs(nx: int, ny: int, nz: int, seed: int = 2, device="cuda", stream=None) -> torch.Tensor:
    """The noise-and-blobs uint8 volume (nz, ny, nx) with the seeded blob
    parameters of synth.blob_params (normalised coordinates, so a smaller
    volume samples the same field)."""
    from .synth import blob_params

    init(torch.device(device).index or 0)
    centres, sigmas, amps = blob_params(seed)
    out = torch.empty((nz, ny, nx), dtype=torch.uint8, device=device)
    fp = C.POINTER(C.c_float)
    cc = np.ascontiguousarray(centres, np.float32)
    ss = np.ascontiguousarray(sigmas, np.float32)
    aa = np.ascontiguousarray(amps, np.float32)
    _check(load_library().mray_hip_synth_noise_blobs(_ptr(out), nx, ny, nz, cc.ctypes.data_as(fp),
                                                     ss.ctypes.data_as(fp), aa.ctypes.data_as(fp), len(aa),
                                                     C.c_uint64(seed), _stream(stream)),
           "mray_hip_synth_noise_blobs")
    return out
