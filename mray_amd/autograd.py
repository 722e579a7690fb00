"""torch.autograd wrappers over the HIP kernels (forward and hand-written
adjoint behind the C ABI).

ReconstructFunction: float RGBA of the moment-image march as a function of the
Lagrange multipliers and of the 256 x 4 transfer-function table. The backward
pass calls mray_hip_reconstruct_backward, which re-marches each ray
(per-ray checkpointing: memory stays O(pixels x moments)).

transfer_table: the reference's preintegrated table T[x] = (rgb^2.2 a s, a s)
from a float RGBA transfer function in [0, 1] (src/transfer_function.cpp:46-119),
in torch ops so that gradients reach the RGBA entries."""
from __future__ import annotations

import torch

from . import hip
from .types import RecParams


class ReconstructFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, lagrange, table, cam, dmin, dmax, W, H, m, rp: RecParams, index, bounds):
        rgba, _ = hip.reconstruct(cam, dmin, dmax, lagrange.detach().contiguous(), W, H, m,
                                  table.detach().contiguous(), rp, form="mixed", index=index, bounds=bounds)
        ctx.save_for_backward(lagrange, table, rgba)
        ctx.args = (cam, dmin, dmax, W, H, m, rp, index, bounds)
        return rgba

    @staticmethod
    def backward(ctx, grad_rgba):
        lagrange, table, rgba = ctx.saved_tensors
        cam, dmin, dmax, W, H, m, rp, index, bounds = ctx.args
        g_lag, g_tf = hip.reconstruct_backward(cam, dmin, dmax, lagrange.detach().contiguous(), W, H, m,
                                               table.detach().contiguous(), rp, rgba,
                                               grad_rgba.detach().to(torch.float32).contiguous(), index=index,
                                               bounds=bounds, want_lagrange=ctx.needs_input_grad[0],
                                               want_tf=ctx.needs_input_grad[1])
        return g_lag, g_tf, None, None, None, None, None, None, None, None, None


def reconstruct(lagrange, table, cam, dmin, dmax, W, H, m, rp: RecParams, index=None, bounds=None):
    """Differentiable float RGBA (H, W, 4)."""
    return ReconstructFunction.apply(lagrange, table, cam, dmin, dmax, W, H, m, rp, index, bounds)


def transfer_table(tf_rgba: torch.Tensor, step_size: float) -> torch.Tensor:
    """(256, 4) table (rgb^2.2 * a * s, a * s) of a float RGBA transfer function."""
    rgb = tf_rgba[:, :3].clamp(0.0, 1.0).pow(2.2)
    a = tf_rgba[:, 3:4].clamp(0.0, 1.0) * step_size
    return torch.cat([rgb * a, a], dim=1)


class PrepareFunction(torch.autograd.Function):
    """Lagrange multipliers of prepare (MIXED) as a function of the moments."""

    @staticmethod
    def forward(ctx, moments, W, H, m, bias, index):
        lag = hip.prepare(moments.detach().contiguous(), W, H, m, bias, form="mixed", index=index)
        ctx.save_for_backward(moments)
        ctx.args = (W, H, m, bias, index)
        return lag

    @staticmethod
    def backward(ctx, grad_lag):
        (moments,) = ctx.saved_tensors
        W, H, m, bias, index = ctx.args
        g = hip.prepare_backward(moments.detach().contiguous(), W, H, m, bias,
                                 grad_lag.detach().to(torch.float32).contiguous(), index=index)
        return g, None, None, None, None, None


class GenerateFunction(torch.autograd.Function):
    """Dense moments of generate (MIXED) as a function of the voxels of an
    F32 volume (the per-ray bounds are constants of the derivative)."""

    @staticmethod
    def forward(ctx, voxels, dmin, dmax, cam, gp, W, H, m):
        vol = hip.Volume(voxels.detach().contiguous(), dmin, dmax)
        # the per-ray normalisation is kept for the reverse pass
        rr = torch.empty((H * W, 2), dtype=torch.float32, device=voxels.device) if gp.transform_to_bounds else None
        mo, bounds, _ = hip.generate(vol, cam, gp, W, H, m, form="mixed", ray_range=rr)
        ctx.vol = vol
        ctx.ray_range = rr
        ctx.args = (cam, gp, W, H, m)
        ctx.mark_non_differentiable(bounds) if bounds is not None else None
        return mo, bounds

    @staticmethod
    def backward(ctx, grad_mo, grad_bounds):
        cam, gp, W, H, m = ctx.args
        g = hip.generate_backward(ctx.vol, cam, gp, W, H, m, grad_mo.detach().to(torch.float32).contiguous(),
                                  ray_range=ctx.ray_range)
        return g, None, None, None, None, None, None, None


def render(voxels, dmin, dmax, cam, gp, W, H, m, table, rp: RecParams):
    """Differentiable image of a volume: generate -> prepare -> reconstruct,
    with gradients to the voxels and to the transfer-function table."""
    mo, bounds = GenerateFunction.apply(voxels, dmin, dmax, cam, gp, W, H, m)
    lag = PrepareFunction.apply(mo, W, H, m, rp.bias, None)
    return reconstruct(lag, table, cam, dmin, dmax, W, H, m, rp, bounds=bounds)
