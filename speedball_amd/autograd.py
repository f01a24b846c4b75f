"""torch.autograd wrapper of the differentiable probe update, for gradient checking.

`DiffProbeUpdate.apply(albedo, intensity, field, intensity0, params)` runs the
two-update differentiable solve on the field (sbgi_diff_forward) and returns the
second update's interior irradiance (P, 36, 3). backward() calls the hand-written
adjoint (sbgi_diff_adjoint). Light intensities enter as scales of the scene's light
records: scale = intensity / intensity0, so d/d intensity = d/d scale / intensity0.

The tensors cross the C ABI through host memory (the wrapper exists to check the
adjoint against finite differences and an FP64 gradient, not as a training path).
"""
from __future__ import annotations

import numpy as np
import torch


class DiffProbeUpdate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, albedo, intensity, field, intensity0, params):
        i0 = np.asarray(intensity0, np.float64)
        scale = (intensity.detach().cpu().double().numpy() / i0).astype(np.float32)
        irr = field.diff_forward(albedo.detach().cpu().float().numpy(), scale, **(params or {}))
        ctx.field = field
        ctx.i0 = i0
        ctx.meta = (albedo.dtype, albedo.device, intensity.dtype, intensity.device)
        return torch.from_numpy(irr).to(device=albedo.device, dtype=albedo.dtype)

    @staticmethod
    def backward(ctx, grad):
        ga, gl = ctx.field.diff_adjoint(grad.detach().cpu().float().numpy())
        adt, adev, idt, idev = ctx.meta
        g_alb = torch.from_numpy(ga).to(device=adev, dtype=adt)
        g_int = torch.from_numpy((gl / ctx.i0).astype(np.float32)).to(device=idev, dtype=idt)
        return g_alb, g_int, None, None, None


def diff_probe_update(albedo, intensity, field, intensity0, params=None):
    return DiffProbeUpdate.apply(albedo, intensity, field, intensity0, params)
