"""Fused lensless RF renderer: per-ray compositing + FMCW signal in one op.

This is the per-ray form of the reference path that runs, for each ray,
SDF -> NeuS alpha -> transmittance -> calibrated transmittance -> sigma
(rf_rendering.py:46-52,145-169,749-759; sdf_network.py:475-481) and then the
signal tracer with the ray's own antenna as transmitter and receiver
(SignalTracer_cuda_kernel.cu:351-450 with A = 1, P = the ray's samples).  The
torch glue of the reference keeps ~15 intermediates per sample for autograd;
here the backward is a hand-written adjoint kernel that recomputes them.

Inputs (R rays, S samples; float32 CUDA tensors):
  sdf [R,S], gradients [R,S,3] (raw SDF gradients: used for iter_cos and,
  normalised, as the reflecting normal), dists [R,S] (sample spacing in the
  SDF's frame), z_vals [R,S] (depth along the ray in metres), ray_o [R,3]
  (antenna positions, metres), ray_d [R,3] (unit directions), inv_s and power
  (scalar tensors), reflectance [R,S] (optional, default 1), trans_offset
  [R,S] (optional calibration new_cdf - prev_cdf, no gradient), n_valid [R]
  int (optional valid-sample count per ray; samples past it are absent).
Output: (real [R,K], imag [R,K]), divided by each ray's valid-sample count.
Gradients: sdf, gradients, reflectance, inv_s, power.
"""
from __future__ import annotations

from typing import Optional, Tuple

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _C
from .. import _native


def _c(t: Optional[torch.Tensor], dtype=torch.float32) -> Optional[torch.Tensor]:
    if t is None:
        return None
    t = t.detach()
    if t.dtype != dtype:
        t = t.to(dtype)
    return t if t.is_contiguous() else t.contiguous()


def _scalar(t, device) -> torch.Tensor:
    if not torch.is_tensor(t):
        t = torch.tensor(float(t), device=device)
    return _c(t.reshape(1))


def _kernel_inputs(sdf, gradients, reflectance, inv_s, power, dists, z_vals, ray_o, ray_d,
                   trans_offset, n_valid):
    """The kernels' operands: detached, float32 (int32 counts), contiguous."""
    dev = sdf.device
    return dict(sdf=_c(sdf), grad=_c(gradients), refl=_c(reflectance), dists=_c(dists),
                z=_c(z_vals), o=_c(ray_o), d=_c(ray_d), off=_c(trans_offset),
                nv=_c(n_valid, torch.int32), inv_s=_scalar(inv_s, dev),
                power=_scalar(power, dev))


class _LenslessTracerFunction(Function):
    @staticmethod
    def forward(ctx, sdf, gradients, reflectance, inv_s, power, dists, z_vals, ray_o, ray_d,
                trans_offset, n_valid, K: int, adcSampleRate: float, fc_start: float,
                As_sci: float, cos_anneal_ratio: float):
        R, S = sdf.shape
        ins = _kernel_inputs(sdf, gradients, reflectance, inv_s, power, dists, z_vals, ray_o,
                             ray_d, trans_offset, n_valid)
        out_r = torch.empty((R, K), dtype=torch.float32, device=sdf.device)
        out_i = torch.empty_like(out_r)
        # transmittance at each 64-sample tile start: the backward then walks
        # the samples once instead of twice
        carry = torch.empty((R, (S + 63) // 64), dtype=torch.float32, device=sdf.device)
        prm = _C.tracer_params(fc_start, As_sci, adcSampleRate)
        with torch.cuda.device(sdf.device):
            _native.call("geraf_lensless_forward_f32",
                         *map(_native.ptr, (ins["sdf"], ins["grad"], ins["refl"], ins["dists"],
                                            ins["z"], ins["o"], ins["d"], ins["off"], ins["nv"],
                                            ins["inv_s"], ins["power"], out_r, out_i, None,
                                            carry)),
                         R, S, K, float(cos_anneal_ratio), prm, _native.stream_of(out_r))
        # the inputs themselves are saved (autograd then refuses a backward
        # after any of them was modified in place); the backward converts
        # them again as the forward did
        ctx.save_for_backward(sdf, gradients, reflectance, inv_s, power, dists, z_vals, ray_o,
                              ray_d, trans_offset, n_valid)
        ctx.cfg = (R, S, K, float(cos_anneal_ratio), prm)
        ctx.carry = carry
        return out_r, out_i

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_real, grad_imag):
        saved = ctx.saved_tensors
        ins = _kernel_inputs(*saved)
        R, S, K, anneal, prm = ctx.cfg
        inv_s, power, reflectance = saved[3], saved[4], saved[2]
        dev = ins["sdf"].device
        g_r = _c(grad_real)
        g_i = _c(grad_imag)
        g_sdf = torch.empty((R, S), dtype=torch.float32, device=dev)
        g_grad = torch.empty((R, S, 3), dtype=torch.float32, device=dev)
        g_refl = (torch.empty((R, S), dtype=torch.float32, device=dev)
                  if reflectance is not None else None)
        g_inv = torch.empty(R, dtype=torch.float32, device=dev)
        g_pow = torch.empty(R, dtype=torch.float32, device=dev)
        with torch.cuda.device(dev):
            _native.call("geraf_lensless_backward_f32",
                         *map(_native.ptr, (g_r, g_i, ins["sdf"], ins["grad"], ins["refl"],
                                            ins["dists"], ins["z"], ins["o"], ins["d"],
                                            ins["off"], ins["nv"], ins["inv_s"], ins["power"],
                                            g_sdf, g_grad, g_refl, g_inv, g_pow, ctx.carry)),
                         R, S, K, anneal, prm, _native.stream_of(g_sdf))
        ctx.carry = None
        need = ctx.needs_input_grad
        g_inv_s = g_inv.sum().reshape(inv_s.shape) if need[3] else None
        g_power = g_pow.sum().reshape(power.shape) if need[4] else None
        return (g_sdf if need[0] else None, g_grad if need[1] else None,
                g_refl if (need[2] and g_refl is not None) else None, g_inv_s, g_power) + \
            (None,) * 11


def lensless_signal_tracer(sdf: torch.Tensor, gradients: torch.Tensor, dists: torch.Tensor,
                           z_vals: torch.Tensor, ray_o: torch.Tensor, ray_d: torch.Tensor,
                           inv_s, power, numADCSample: int, adcSampleRate: float,
                           fc_start: float, As_sci: float,
                           reflectance: Optional[torch.Tensor] = None,
                           trans_offset: Optional[torch.Tensor] = None,
                           n_valid: Optional[torch.Tensor] = None,
                           cos_anneal_ratio: float = 1.0) -> Tuple[torch.Tensor, torch.Tensor]:
    """Render the per-ray FMCW signal [R, numADCSample] (real, imag)."""
    tensors = [sdf, gradients, dists, z_vals, ray_o, ray_d]
    if not all(t.is_cuda for t in tensors):
        raise ValueError("All lensless tracer inputs must be CUDA tensors.")
    if sdf.dim() != 2:
        raise ValueError(f"`sdf` must have shape [R, S], got {tuple(sdf.shape)}.")
    R, S = sdf.shape
    for t, name, shape in ((gradients, "gradients", (R, S, 3)), (dists, "dists", (R, S)),
                           (z_vals, "z_vals", (R, S)), (ray_o, "ray_o", (R, 3)),
                           (ray_d, "ray_d", (R, 3))):
        if tuple(t.shape) != shape:
            raise ValueError(f"`{name}` must have shape {list(shape)}, got {tuple(t.shape)}.")
    for t, name in ((reflectance, "reflectance"), (trans_offset, "trans_offset")):
        if t is not None and tuple(t.shape) != (R, S):
            raise ValueError(f"`{name}` must have shape [R, S], got {tuple(t.shape)}.")
    if n_valid is not None and tuple(n_valid.shape) != (R,):
        raise ValueError(f"`n_valid` must have shape [R], got {tuple(n_valid.shape)}.")
    dev = sdf.device
    inv_s_t = inv_s if torch.is_tensor(inv_s) else torch.tensor(float(inv_s), device=dev)
    power_t = power if torch.is_tensor(power) else torch.tensor(float(power), device=dev)
    if inv_s_t.numel() != 1 or power_t.numel() != 1:
        raise ValueError("`inv_s` and `power` must be scalars.")
    return _LenslessTracerFunction.apply(sdf, gradients, reflectance, inv_s_t, power_t, dists,
                                         z_vals, ray_o, ray_d, trans_offset, n_valid,
                                         int(numADCSample), adcSampleRate, fc_start, As_sci,
                                         cos_anneal_ratio)

