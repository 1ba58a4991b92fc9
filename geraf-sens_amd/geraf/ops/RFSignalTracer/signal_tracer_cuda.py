"""Differentiable [A,P] signal tracer (drop-in for
geraf/ops/RFSignalTracer/signal_tracer_cuda.py:16-188).

``signal_tracer_cuda(normals[P,3], sigmas[A,P], points[P,3], antenna_t[A,3],
antenna_r[A,3], inbounds[P], numADCSample, adcSampleRate, fc_start, As_sci,
total_ant_size, total_point_size, use_diff=True, nomralization=True)
-> (real[A,K], imag[A,K])`` with the reference's ValueError checks, gradients
into ``normals`` and ``sigmas`` only, and an ``once_differentiable`` backward.
``nomralization`` is accepted and stored but, as in the reference, unused.
"""
from __future__ import annotations

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _C


def _ensure_contiguous(t: torch.Tensor) -> torch.Tensor:
    return t if t.is_contiguous() else t.contiguous()


class _SignalTracerFunction(Function):
    @staticmethod
    def forward(ctx, normals, sigmas, points, antenna_t, antenna_r, inbounds, numADCSample: int,
                adcSampleRate: float, fc_start: float, As_sci: float, total_ant_size: int,
                total_point_size: int, use_diff: bool, nomralization: bool):
        if not use_diff:
            raise ValueError("The standalone signal tracer retains only the differentiable "
                             "CUDA path.")
        for t, name, shape in ((normals, "normals", "[num_points, 3]"),
                               (points, "points", "[num_points, 3]"),
                               (antenna_t, "antenna_t", "[num_antennas, 3]"),
                               (antenna_r, "antenna_r", "[num_antennas, 3]")):
            if t.dim() != 2 or t.size(1) != 3:
                raise ValueError(f"`{name}` must have shape {shape}, got {tuple(t.shape)}.")
        if sigmas.dim() != 2:
            raise ValueError(f"`sigmas` must have shape [num_antennas, num_points], got "
                             f"{tuple(sigmas.shape)}.")
        if inbounds.dim() != 1:
            raise ValueError(f"`inbounds` must have shape [num_points], got "
                             f"{tuple(inbounds.shape)}.")
        normals, sigmas, points, antenna_t, antenna_r, inbounds = map(
            _ensure_contiguous, (normals, sigmas, points, antenna_t, antenna_r, inbounds))
        out_real = torch.empty((total_ant_size, numADCSample), dtype=normals.dtype,
                               device=normals.device)
        out_imag = torch.empty_like(out_real)
        _C.SignalTracer_forward(normals, sigmas, points, antenna_t, antenna_r, inbounds,
                                out_real, out_imag, numADCSample, adcSampleRate, fc_start,
                                As_sci, total_ant_size, total_point_size)
        ctx.save_for_backward(normals, sigmas, points, antenna_t, antenna_r, inbounds)
        ctx.cfg = (numADCSample, adcSampleRate, fc_start, As_sci, total_ant_size,
                   total_point_size)
        ctx.nomralization = nomralization
        return out_real, out_imag

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output_real, grad_output_imag):
        normals, sigmas, points, antenna_t, antenna_r, inbounds = ctx.saved_tensors
        grad_normals = grad_sigmas = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            full_n = torch.empty_like(normals)
            full_s = torch.empty_like(sigmas)
            _C.SignalTracer_backward_normal_sigma(
                _ensure_contiguous(grad_output_real), _ensure_contiguous(grad_output_imag),
                normals, sigmas, points, antenna_t, antenna_r, inbounds, full_n, full_s,
                *ctx.cfg)
            if ctx.needs_input_grad[0]:
                grad_normals = full_n
            if ctx.needs_input_grad[1]:
                grad_sigmas = full_s
        return (grad_normals, grad_sigmas) + (None,) * 12


def signal_tracer_cuda(normals: torch.Tensor, sigmas: torch.Tensor, points: torch.Tensor,
                       antenna_t: torch.Tensor, antenna_r: torch.Tensor, inbounds: torch.Tensor,
                       numADCSample: int, adcSampleRate: float, fc_start: float, As_sci: float,
                       total_ant_size: int, total_point_size: int, use_diff: bool = True,
                       nomralization: bool = True):
    if not all(t.is_cuda for t in (normals, sigmas, points, antenna_t, antenna_r, inbounds)):
        raise ValueError("All signal tracer inputs must be CUDA tensors.")
    return _SignalTracerFunction.apply(normals, sigmas, points, antenna_t, antenna_r, inbounds,
                                       numADCSample, adcSampleRate, fc_start, As_sci,
                                       total_ant_size, total_point_size, use_diff, nomralization)
