"""Matched-filter backprojection op (drop-in for geraf/ops/RFMatchedFilter/matched_filter_cuda.py).

``matched_filter_cuda(real, image, tpos, rpos, voxel, total_ant_size,
numADCSample, adcSampleRate, As_sci, N) -> (out_real[N], out_imag[N])`` with
the reference's argument meaning, shape assertions and autograd behaviour
(matched_filter_cuda.py:12-255):

* ``voxel`` is given as [N, 3] and handed to the kernel as [3, N];
* gradients flow only into ``real`` and ``image`` and only when both need
  them (matched_filter_cuda.py:175); positions and voxels get ``None``;
* the backward is ``once_differentiable``.

Forward and adjoint run on the gfx950 kernels of libgeraf_rf.so.
"""
from __future__ import annotations

from typing import Tuple

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _C


def _contig(t: torch.Tensor) -> torch.Tensor:
    return t if t.is_contiguous() else t.contiguous()


class _MatchedFilterKernel(Function):
    @staticmethod
    def forward(ctx, real, image, tpos, rpos, voxel, total_ant_size: int, numADCSample: int,
                adcSampleRate: float, As_sci: float, N: int):
        assert real.dim() == 2, (
            f"Expected 2D tensor (total_ant_size, numADCSample) for real part, but got: {real.dim()}D")
        assert image.dim() == 2, (
            f"Expected 2D tensor(total_ant_size, numADCSample) for image part, but got: {image.dim()}D")
        assert tpos.dim() == 2, (
            f"Expected 2D tensor (total_ant_size, 3) for position data, but got: {tpos.dim()}D")
        assert rpos.dim() == 2, (
            f"Expected 2D tensor (total_ant_size, 3) for position data, but got: {rpos.dim()}D")
        assert voxel.dim() == 2, (
            f"Expected 3D tensor (numPoints, 3) for voxel data, but got: {voxel.dim()}D")
        for t, name, rows, cols in ((real, "real", total_ant_size, numADCSample),
                                    (image, "image", total_ant_size, numADCSample),
                                    (tpos, "pos", total_ant_size, 3),
                                    (rpos, "pos", total_ant_size, 3),
                                    (voxel, "pos", N, 3)):
            assert t.size(0) == rows and t.size(1) == cols, (
                f"Size mismatch for {name} tensor: expected ({rows}, {cols}), "
                f"but got: ({t.size(0)}, {t.size(1)})")

        voxel_soa = _contig(voxel.transpose(0, 1))
        real, image, tpos, rpos = map(_contig, (real, image, tpos, rpos))
        out_real = torch.empty((N,), dtype=real.dtype, device=real.device)
        out_imag = torch.empty((N,), dtype=real.dtype, device=real.device)
        _C.MatchedFilter_forward(real, image, tpos, rpos, out_real, out_imag, voxel_soa,
                                 total_ant_size, numADCSample, adcSampleRate, As_sci, N)
        ctx.save_for_backward(real, image, tpos, rpos, voxel_soa)
        ctx.cfg = (total_ant_size, numADCSample, adcSampleRate, As_sci, N)
        return out_real, out_imag

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output_real, grad_output_image):
        real, image, tpos, rpos, voxel_soa = ctx.saved_tensors
        total_ant_size, numADCSample, adcSampleRate, As_sci, N = ctx.cfg
        assert grad_output_real.is_cuda and grad_output_image.is_cuda, \
            "grad_output_real and grad_output_image must be CUDA tensors."
        grad_real = grad_image = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            grad_real = torch.empty_like(real)
            grad_image = torch.empty_like(image)
            _C.MatchedFilter_backward_realimag(
                _contig(grad_output_real), _contig(grad_output_image), real, image, tpos, rpos,
                grad_real, grad_image, voxel_soa, total_ant_size, numADCSample, adcSampleRate,
                As_sci, N)
        return (grad_real, grad_image) + (None,) * 8


def matched_filter_cuda(real: torch.Tensor, image: torch.Tensor, tpos: torch.Tensor,
                        rpos: torch.Tensor, voxel: torch.Tensor, total_ant_size: int,
                        numADCSample: int, adcSampleRate: float, As_sci: float,
                        N: int) -> Tuple[torch.Tensor, torch.Tensor]:
    """Backproject echoes ``real + j*image`` [A, K] onto voxels [N, 3].

    out[v] = (1/A) * sum_a sum_k s[a,k] * exp(-j*2*pi*(773.704 + As_sci*k/fs)*tau)
    with tau = (|t_a - v| + |r_a - v| + 0.15) / 2.9979 (reference units).
    """
    assert real.is_cuda and image.is_cuda and tpos.is_cuda and rpos.is_cuda and voxel.is_cuda, \
        "All tensors must be on CUDA."
    return _MatchedFilterKernel.apply(real, image, tpos, rpos, voxel, total_ant_size,
                                      numADCSample, adcSampleRate, As_sci, N)
