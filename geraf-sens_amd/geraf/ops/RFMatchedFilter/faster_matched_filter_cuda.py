"""Selected/unselected split matched filter (drop-in for
geraf/ops/RFMatchedFilter/faster_matched_filter_cuda.py:12-319).

The forward backprojects the concatenation of the selected and unselected
antenna banks; the backward computes the adjoint only for the ``num_sel``
selected rows and divides by the total antenna count, which is what the
reference obtains by dividing by ``num_sel`` and rescaling by
``num_sel / total_ant_size`` (faster_matched_filter_cuda.py:229-238).  Here the
adjoint kernel takes the divisor directly, so the unselected rows cost
nothing and no rescale pass runs.
"""
from __future__ import annotations

from typing import Tuple

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _C
from .. import _native


def _contig(t: torch.Tensor) -> torch.Tensor:
    return t if t.is_contiguous() else t.contiguous()


class _FasterMatchedFilterKernel(Function):
    @staticmethod
    def forward(ctx, real_sel, image_sel, real_unsel, image_unsel, tpos_sel, rpos_sel,
                tpos_unsel, rpos_unsel, voxel, total_ant_size: int, num_sel: int,
                numADCSample: int, adcSampleRate: float, As_sci: float, N: int):
        num_unsel = total_ant_size - num_sel
        for t, name in ((real_sel, "real"), (image_sel, "image"), (real_unsel, "real"),
                        (image_unsel, "image"), (tpos_sel, "position"), (rpos_sel, "position"),
                        (tpos_unsel, "position"), (rpos_unsel, "position"), (voxel, "voxel")):
            assert t.dim() == 2, f"Expected 2D tensor for {name} data, but got: {t.dim()}D"
        for t, name, rows, cols in ((real_sel, "real", num_sel, numADCSample),
                                    (image_sel, "image", num_sel, numADCSample),
                                    (real_unsel, "real", num_unsel, numADCSample),
                                    (image_unsel, "image", num_unsel, numADCSample),
                                    (tpos_sel, "pos", num_sel, 3), (rpos_sel, "pos", num_sel, 3),
                                    (tpos_unsel, "pos", num_unsel, 3),
                                    (rpos_unsel, "pos", num_unsel, 3), (voxel, "pos", N, 3)):
            assert t.size(0) == rows and t.size(1) == cols, (
                f"Size mismatch for {name} tensor: expected ({rows}, {cols}), "
                f"but got: ({t.size(0)}, {t.size(1)})")

        voxel_soa = _contig(voxel.transpose(0, 1))
        real = torch.cat([real_sel, real_unsel], dim=0)
        image = torch.cat([image_sel, image_unsel], dim=0)
        tpos = torch.cat([tpos_sel, tpos_unsel], dim=0)
        rpos = torch.cat([rpos_sel, rpos_unsel], dim=0)
        out_real = torch.empty((N,), dtype=real_sel.dtype, device=real_sel.device)
        out_imag = torch.empty((N,), dtype=real_sel.dtype, device=real_sel.device)
        _C.MatchedFilter_forward(real, image, tpos, rpos, out_real, out_imag, voxel_soa,
                                 total_ant_size, numADCSample, adcSampleRate, As_sci, N)
        ctx.save_for_backward(_contig(tpos_sel), _contig(rpos_sel), voxel_soa)
        ctx.sel_shape = tuple(real_sel.shape)
        ctx.cfg = (total_ant_size, num_sel, numADCSample, adcSampleRate, As_sci, N)
        return out_real, out_imag

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output_real, grad_output_image):
        tpos_sel, rpos_sel, voxel_soa = ctx.saved_tensors
        total_ant_size, num_sel, numADCSample, adcSampleRate, As_sci, N = ctx.cfg
        assert grad_output_real.is_cuda and grad_output_image.is_cuda, \
            "grad_output_real and grad_output_image must be CUDA tensors."
        grad_real = grad_image = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            grad_real = torch.empty(ctx.sel_shape, dtype=grad_output_real.dtype,
                                    device=grad_output_real.device)
            grad_image = torch.empty_like(grad_real)
            backward_rows(_contig(grad_output_real), _contig(grad_output_image), tpos_sel,
                          rpos_sel, grad_real, grad_image, voxel_soa, num_sel, numADCSample,
                          N, total_ant_size, adcSampleRate, As_sci)
        return (grad_real, grad_image) + (None,) * 13


def backward_rows(grad_out_r, grad_out_i, tpos, rpos, grad_r, grad_i, voxel_soa, A, K, N,
                  norm_count, adcSampleRate, As_sci) -> None:
    """Adjoint for A antenna rows, divided by ``norm_count``."""
    params = _C._params(adcSampleRate, As_sci)
    dtype = grad_r.dtype
    if dtype == torch.float16:
        g_r = torch.empty(grad_r.shape, dtype=torch.float32, device=grad_r.device)
        g_i = torch.empty_like(g_r)
        backward_rows(grad_out_r.float(), grad_out_i.float(), tpos.float(), rpos.float(), g_r,
                      g_i, voxel_soa.float(), A, K, N, norm_count, adcSampleRate, As_sci)
        grad_r.copy_(g_r)
        grad_i.copy_(g_i)
        return
    name = "geraf_mf_backward_f32" if dtype == torch.float32 else "geraf_mf_backward_f64"
    with torch.cuda.device(grad_r.device):
        _native.call(name, *map(_native.ptr, (grad_out_r, grad_out_i, tpos, rpos, grad_r, grad_i,
                                             voxel_soa)),
                     int(A), int(K), int(N), int(norm_count), params, _native.stream_of(grad_r))


def faster_matched_filter_cuda(real_sel, image_sel, real_unsel, image_unsel, tpos_sel, rpos_sel,
                               tpos_unsel, rpos_unsel, voxel, total_ant_size: int, num_sel: int,
                               numADCSample: int, adcSampleRate: float, As_sci: float,
                               N: int) -> Tuple[torch.Tensor, torch.Tensor]:
    assert real_sel.is_cuda and image_sel.is_cuda and tpos_sel.is_cuda and rpos_sel.is_cuda \
        and voxel.is_cuda, "All tensors must be on CUDA."
    assert real_unsel.is_cuda and image_unsel.is_cuda and tpos_unsel.is_cuda \
        and rpos_unsel.is_cuda, "All tensors must be on CUDA."
    return _FasterMatchedFilterKernel.apply(real_sel, image_sel, real_unsel, image_unsel,
                                            tpos_sel, rpos_sel, tpos_unsel, rpos_unsel, voxel,
                                            total_ant_size, num_sel, numADCSample,
                                            adcSampleRate, As_sci, N)
