"""Linear depth resampling of per-ray normals and sigmas (drop-in for
geraf/ops/RFSignalTracer/bilinear_raysample_cuda.py:12-145).

``bilinear_raysample_cuda(normals[R,S,3], sigmas[A,R,S], z_vals[R,S],
tgt_z_vals[R,T], N_ant, N_ray, Nz, Nz_tgt) -> (normals[R,T,3],
sigmas[A,R,T], inbounds[R,T] bool)``.  The backward is an atomic-free gather,
so its result is deterministic.
"""
from __future__ import annotations

from typing import Tuple

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _C


class BilinearRaysampleKernel(Function):
    @staticmethod
    def forward(ctx, normals, sigmas, z_vals, tgt_z_vals, N_ant: int, N_ray: int, Nz: int,
                Nz_tgt: int):
        normals, sigmas = normals.contiguous(), sigmas.contiguous()
        z_vals, tgt_z_vals = z_vals.contiguous(), tgt_z_vals.contiguous()
        out_n = torch.empty((N_ray, Nz_tgt, 3), dtype=normals.dtype, device=normals.device)
        out_s = torch.empty((N_ant, N_ray, Nz_tgt), dtype=sigmas.dtype, device=sigmas.device)
        out_inb = torch.empty((N_ray, Nz_tgt), dtype=torch.bool, device=sigmas.device)
        _C.BilinearRaysample_forward(normals, sigmas, z_vals, tgt_z_vals, out_n, out_s, out_inb,
                                     N_ant, N_ray, Nz, Nz_tgt)
        ctx.save_for_backward(z_vals, tgt_z_vals)
        ctx.dims = (N_ant, N_ray, Nz, Nz_tgt)
        ctx.mark_non_differentiable(out_inb)
        return out_n, out_s, out_inb

    @staticmethod
    @once_differentiable
    def backward(ctx, grad_output_normals, grad_output_sigmas, _grad_inbounds):
        z_vals, tgt_z_vals = ctx.saved_tensors
        N_ant, N_ray, Nz, Nz_tgt = ctx.dims
        g_n = grad_output_normals.contiguous()
        g_s = grad_output_sigmas.contiguous()
        normals_grad = torch.empty((N_ray, Nz, 3), dtype=g_n.dtype, device=g_n.device)
        sigmas_grad = torch.empty((N_ant, N_ray, Nz), dtype=g_s.dtype, device=g_s.device)
        _C.BilinearRaysample_backward(g_n, g_s, z_vals, tgt_z_vals, normals_grad, sigmas_grad,
                                      N_ant, N_ray, Nz, Nz_tgt)
        return normals_grad, sigmas_grad, None, None, None, None, None, None


def bilinear_raysample_cuda(normals: torch.Tensor, sigmas: torch.Tensor, z_vals: torch.Tensor,
                            tgt_z_vals: torch.Tensor, N_ant: int, N_ray: int, Nz: int,
                            Nz_tgt: int) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Interpolate normals [N_ray, Nz, 3] and sigmas [N_ant, N_ray, Nz] from
    z_vals [N_ray, Nz] to tgt_z_vals [N_ray, Nz_tgt]; targets outside a ray's
    [z_first, z_last] give zeros and ``inbounds = False``."""
    return BilinearRaysampleKernel.apply(normals, sigmas, z_vals, tgt_z_vals, N_ant, N_ray, Nz,
                                         Nz_tgt)
