"""Depth-resampler dispatch helper (drop-in for geraf/models/model_utils/ray_sampler.py:7-66)."""
from typing import Tuple

import torch


def bilinear_ray_sampler(normals: torch.Tensor, sigmas: torch.Tensor, z_vals: torch.Tensor,
                         tgt_z_vals: torch.Tensor,
                         backend: str = "cuda") -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """normals [R,S,3], sigmas [A,R,S], z_vals [R,S], tgt_z_vals [R,T] ->
    (normals [R,T,3], sigmas [A,R,T], inbounds [R,T] bool)."""
    N_ray, Nz, _ = normals.shape
    N_ant = sigmas.shape[0]
    Nz_tgt = tgt_z_vals.shape[1]
    if backend != "cuda":
        raise ValueError(f"Unsupported backend: {backend}. Use 'cuda' or 'torch'.")
    from geraf.ops.RFSignalTracer.bilinear_raysample_cuda import bilinear_raysample_cuda

    return bilinear_raysample_cuda(normals=normals, sigmas=sigmas, z_vals=z_vals,
                                   tgt_z_vals=tgt_z_vals, N_ant=N_ant, N_ray=N_ray, Nz=Nz,
                                   Nz_tgt=Nz_tgt)
