"""Tracer dispatch helpers (drop-in for geraf/models/model_utils/ray_tracer.py:7-58).

``lensless_ray_tracer`` unpacks ``radar_cfg`` into the [A,P] signal tracer as
the reference does; ``lensless_volume_tracer`` is the fused per-ray renderer
(compositing + signal in one op).  Only ``backend="cuda"`` exists (on ROCm
torch, HIP tensors report ``is_cuda``).
"""
from typing import Dict, Optional

import torch


def lensless_ray_tracer(normals: torch.Tensor, sigmas: torch.Tensor, points: torch.Tensor,
                        antenna_t: torch.Tensor, antenna_r: torch.Tensor,
                        radar_cfg: Dict[str, float], backend: str = "cuda",
                        inbounds: Optional[torch.Tensor] = None, use_diff: bool = True,
                        nomralization: bool = True):
    """Trace points [P,3] with normals [P,3] and per-antenna sigmas [A,P] to the
    antenna pairs (antenna_t, antenna_r) [A,3]; returns (real, imag) [A, numADCSample]."""
    num_points = normals.shape[0]
    if inbounds is None:
        inbounds = torch.ones(num_points, dtype=torch.bool, device=points.device)
    if backend != "cuda":
        raise ValueError(f"Unsupported ray tracer backend: {backend!r}. Only 'cuda' is retained.")
    from geraf.ops.RFSignalTracer.signal_tracer_cuda import signal_tracer_cuda

    return signal_tracer_cuda(
        normals=normals, sigmas=sigmas, points=points, antenna_t=antenna_t, antenna_r=antenna_r,
        inbounds=inbounds, numADCSample=radar_cfg["numADCSample"],
        adcSampleRate=radar_cfg["adcSampleRate"], fc_start=radar_cfg["fc_start"],
        As_sci=radar_cfg["As_sci"], total_ant_size=antenna_t.shape[0],
        total_point_size=num_points, use_diff=use_diff, nomralization=nomralization)


def lensless_volume_tracer(sdf: torch.Tensor, gradients: torch.Tensor, dists: torch.Tensor,
                           z_vals: torch.Tensor, ray_o: torch.Tensor, ray_d: torch.Tensor,
                           inv_s, power, radar_cfg: Dict[str, float],
                           reflectance: Optional[torch.Tensor] = None,
                           trans_offset: Optional[torch.Tensor] = None,
                           n_valid: Optional[torch.Tensor] = None, cos_anneal_ratio: float = 1.0,
                           backend: str = "cuda"):
    """Fused per-ray renderer: SDF samples [R,S] -> FMCW signal [R, numADCSample]."""
    if backend != "cuda":
        raise ValueError(f"Unsupported ray tracer backend: {backend!r}. Only 'cuda' is retained.")
    from geraf.ops.RFSignalTracer.lensless_tracer import lensless_signal_tracer

    return lensless_signal_tracer(
        sdf, gradients, dists, z_vals, ray_o, ray_d, inv_s, power,
        numADCSample=radar_cfg["numADCSample"], adcSampleRate=radar_cfg["adcSampleRate"],
        fc_start=radar_cfg["fc_start"], As_sci=radar_cfg["As_sci"], reflectance=reflectance,
        trans_offset=trans_offset, n_valid=n_valid, cos_anneal_ratio=cos_anneal_ratio)
