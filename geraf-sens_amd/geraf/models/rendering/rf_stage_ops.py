"""The per-step operations of GeRaFStage1/2.loss that sit on the RF path,
besides the networks and the two RF ops themselves.

* ``intersect_sphere`` and ``transmission_offset``: the transmittance
  calibration (rf_rendering.py:455-495,611-626,749-750).  The sphere hits and
  the offset are HIP kernels; the SDF evaluations between them are the
  caller's network calls.  ``transmission_offset`` gives the fused lensless
  renderer's ``trans_offset`` (new_cdf - prev_cdf).
* ``matched_filter_prediction`` (rf_rendering.py:686-716) and
  ``compute_mf_loss`` (rf_rendering.py:196-209, used at :814-817 as
  ``mean(compute_mf_loss(pred, gt * loss_scale))``).

Semantics are the reference's; the calibration runs in FP32 where the
reference evaluates its SDF under FP16 autocast (the networks here run BF16).
"""
from __future__ import annotations

from typing import Callable, Dict, Optional, Tuple

import torch
import torch.nn.functional as F

from geraf.ops import _native

SPHERE_RADIUS = 1.0  # rf_rendering.py:27


def intersect_sphere(r_pos_norm: torch.Tensor, norm_pts: torch.Tensor,
                     radius: float = SPHERE_RADIUS) -> Tuple[torch.Tensor, torch.Tensor]:
    """Rays from the centre of the (normalised) antennas r_pos_norm [A,3]
    through every sample norm_pts [R,S,3] against the sphere |x| = radius
    (GeRaFStage1.intersect_sphere, rf_rendering.py:455-495).  Returns
    (valid [1,R,S] bool, hit [1,R,S,3]).  A hit that the reference leaves at
    c + inf * dir (no positive root) is the centre c here; callers use hits
    only where valid."""
    if not (r_pos_norm.is_cuda and norm_pts.is_cuda):
        raise ValueError("intersect_sphere inputs must be CUDA tensors.")
    if norm_pts.dim() != 3 or norm_pts.shape[-1] != 3:
        raise ValueError(f"`norm_pts` must have shape [R, S, 3], got {tuple(norm_pts.shape)}.")
    R, S, _ = norm_pts.shape
    center = r_pos_norm.float().mean(dim=0).contiguous()
    pts = norm_pts.detach().float().contiguous()
    hit = torch.empty((R, S, 3), dtype=torch.float32, device=pts.device)
    valid = torch.empty((R, S), dtype=torch.bool, device=pts.device)
    if R * S:
        with torch.cuda.device(pts.device):
            _native.call("geraf_intersect_sphere_f32", _native.ptr(center), _native.ptr(pts),
                         _native.ptr(hit), _native.ptr(valid), R * S, float(radius),
                         _native.stream_of(pts))
    return valid[None], hit[None]


def transmission_offset(sdf_fn: Callable[[torch.Tensor], torch.Tensor], inv_s: torch.Tensor,
                        sampled_poses_norm: torch.Tensor,
                        r_pos_norm_sel: torch.Tensor) -> torch.Tensor:
    """new_cdf - prev_cdf [R,S] of _calibrate_transmission
    (rf_rendering.py:611-626): prev_cdf = sigmoid(inv_s * sdf(first sample of
    each ray)), new_cdf = sigmoid(inv_s * sdf(sphere hit)) where the hit is
    valid, prev_cdf elsewhere.  ``sdf_fn`` maps points [M,3] to SDF values
    [M] (or [M,1]); ``inv_s`` is the clipped inverse deviation (a scalar
    tensor).  No gradient flows (the reference runs it under no_grad)."""
    R, S, _ = sampled_poses_norm.shape
    dev = sampled_poses_norm.device
    with torch.no_grad():
        valid, hit = intersect_sphere(r_pos_norm_sel, sampled_poses_norm)
        sdf_hit = sdf_fn(hit.reshape(R * S, 3)).reshape(R, S).float().contiguous()
        sdf_first = sdf_fn(sampled_poses_norm[:, 0, :].float().contiguous()).reshape(R)
        sdf_first = sdf_first.float().contiguous()
        inv = inv_s.detach().float().reshape(1).contiguous()
        off = torch.empty((R, S), dtype=torch.float32, device=dev)
        if R * S:
            with torch.cuda.device(dev):
                _native.call("geraf_calibration_offset_f32", _native.ptr(sdf_hit),
                             _native.ptr(sdf_first), _native.ptr(valid[0].contiguous()),
                             _native.ptr(inv), _native.ptr(off), R, S, _native.stream_of(off))
    return off


def matched_filter_prediction(mix_real: torch.Tensor, mix_imag: torch.Tensor,
                              mix_t_pos: torch.Tensor, mix_r_pos: torch.Tensor,
                              tgt_sampled_poses_glb: torch.Tensor, gt_sar_value: torch.Tensor,
                              inbounds: Optional[torch.Tensor], radar_cfg: Dict[str, float],
                              backend: str = "cuda",
                              matched_filter: Optional[Callable] = None
                              ) -> Tuple[torch.Tensor, torch.Tensor]:
    """|MF| of the mixed echoes at the in-bound target points and the matching
    ground truth (GeRaFStage1._matched_filter_prediction,
    rf_rendering.py:686-716): returns (pred [n,1], gt [n,1]) with
    pred = hypot(MF) and gt = gt_sar_value[inbounds] / sub_adc (when the radar
    config has ``sub_adc``).  ``inbounds=None`` takes every point (no mask
    gather, no host synchronisation).  ``matched_filter(real, imag, t_pos,
    r_pos, points)`` replaces base_matched_filter (e.g. a tiled form with a
    prebuilt plan)."""
    if inbounds is not None:
        pts = tgt_sampled_poses_glb[inbounds, :]
        gt = gt_sar_value[inbounds]
    else:
        pts = tgt_sampled_poses_glb.reshape(-1, 3)
        gt = gt_sar_value.reshape(-1)
    n = gt.shape[0]
    if matched_filter is None:
        from geraf.models.model_utils.matched_filter import base_matched_filter
        sar_real, sar_imag = base_matched_filter(real=mix_real, image=mix_imag, t_pos=mix_t_pos,
                                                 r_pos=mix_r_pos, points=pts,
                                                 radar_cfg=radar_cfg, backend=backend)
    else:
        sar_real, sar_imag = matched_filter(mix_real, mix_imag, mix_t_pos, mix_r_pos, pts)
    pred = torch.hypot(sar_real, sar_imag).reshape(n, 1)
    gt = gt.reshape(n, 1)
    if "sub_adc" in radar_cfg:
        gt = gt / radar_cfg["sub_adc"]
    return pred, gt


CHARBONNIER_EPS = 0.001  # rf_rendering.py:207-209


def compute_mf_loss(pred: torch.Tensor, gt: torch.Tensor, loss_mode: str) -> torch.Tensor:
    """Per-point loss, summed over the last axis (GeRaFStage*.compute_mf_loss,
    rf_rendering.py:196-209): "l2", "l1", "smooth_l1" (beta 0.25) or
    "charbonier" (sqrt(sum (gt - pred)^2 + 1e-3)); any other mode raises
    NotImplementedError."""
    if loss_mode == "l2":
        return (pred - gt).square().sum(dim=-1)
    if loss_mode == "l1":
        return (pred - gt).abs().sum(dim=-1)
    if loss_mode == "smooth_l1":
        return F.smooth_l1_loss(pred, gt, reduction="none", beta=0.25).sum(dim=-1)
    if loss_mode == "charbonier":
        return torch.sqrt((gt - pred).square().sum(dim=-1) + CHARBONNIER_EPS)
    raise NotImplementedError(f"loss_mode {loss_mode!r}")
