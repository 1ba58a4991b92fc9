"""FP64 restatement of the lensless compositing glue and the per-ray signal.

TEST INFRASTRUCTURE ONLY (the checker of the fused lensless kernels).

The composition follows the reference's torch glue line by line, in float64:
  iter_cos            rf_rendering.py:145-155 (_compute_iter_cos)
  alpha               rf_rendering.py:157-169 (_compute_alpha_from_sdf), EPS = 1e-5;
                      the cdf difference is taken from the complements in front
                      of the surface (same value, no cancellation)
  transmittance       rf_rendering.py:46-52   (_compute_transmission), TRANS_EPS = 1e-7
  calibration         rf_rendering.py:750     (T - prev_cdf + new_cdf = T + offset)
  sigma               sdf_network.py:475-481  (clamp(refl * exp(light_power) * T, 1e-6) * alpha * T)
  normals             F.normalize(gradients)  (rf_rendering.py:638,1054)
  signal              SignalTracer_cuda_kernel.cu:172-197,351-450 with tx = rx = ray origin
Gradients of this function by torch autograd are the reference composition's
gradients; tests pin them with central finite differences.
"""
from __future__ import annotations

import math

import torch
import torch.nn.functional as F

EPS = 1e-5
TRANS_EPS = 1e-7


def _normalize_eps(v: torch.Tensor) -> torch.Tensor:
    # SignalTracer_cuda_kernel.cu:40-47: v / (|v| + 1e-7)
    return v / (torch.linalg.norm(v, dim=-1, keepdim=True) + 1e-7)


def points_f32(o32: torch.Tensor, z32: torch.Tensor, d32: torch.Tensor) -> torch.Tensor:
    """Sample positions o + z*d formed in float32 (separately rounded product
    and sum), as a float32 tensor op produces them."""
    return (o32[:, None, :].float() + z32[:, :, None].float() * d32[:, None, :].float()).double()


def lensless_forward(sdf, grad, dists, pts, o, d, inv_s, power, K, fs, As_sci, f0, c, bias,
                     refl=None, offset=None, n_valid=None, anneal=1.0, return_parts=False):
    """All tensors float64 on the CPU; pts [R,S,3] from points_f32.  Returns the
    complex signal [R,K] (complex128)."""
    R, S = sdf.shape
    true_cos = (d[:, None, :] * grad).sum(-1)
    iter_cos = -(F.relu(-true_cos * 0.5 + 0.5) * (1.0 - anneal) + F.relu(-true_cos) * anneal)
    next_sdf = sdf + iter_cos * dists * 0.5
    prev_sdf = sdf - iter_cos * dists * 0.5
    tp = prev_sdf * inv_s
    tn = next_sdf * inv_s
    # prev_cdf - next_cdf, formed as (1 - next_cdf) - (1 - prev_cdf) where
    # prev_cdf >= 1/2, and prev_cdf as 1 - (1 - prev_cdf) there: equal in exact
    # arithmetic, but the literal difference of two cdfs within 1e-14 of 1 (and
    # autograd's slope y (1 - y) of a cdf near 1) keeps only a few digits even
    # in float64.
    front = tp >= 0.0
    cp = torch.sigmoid(-tp)
    prev_cdf = torch.where(front, 1.0 - cp, torch.sigmoid(tp))
    dcdf = torch.where(front, torch.sigmoid(-tn) - cp, prev_cdf - torch.sigmoid(tn))
    alpha = ((dcdf + EPS) / (prev_cdf + EPS)).clip(0.0, 1.0)
    if n_valid is not None:
        valid = torch.arange(S)[None, :] < n_valid[:, None]
    else:
        valid = torch.ones((R, S), dtype=torch.bool)
    ones = torch.ones((R, 1), dtype=sdf.dtype)
    trans = torch.cumprod(torch.cat([ones, 1.0 - alpha + TRANS_EPS], dim=-1), dim=-1)[..., :-1]
    tc = trans + (offset if offset is not None else 0.0)
    rp = (refl if refl is not None else torch.ones_like(sdf)) * power
    sigma = torch.clamp(rp * tc, min=1e-6) * (alpha * tc)
    n = F.normalize(grad, dim=-1)
    # monostatic specular geometry (the reference tracer with tx = rx = origin)
    oo = o[:, None, :].expand_as(pts)
    inc = _normalize_eps(pts - oo)
    idn = (inc * n).sum(-1)
    refl_raw = inc - 2.0 * idn[..., None] * n
    refl_dir = _normalize_eps(refl_raw)
    ret = _normalize_eps(oo - pts)
    spec = (ret * refl_dir).sum(-1)
    live = valid & (spec >= 1e-6) & (idn <= 0.0)
    dpath = 2.0 * torch.linalg.norm(pts - oo, dim=-1)
    amp = torch.where(live, sigma * spec / dpath ** 2, torch.zeros_like(sigma))
    tau = (dpath + bias) / c
    k = torch.arange(K, dtype=sdf.dtype)
    phase = 2.0 * math.pi * (f0 + As_sci * (k / fs))[None, None, :] * tau[..., None]
    sig = (amp[..., None] * torch.polar(torch.ones_like(phase), phase)).sum(1)
    nv = valid.sum(-1).clamp(min=1).to(sdf.dtype)
    out = sig / nv[:, None]
    if return_parts:
        return out, dict(alpha=alpha, trans=trans, sigma=torch.where(valid, sigma, 0 * sigma),
                         spec=spec, live=live, true_cos=true_cos, amp=amp)
    return out


def masks(sdf, grad, dists, pts, o, d, inv_s, power, refl=None, offset=None, n_valid=None,
          anneal=1.0):
    """The piecewise branches a finite difference must not cross."""
    _, parts = lensless_forward(sdf, grad, dists, pts, o, d, inv_s, power, 1, 1.0, 0.0, 1.0, 1.0,
                                0.0, refl=refl, offset=offset, n_valid=n_valid, anneal=anneal,
                                return_parts=True)
    return parts


# ----------------------------------------------------------- transmittance calibration

SPHERE_RADIUS = 1.0  # rf_rendering.py:27
HIT_EPS = 1e-5       # rf_rendering.py:484-490 (t1 > 1e-5 ? t1 : t2 > 1e-5 ? t2 : inf)
DIR_EPS = 1e-8       # rf_rendering.py:466-468


def intersect_sphere(r_pos_norm, norm_pts, radius=SPHERE_RADIUS):
    """FP64 numpy restatement of GeRaFStage1.intersect_sphere
    (rf_rendering.py:455-495): rays from the centre of the antennas
    (the mean of r_pos_norm [A,3]) through every sample norm_pts [R,S,3] meet
    the sphere |x| = radius at t1 <= t2; t is t1 if t1 > 1e-5, else t2 if
    t2 > 1e-5, else none.  Returns (valid [1,R,S] bool, hit [1,R,S,3]); hit
    rows of invalid samples are +-inf / nan, as the reference's c + inf * dir."""
    import numpy as np

    c = np.mean(np.asarray(r_pos_norm, dtype=np.float64), axis=0)
    p = np.asarray(norm_pts, dtype=np.float64)
    v = p - c
    dirs = v / (np.linalg.norm(v, axis=-1, keepdims=True) + DIR_EPS)
    b = 2.0 * np.sum(c * dirs, axis=-1)
    cc = np.dot(c, c) - radius ** 2
    disc = b * b - 4.0 * cc
    sq = np.sqrt(np.maximum(disc, 0.0))
    t1 = (-b - sq) / 2.0
    t2 = (-b + sq) / 2.0
    t = np.where(t1 > HIT_EPS, t1, np.where(t2 > HIT_EPS, t2, np.inf))
    valid = (disc >= 0.0) & np.isfinite(t)
    with np.errstate(invalid="ignore"):
        hit = c + t[..., None] * dirs
    return valid[None], hit[None]


def transmission_offset(sdf_hit, sdf_first, valid, inv_s):
    """FP64 numpy restatement of _calibrate_transmission (rf_rendering.py:611-626)
    and its use at :749-750: prev_cdf = sigmoid(inv_s * sdf at each ray's
    first sample) [R,1], new_cdf = sigmoid(inv_s * sdf at the sphere hit) where
    the hit is valid, prev_cdf elsewhere; the calibrated transmittance
    (T - prev_cdf) + new_cdf is T + offset with offset = new_cdf - prev_cdf
    [R,S].  sdf_hit [R,S] (any value where not valid), sdf_first [R]."""
    import numpy as np

    sig = lambda x: 1.0 / (1.0 + np.exp(-x))  # noqa: E731
    prev = sig(np.asarray(sdf_first, dtype=np.float64) * inv_s)[:, None]
    with np.errstate(invalid="ignore", over="ignore"):
        hit = sig(np.asarray(sdf_hit, dtype=np.float64) * inv_s)
    new = np.where(np.asarray(valid, dtype=bool), hit, prev)
    return new - prev
