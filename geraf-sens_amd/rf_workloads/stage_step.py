"""Stage-1/2-style training steps built on the RF ops (the step workloads of
the benchmark).

One step (GeRaFStage1.loss, rf_rendering.py:721-820, in its per-ray form):
  1. SDF MLP (8 x 256, softplus(beta=100), skip at 4, 6-frequency PE,
     geometric init) on the normalised sample points -> sdf, features, and the
     SDF gradient with create_graph (the normals, differentiated again);
  2. inv_s from the variance network; reflectance from the 4 x 256 signal MLP
     on [features, points]; power = exp(light_power);
  3. the transmittance calibration (rf_rendering.py:611-626,749-750): rays
     from the array centre through every sample hit the unit sphere; the SDF
     there and at each ray's first sample give the offset new_cdf - prev_cdf;
  4. the fused lensless tracer renders each ray's FMCW signal (its antenna is
     the ray origin);
  5. the matched filter backprojects all rays' signals onto the valid sample
     points (rays from one aperture position are one antenna, whose echo is
     the sum of their signals); the prediction |MF| against the target volume
     sampled there (trilinear, divided by trans_power as the dataset loads it,
     then by sub_adc), the Charbonnier loss (rf_rendering.py:686-716,814-817),
     plus the eikonal term (rf_rendering.py:645-650);
  6. backward (matched-filter adjoint, tracer adjoint, MLP double-backward)
     and an AdamW step.

``form="fused"`` is the MI355X path (the SDF network, its normals and their
backward as the library's BF16 matrix-core products, the reflectance MLP in
BF16 with split-K weight gradients, the hand-written tracer kernels);
``form="fp32_autograd"`` is the
all-FP32 autograd form of the same step: FP32 nn.Linear MLPs and the
compositing + signal as torch FP32 ops (the reference's glue style), with the
same matched filter.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch
import torch.nn.functional as F

from geraf.models.networks.fused_reflectance import fused_reflectance
from geraf.models.networks.fused_sdf import fused_sdf, fused_sdf_value
from geraf.models.networks.sdf_network import (ReflectivePowerNetwork, SDFNetwork,
                                                SingleVarianceNetwork)
from geraf.models.rendering.rf_stage_ops import (compute_mf_loss, matched_filter_prediction,
                                                 transmission_offset)
from geraf.ops.RFMatchedFilter import matched_filter_cuda
from geraf.ops.RFMatchedFilter.tiled_matched_filter import make_plan, matched_filter_tiled
from geraf.ops.RFSignalTracer.lensless_tracer import lensless_signal_tracer

from . import tracer_scenes as ts
from .mf_scenes import PATH_BIAS, SPEED_OF_LIGHT

CENTER = (0.0, 0.0, 0.30)
RADIUS = 0.3
TRANS_POWER = 0.0158489319  # radar config trans_power: MF volumes are divided by it
LOSS_MODE = "charbonier"    # stage configs: loss_mode, loss_scale
LOSS_SCALE = 1.0


def torch_lensless(sdf, grad, dists, z, o, d, inv_s, power, refl, n_valid, K, fs, As_sci, f0,
                   anneal=1.0, offset=None):
    """The per-ray compositing + signal as plain torch ops (FP32, autograd);
    ``offset`` is the calibration new_cdf - prev_cdf added to the
    transmittance (rf_rendering.py:750)."""
    R, S = sdf.shape
    true_cos = (d[:, None, :] * grad).sum(-1)
    iter_cos = -(F.relu(-true_cos * 0.5 + 0.5) * (1.0 - anneal) + F.relu(-true_cos) * anneal)
    nxt = sdf + iter_cos * dists * 0.5
    prv = sdf - iter_cos * dists * 0.5
    pc = torch.sigmoid(prv * inv_s)
    nc = torch.sigmoid(nxt * inv_s)
    alpha = ((pc - nc + 1e-5) / (pc + 1e-5)).clip(0.0, 1.0)
    trans = torch.cumprod(torch.cat([torch.ones_like(alpha[:, :1]), 1.0 - alpha + 1e-7], -1),
                          -1)[:, :-1]
    if offset is not None:
        trans = trans + offset
    sigma = torch.clamp(refl * power * trans, min=1e-6) * (alpha * trans)
    n = F.normalize(grad, dim=-1)
    pts = o[:, None, :] + z[..., None] * d[:, None, :]
    v = pts - o[:, None, :]
    inc = v / (v.norm(dim=-1, keepdim=True) + 1e-7)
    idn = (inc * n).sum(-1)
    rr = inc - 2.0 * idn[..., None] * n
    rd = rr / (rr.norm(dim=-1, keepdim=True) + 1e-7)
    spec = (-inc * rd).sum(-1)
    valid = torch.arange(S, device=sdf.device)[None, :] < n_valid[:, None]
    live = valid & (spec >= 1e-6) & (idn <= 0)
    dpath = 2.0 * v.norm(dim=-1)
    amp = torch.where(live, sigma * spec / dpath ** 2, torch.zeros_like(sigma))
    tau = (dpath + PATH_BIAS) / SPEED_OF_LIGHT
    k = torch.arange(K, device=sdf.device, dtype=sdf.dtype)
    ph = 2.0 * math.pi * (f0 + As_sci * k / fs)[None, None, :] * tau[..., None]
    out_r = (amp[..., None] * torch.cos(ph)).sum(1) / n_valid[:, None]
    out_i = (amp[..., None] * torch.sin(ph)).sum(1) / n_valid[:, None]
    return out_r, out_i


def antenna_groups(o: torch.Tensor):
    """Rays that leave the same aperture position are one antenna: the matched
    filter is linear in the echo and its phase depends on the antenna position
    only, so backprojecting the sum of their signals from that position equals
    backprojecting each ray's signal from its own origin.  Returns the distinct
    positions [U,3] and an index matrix [U, max rays per position] into the
    rays, padded with R (a zero row appended to the signals), and each ray's
    antenna [R]; the sum over a row of the matrix is taken in a fixed order
    (deterministic, no atomics)."""
    R = o.shape[0]
    pos, inv = torch.unique(o, dim=0, return_inverse=True)
    U = pos.shape[0]
    counts = torch.bincount(inv, minlength=U)
    order = torch.argsort(inv, stable=True)
    start = torch.cumsum(counts, 0) - counts
    slot = torch.arange(R, device=o.device) - start[inv[order]]
    idx = torch.full((U, int(counts.max())), R, dtype=torch.long, device=o.device)
    idx[inv[order], slot] = order
    return pos.contiguous(), idx, inv


class _MergeRays(torch.autograd.Function):
    """Sum of each antenna's rays' signals; its adjoint hands every ray its
    antenna's gradient row (a gather: no accumulation, unlike the autograd of
    the padded gather, which serialises on the padding index)."""

    @staticmethod
    def forward(ctx, sig, idx, inv):
        ctx.save_for_backward(inv)
        ext = torch.cat([sig, sig.new_zeros(1, sig.shape[1])], 0)
        return ext[idx].sum(1)

    @staticmethod
    def backward(ctx, g):
        (inv,) = ctx.saved_tensors
        return g[inv], None, None


def merge_rays(sig: torch.Tensor, idx: torch.Tensor, inv: torch.Tensor) -> torch.Tensor:
    """Per-antenna echo: the sum of its rays' signals [R,K] -> [U,K]."""
    return _MergeRays.apply(sig, idx, inv)


@dataclass
class StageConfig:
    rays: int
    samples: int
    K: int
    min_valid: int
    consistency_points: int = 0


STAGE1 = StageConfig(rays=8192, samples=64, K=64, min_valid=16)
STAGE2 = StageConfig(rays=32768, samples=256, K=512, min_valid=256, consistency_points=262144)


class StageStep:
    def __init__(self, cfg: StageConfig, target_volume: torch.Tensor, volume_axes,
                 device: torch.device, form: str = "fused", seed: int = 0,
                 mlp_dtype: Optional[torch.dtype] = "auto", mf_algo: str = "auto"):
        """``mlp_dtype``: compute dtype of the MLP GEMMs; "auto" is BF16 for the
        fused form and FP32 for the autograd form.  A BF16 fused form runs the
        SDF network on the hand-written kernels (``fused_sdf``).  ``mf_algo``: "direct" or
        "tiled" (the per-tile expansion matched filter, truncation bound 1e-6);
        "auto" is tiled for the fused form and direct for the autograd form."""
        if form not in ("fused", "fp32_autograd"):
            raise ValueError(f"unknown step form {form!r}")
        self.cfg, self.form, self.dev = cfg, form, device
        torch.manual_seed(seed)
        if mlp_dtype == "auto":
            mlp_dtype = torch.bfloat16 if form == "fused" else None
        cd = mlp_dtype
        tall = form == "fused"  # the autograd form keeps plain nn.Linear layers
        self.hand_sdf = form == "fused" and cd == torch.bfloat16
        self.sdf_net = SDFNetwork(3, 257, 256, 8, skip_in=(4,), multires=6, compute_dtype=cd,
                                  tall_products=tall).to(device)
        self.var_net = SingleVarianceNetwork(0.3).to(device)
        self.sig_net = ReflectivePowerNetwork(256, light_power=0.0, compute_dtype=cd,
                                              tall_products=tall).to(device)
        params = list(self.sdf_net.parameters()) + list(self.var_net.parameters()) + \
            list(self.sig_net.parameters())
        self.opt = torch.optim.AdamW(params, lr=5e-4)
        uneven = None if cfg.min_valid >= cfg.samples else (cfg.min_valid, cfg.samples)
        b = ts.rays(cfg.rays, cfg.samples, cfg.K, seed=seed, uneven=uneven).to(device)
        if b.n_valid is None:
            b.n_valid = torch.full((cfg.rays,), cfg.samples, dtype=torch.int32, device=device)
        self.b = b
        center = torch.tensor(CENTER, device=device)
        self.pts = b.o[:, None, :] + b.z[..., None] * b.d[:, None, :]           # [R,S,3] m
        self.pts_n = ((self.pts - center) / RADIUS).reshape(-1, 3)              # SDF frame
        self.valid = (torch.arange(cfg.samples, device=device)[None, :] <
                      b.n_valid[:, None].long())
        self.tgt = self.pts[self.valid].contiguous()                           # [N,3]
        # target |MF| at the valid sample points: trilinear in the volume
        # (grid_sample axis order (z, y, x), align_corners=True, as loading.py:20-25)
        lo = torch.tensor([float(a[0]) for a in volume_axes], device=device)
        hi = torch.tensor([float(a[-1]) for a in volume_axes], device=device)
        g = (self.tgt - lo) / (hi - lo) * 2.0 - 1.0
        vol = target_volume.abs().float()[None, None]                          # [1,1,X,Y,Z]
        samp = F.grid_sample(vol, g[:, [2, 1, 0]].view(1, -1, 1, 1, 3), align_corners=True)
        # the dataset divides the volume by trans_power (loading.py:51-58) and
        # samples it at the targets; the prediction divides by sub_adc
        self.gt = samp.view(-1) / TRANS_POWER
        if cfg.consistency_points:
            gen = torch.Generator(device="cpu").manual_seed(seed + 1)
            self.cons_pts = (torch.rand(cfg.consistency_points, 3, generator=gen) * 2 - 1) \
                .to(device) * 0.6
        # sub_adc: the K samples are every (512 / K)-th of the 512-sample chirp
        self.radar = dict(b.radar, sub_adc=max(1, 512 // cfg.K))
        if mf_algo == "auto":
            mf_algo = "tiled" if form == "fused" else "direct"
        if mf_algo not in ("direct", "tiled"):
            raise ValueError(f"unknown matched-filter algorithm {mf_algo!r}")
        self.mf_algo = mf_algo
        self.plan = (make_plan(self.tgt, cfg.K, self.radar["adcSampleRate"], self.radar["As_sci"])
                     if mf_algo == "tiled" else None)
        # the matched filter's antennas: the distinct ray origins (their rays'
        # signals summed).  Scaling the sums by U / R keeps the 1/A
        # normalisation of the per-ray form, so the prediction is unchanged.
        self.ant_pos, self.ant_idx, self.ant_of_ray = antenna_groups(b.o)
        self.ant_scale = self.ant_pos.shape[0] / cfg.rays
        self.ant_pos_norm = ((self.ant_pos - center) / RADIUS).contiguous()

    def calibration_offset(self) -> torch.Tensor:
        """trans_offset [R,S] = new_cdf - prev_cdf from the sphere hits of rays
        leaving the centre of the (normalised) antennas (no gradient)."""
        R, S = self.cfg.rays, self.cfg.samples
        inv_s = self.var_net.inv_s().clip(1e-6, 1e6)
        sdf_fn = ((lambda x: fused_sdf_value(self.sdf_net, x)) if self.hand_sdf
                  else self.sdf_net.sdf_value)
        return transmission_offset(sdf_fn, inv_s,
                                   self.pts_n.view(R, S, 3), self.ant_pos_norm)

    def parts(self) -> dict:
        """The step's forward, returning its intermediates: network outputs
        (sdf, grad, refl, inv_s, power), the calibration offset, the per-ray
        signals (sim_r, sim_i), the merged antenna echoes (ant_r, ant_i), the
        matched filter at the targets (mf_r, mf_i), pred and gt, the loss
        terms and the total ``loss``."""
        cfg, b = self.cfg, self.b
        R, S = cfg.rays, cfg.samples
        offset = self.calibration_offset()
        # the features stay in the networks' compute dtype (BF16 in the fused
        # form): the reflectance network takes them as they are
        if self.hand_sdf:
            sdf, feat, grad = fused_sdf(self.sdf_net, self.pts_n)
            sdf = sdf.reshape(R, S)
        else:
            out, grad = self.sdf_net.sdf_and_gradient(self.pts_n.clone(), float_out=False)
            sdf = out[:, 0].float().reshape(R, S)
            feat = out[:, 1:]
        grad = grad.reshape(R, S, 3)
        inv_s = self.var_net.inv_s().clip(1e-6, 1e6)
        refl = (fused_reflectance(self.sig_net, self.pts_n, feat) if self.hand_sdf
                else self.sig_net.reflectance(self.pts_n, feat)).reshape(R, S)
        power = self.sig_net.power()
        r = self.radar
        if self.form == "fused":
            sim_r, sim_i = lensless_signal_tracer(sdf, grad, b.dists, b.z, b.o, b.d, inv_s, power,
                                                  reflectance=refl, trans_offset=offset,
                                                  n_valid=b.n_valid, numADCSample=r["numADCSample"],
                                                  adcSampleRate=r["adcSampleRate"],
                                                  fc_start=r["fc_start"], As_sci=r["As_sci"])
        else:
            sim_r, sim_i = torch_lensless(sdf, grad, b.dists, b.z, b.o, b.d, inv_s, power, refl,
                                          b.n_valid.float(), r["numADCSample"],
                                          r["adcSampleRate"], r["As_sci"], r["fc_start"],
                                          offset=offset)
        U = self.ant_pos.shape[0]
        ant_r = merge_rays(sim_r, self.ant_idx, self.ant_of_ray) * self.ant_scale
        ant_i = merge_rays(sim_i, self.ant_idx, self.ant_of_ray) * self.ant_scale
        keep = {}

        def mf(re, im, tp, rp, pts):
            if self.mf_algo == "tiled":
                o = matched_filter_tiled(re, im, tp, rp, pts, U, r["numADCSample"],
                                         r["adcSampleRate"], r["As_sci"], pts.shape[0],
                                         plan=self.plan)
            else:
                o = matched_filter_cuda(re, im, tp, rp, pts, U, r["numADCSample"],
                                        r["adcSampleRate"], r["As_sci"], pts.shape[0])
            keep["mf"] = o
            return o

        # the targets are the valid samples already (the fused tracer's
        # inbounds), so no mask gather
        pred, gt = matched_filter_prediction(ant_r, ant_i, self.ant_pos, self.ant_pos, self.tgt,
                                             self.gt, None, r, matched_filter=mf)
        mf_loss = compute_mf_loss(pred, gt * LOSS_SCALE, LOSS_MODE).mean()
        eik = ((grad.norm(dim=-1) - 1.0) ** 2)[self.valid].mean()
        loss = mf_loss + 0.1 * eik
        cons = None
        if self.cfg.consistency_points:
            # vision-SDF consistency vs an analytic box SDF (half side 0.16/0.3/2)
            h = 0.08 / RADIUS
            q = self.cons_pts.abs() - h
            box = q.clamp(min=0).norm(dim=-1) + q.max(dim=-1).values.clamp(max=0)
            cons_sdf = (fused_sdf(self.sdf_net, self.cons_pts, with_normals=False)[0]
                        if self.hand_sdf else self.sdf_net.sdf(self.cons_pts)[:, 0])
            cons = (cons_sdf - box).abs().mean()
            loss = loss + cons
        return dict(sdf=sdf, grad=grad, refl=refl, inv_s=inv_s, power=power, offset=offset,
                    sim_r=sim_r, sim_i=sim_i, ant_r=ant_r, ant_i=ant_i, mf_r=keep["mf"][0],
                    mf_i=keep["mf"][1], pred=pred, gt=gt, mf_loss=mf_loss, eikonal=eik,
                    consistency=cons, loss=loss)

    def loss(self) -> torch.Tensor:
        return self.parts()["loss"]

    def step(self) -> float:
        self.opt.zero_grad(set_to_none=True)
        loss = self.loss()
        loss.backward()
        self.opt.step()
        return float(loss.detach())
