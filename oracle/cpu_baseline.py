"""The reference's matched-filter formula restated in torch-CPU FP32.

TEST AND CPU-BASELINE INFRASTRUCTURE ONLY: timed by bench.py as the CPU baseline
(the reference itself has no CPU path: its ops assert CUDA,
matched_filter_cuda.py:247).  Per term it does what
MatchedFilter_cuda_kernel.cu:87-111 does - FP32 distance, FP32 phase, one
cos and one sin - vectorised over (voxel, antenna, sample) chunks.
"""
from __future__ import annotations

import math
import time

import numpy as np
import torch


def mf_forward_fp32(echo: np.ndarray, tx, rx, vox, f0, slope, fs, c, bias,
                    chunk_voxels: int = 8) -> np.ndarray:
    re = torch.as_tensor(np.ascontiguousarray(echo.real), dtype=torch.float32)
    im = torch.as_tensor(np.ascontiguousarray(echo.imag), dtype=torch.float32)
    t = torch.as_tensor(np.asarray(tx), dtype=torch.float32)
    r = torch.as_tensor(np.asarray(rx), dtype=torch.float32)
    v = torch.as_tensor(np.asarray(vox), dtype=torch.float32)
    A, K = re.shape
    two_pi = torch.tensor(2.0 * math.pi, dtype=torch.float32)
    st = torch.arange(K, dtype=torch.float32) / torch.tensor(fs, dtype=torch.float32)
    slope_t = two_pi * torch.tensor(slope, dtype=torch.float32) * st          # [K]
    out = np.empty(v.shape[0], np.complex64)
    for i in range(0, v.shape[0], chunk_voxels):
        vv = v[i:i + chunk_voxels]
        d = torch.linalg.norm(t[None] - vv[:, None], dim=-1) + \
            torch.linalg.norm(r[None] - vv[:, None], dim=-1) + bias            # [n,A]
        tau = d / c
        base = two_pi * f0 * tau
        ph = base[:, :, None] + slope_t[None, None, :] * tau[:, :, None]      # [n,A,K]
        cs, sn = torch.cos(ph), torch.sin(ph)
        o_r = (cs * re + sn * im).sum(dim=(1, 2)) / A
        o_i = (cs * im - sn * re).sum(dim=(1, 2)) / A
        out[i:i + chunk_voxels] = (o_r + 1j * o_i).numpy()
    return out


def time_mf_sample(scene, target_seconds: float = 15.0, threads: int = 0, seed: int = 0):
    """Time the FP32 CPU restatement on a random voxel sample sized for about
    `target_seconds`; returns (terms_per_s, threads, sample_description)."""
    if threads:
        torch.set_num_threads(threads)
    nthreads = torch.get_num_threads()
    rng = np.random.default_rng(seed)
    kw = dict(f0=scene.f0, slope=scene.As_sci, fs=scene.fs, c=scene.c, bias=scene.bias)
    pilot = scene.voxels[rng.integers(0, scene.N, 8)]
    t0 = time.perf_counter()
    mf_forward_fp32(scene.echo, scene.tx, scene.rx, pilot, **kw)
    dt = max(time.perf_counter() - t0, 1e-3)
    n = int(min(scene.N, max(8, 8 * target_seconds / dt)))
    sample = scene.voxels[rng.integers(0, scene.N, n)]
    t0 = time.perf_counter()
    mf_forward_fp32(scene.echo, scene.tx, scene.rx, sample, **kw)
    dt = time.perf_counter() - t0
    terms = n * scene.A * scene.K
    desc = (f"{n} random voxels of {scene.name} x {scene.A} antennas x {scene.K} samples "
            f"({terms:.3e} terms, {dt:.1f} s)")
    return terms / dt, nthreads, desc


def time_tracer_sample(batch, target_seconds: float = 10.0, threads: int = 0):
    """Time the reference composition (torch-CPU FP32 glue + per-ray signal with
    a per-term cos/sin, forward and autograd backward) on a ray subset sized for
    about `target_seconds`; returns (samples_per_s, threads, description)."""
    from oracle import lensless as L

    if threads:
        torch.set_num_threads(threads)
    nthreads = torch.get_num_threads()
    r = batch.radar

    def run(n):
        leaves = {k: getattr(batch, k)[:n].clone().float().requires_grad_(True)
                  for k in ("sdf", "grad", "refl")}
        inv_s = batch.inv_s.clone().float().requires_grad_(True)
        power = batch.power.clone().float().requires_grad_(True)
        pts = (batch.o[:n, None, :] + batch.z[:n, :, None] * batch.d[:n, None, :]).float()
        out = L.lensless_forward(leaves["sdf"], leaves["grad"], batch.dists[:n], pts,
                                 batch.o[:n], batch.d[:n], inv_s, power, r["numADCSample"],
                                 r["adcSampleRate"], r["As_sci"], r["fc_start"], 2.9979, 0.15,
                                 refl=leaves["refl"], anneal=batch.anneal)
        (out.real.sum() + out.imag.sum()).backward()

    n0 = 16
    while True:  # grow the pilot until it takes a measurable time
        t0 = time.perf_counter()
        run(n0)
        dt = time.perf_counter() - t0
        if dt > 0.5 or n0 >= batch.R:
            break
        n0 = min(batch.R, 4 * n0)
    n = int(min(batch.R, max(n0, n0 * target_seconds / dt)))
    t0 = time.perf_counter()
    run(n)
    dt = time.perf_counter() - t0
    samples = n * batch.S
    desc = (f"{n} rays x {batch.S} samples, K={r['numADCSample']} (fwd+bwd, torch autograd), "
            f"{dt:.1f} s")
    return samples / dt, nthreads, desc
