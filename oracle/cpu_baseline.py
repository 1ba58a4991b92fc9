"""The reference's matched-filter formula restated in torch-CPU FP32.

TEST / BASELINE INFRASTRUCTURE ONLY: timed by bench.py as the CPU baseline
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
