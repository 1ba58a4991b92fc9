"""Matched-filter scenes: antenna arrays, voxel grids, scatterers and echoes.

Units follow the reference kernels: positions in metres, frequencies in 1e8 Hz,
chirp slope As_sci in 1e8 Hz/s, ADC rate fs in Hz, c in 1e8 m/s.  The matched
filter's start frequency is the reference's hard-coded 773.704 (77.37 GHz,
MatchedFilter_cuda_kernel.cu:15), so a "77-81 GHz, K samples" sweep is
realised as 77.37 GHz + k * 4 GHz/(K-1), k = 0..K-1 (inclusive end points),
with fs = 10 MHz and As_sci = df * fs / 1e8.  Scalar parameters are rounded to
float32, as the reference's C++ bindings receive them.

Echo model (FP64): the forward model the matched filter is matched to, i.e. the
signal tracer's convention (SignalTracer_cuda_kernel.cu:438-443),
    s[a,k] = sum_m amp_m * exp(+j*2*pi*(f0 + k*df)*tau_am) / (R_t,am * R_r,am)
    tau_am = (R_t,am + R_r,am + 0.15) / c
plus complex Gaussian noise at the stated SNR (relative to mean |s|^2).  The
workload text writes the echo with exp(-j...) and no path bias; with the
reference's exp(-j*phi) matched filter that sign would not focus, so the
matched sign is used.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Tuple

import numpy as np
import torch

SPEED_OF_LIGHT = float(np.float32(2.9979))
MF_F0 = float(np.float32(773.704))
PATH_BIAS = float(np.float32(0.15))


def f32(x: float) -> float:
    return float(np.float32(x))


@dataclass
class MFScene:
    name: str
    tx: np.ndarray          # [A,3] float32
    rx: np.ndarray          # [A,3] float32
    voxels: np.ndarray      # [N,3] float32, C order over grid_shape
    grid_shape: Tuple[int, int, int]
    axes: Tuple[np.ndarray, np.ndarray, np.ndarray]
    scatterers: np.ndarray  # [M,3] float64
    amplitudes: np.ndarray  # [M] complex128
    K: int
    fs: float
    As_sci: float
    f0: float = MF_F0
    c: float = SPEED_OF_LIGHT
    bias: float = PATH_BIAS
    echo: np.ndarray = field(default=None, repr=False)  # [A,K] complex64

    @property
    def A(self) -> int:
        return self.tx.shape[0]

    @property
    def N(self) -> int:
        return self.voxels.shape[0]

    def params(self):
        return dict(f0=self.f0, slope=self.As_sci, fs=self.fs, c=self.c, bias=self.bias)

    def terms(self) -> int:
        return self.A * self.K * self.N


def sweep(K: int, bandwidth_hz: float = 4.0e9, fs: float = 10.0e6) -> Tuple[float, float]:
    """(fs, As_sci) for K samples spanning `bandwidth_hz` end to end."""
    df_hz = bandwidth_hz / (K - 1)
    return f32(fs), f32(df_hz * fs / 1.0e8)


def planar_grid(n: int, pitch: float) -> np.ndarray:
    c = (np.arange(n) - (n - 1) / 2.0) * pitch
    gx, gy = np.meshgrid(c, c, indexing="ij")
    return np.stack([gx.ravel(), gy.ravel(), np.zeros(n * n)], axis=1)


def voxel_grid(n: int, side: float, center) -> Tuple[np.ndarray, tuple]:
    ax = np.linspace(-side / 2, side / 2, n)
    axes = (ax + center[0], ax + center[1], ax + center[2])
    gx, gy, gz = np.meshgrid(*axes, indexing="ij")
    vox = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], axis=1).astype(np.float32)
    return vox, tuple(a.astype(np.float32) for a in axes)


def synth_echo(tx, rx, scat, amp, K, fs, As_sci, f0=MF_F0, c=SPEED_OF_LIGHT, bias=PATH_BIAS,
               device="cpu", chunk_terms=1 << 27) -> torch.Tensor:
    """Noise-free echo [A,K] complex128, evaluated in FP64 on `device`."""
    dev = torch.device(device)
    t = torch.as_tensor(np.asarray(tx, np.float64), device=dev)
    r = torch.as_tensor(np.asarray(rx, np.float64), device=dev)
    x = torch.as_tensor(np.asarray(scat, np.float64), device=dev)
    am = torch.as_tensor(np.asarray(amp, np.complex128), device=dev)
    k = torch.arange(K, dtype=torch.float64, device=dev)
    freq = f0 + k * (As_sci / fs)                           # [K] in 1e8 Hz
    A, M = t.shape[0], x.shape[0]
    out = torch.empty((A, K), dtype=torch.complex128, device=dev)
    rows = max(1, chunk_terms // max(1, K * M))
    for a0 in range(0, A, rows):
        a1 = min(A, a0 + rows)
        rt = torch.linalg.norm(t[a0:a1, None, :] - x[None], dim=-1)   # [a,M]
        rr = torch.linalg.norm(r[a0:a1, None, :] - x[None], dim=-1)
        tau = (rt + rr + bias) / c
        w = am[None] / (rt * rr)                                         # [a,M]
        cyc = tau[:, :, None] * freq[None, None, :]                       # [a,M,K]
        cyc = cyc - torch.round(cyc)
        ph = torch.polar(torch.ones_like(cyc), 2.0 * np.pi * cyc)
        out[a0:a1] = torch.einsum("am,amk->ak", w, ph)
    return out


def add_noise(echo: np.ndarray, snr_db: float, rng: np.random.Generator) -> np.ndarray:
    p = np.mean(np.abs(echo) ** 2) / (10.0 ** (snr_db / 10.0))
    n = rng.standard_normal(echo.shape) + 1j * rng.standard_normal(echo.shape)
    return echo + np.sqrt(p / 2.0) * n


def config1(seed: int = 0, device: str = "cpu") -> MFScene:
    """Monostatic 16x16 aperture, 1.95 mm pitch, K=128, 32^3 voxels over a
    10 cm cube at z=25 cm, 200 scatterers on a 3 cm sphere, 20 dB SNR."""
    rng = np.random.Generator(np.random.PCG64(seed))
    pos = planar_grid(16, 1.95e-3).astype(np.float32)
    center = np.array([0.0, 0.0, 0.25])
    vox, axes = voxel_grid(32, 0.10, center)
    M = 200
    d = rng.standard_normal((M, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    scat = center + 0.03 * d
    amp = rng.uniform(0.5, 1.0, M) * np.exp(1j * rng.uniform(0.0, 2 * np.pi, M))
    K = 128
    fs, As = sweep(K)
    clean = synth_echo(pos, pos, scat, amp, K, fs, As, device=device).cpu().numpy()
    echo = add_noise(clean, 20.0, rng).astype(np.complex64)
    return MFScene("mf_cfg1_mono16x16_K128_32c", pos, pos.copy(), vox, (32, 32, 32), axes, scat,
                   amp, K, fs, As, echo=echo)


def box_torus_scatterers(rng: np.random.Generator, M: int, center, box=0.16, wall=0.005,
                         major=0.06, minor=0.015, box_share=0.6) -> np.ndarray:
    """Points in the 5 mm walls of a hollow cube (side `box`) and on a torus
    (major radius `major`, tube radius `minor`) centred inside it."""
    mb = int(round(M * box_share))
    pts = []
    while len(pts) < mb:  # rejection sample the wall shell
        p = rng.uniform(-box / 2, box / 2, (4 * mb, 3))
        inner = np.all(np.abs(p) < box / 2 - wall, axis=1)
        pts.extend(p[~inner][: mb - len(pts)])
    wall_pts = np.asarray(pts[:mb])
    mt = M - mb
    u = rng.uniform(0, 2 * np.pi, mt)
    v = rng.uniform(0, 2 * np.pi, mt)
    tor = np.stack([(major + minor * np.cos(v)) * np.cos(u), (major + minor * np.cos(v)) * np.sin(u),
                    minor * np.sin(v)], axis=1)
    return np.concatenate([wall_pts, tor], axis=0) + np.asarray(center)


def config2(seed: int = 0, device: str = "cpu", n_side: int = 128, n_ap: int = 64,
            K: int = 512, slab_z=None) -> MFScene:
    """MIMO 64x64 virtual array (tx/rx offset -/+1 mm in x, 1.95 mm pitch),
    K=512, 128^3 voxels over a 20 cm cube at z=30 cm, 5,000 scatterers on a
    16 cm hollow box (5 mm walls) enclosing a 6 cm torus, log-normal(0,0.5)
    amplitudes, 10 dB SNR."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = planar_grid(n_ap, 1.95e-3)
    off = np.array([1.0e-3, 0.0, 0.0])
    tx = (p - off).astype(np.float32)
    rx = (p + off).astype(np.float32)
    center = np.array([0.0, 0.0, 0.30])
    vox, axes = voxel_grid(n_side, 0.20, center)
    if slab_z is not None:  # keep the z-planes [slab_z[0], slab_z[1]) of the grid
        z0, z1 = slab_z
        vox = vox.reshape(n_side, n_side, n_side, 3)[:, :, z0:z1].reshape(-1, 3).copy()
        axes = (axes[0], axes[1], axes[2][z0:z1])
    M = 5000
    scat = box_torus_scatterers(rng, M, center)
    amp = rng.lognormal(0.0, 0.5, M) * np.exp(1j * rng.uniform(0.0, 2 * np.pi, M))
    fs, As = sweep(K)
    clean = synth_echo(tx, rx, scat, amp, K, fs, As, device=device).cpu().numpy()
    echo = add_noise(clean, 10.0, rng).astype(np.complex64)
    shape = (n_side, n_side, n_side if slab_z is None else slab_z[1] - slab_z[0])
    return MFScene(f"mf_cfg2_mimo{n_ap}x{n_ap}_K{K}_{n_side}c", tx, rx, vox, shape, axes, scat,
                   amp, K, fs, As, echo=echo)


def nearest_voxels(scene: MFScene, pts: np.ndarray) -> np.ndarray:
    """Flat indices of the grid voxels nearest to `pts` [M,3]."""
    idx = []
    for d in range(3):
        ax = scene.axes[d].astype(np.float64)
        step = ax[1] - ax[0]
        i = np.clip(np.rint((pts[:, d] - ax[0]) / step).astype(np.int64), 0, len(ax) - 1)
        idx.append(i)
    nx, ny, nz = scene.grid_shape
    return (idx[0] * ny + idx[1]) * nz + idx[2]
