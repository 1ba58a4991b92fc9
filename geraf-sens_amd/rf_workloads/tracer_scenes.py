"""Ray batches for the lensless tracer: rays from an antenna aperture into the
unit sphere of the SDF's frame, an analytic noisy-sphere SDF and its gradient.

Frames: positions in metres in the radar frame; the SDF lives in a normalised
frame x_n = (x - center) / radius (the reference maps the scene to the unit
sphere the same way, radius 0.3 m).  Directions are the same in both frames;
``dists`` are sample spacings in the normalised frame (as the reference's
``dists``), ``z_vals`` depths in metres (for the RF geometry).

Radar constants (reference stage config): fc_start = 773.704, As_sci =
70.15e12 / 1e8, fs = 10 MHz / (512 / K) so that K samples span the 512-sample
chirp, c = 2.9979, path bias 0.15 m; all rounded to float32.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np
import torch

from .mf_scenes import PATH_BIAS, SPEED_OF_LIGHT, f32, planar_grid

FC_START = f32(773.704)
AS_SCI = f32(70.15e12 / 1e8)


def radar_for_k(K: int) -> Dict[str, float]:
    return dict(numADCSample=K, adcSampleRate=f32(10.0e6 * K / 512.0), fc_start=FC_START,
                As_sci=AS_SCI)


class NoisySphereSDF:
    """sdf(x) = |x| - r0 + amp * noise(x), noise a sum of random Fourier modes
    with unit variance (a smooth N(0,1) field); exact gradient."""

    def __init__(self, rng: np.random.Generator, r0=0.5, amp=0.02, modes=16, freq=4.0):
        self.r0, self.amp = r0, amp
        self.k = rng.standard_normal((modes, 3)) * freq
        self.ph = rng.uniform(0, 2 * np.pi, modes)
        self.a = np.full(modes, np.sqrt(2.0 / modes))

    def __call__(self, x: np.ndarray):
        r = np.linalg.norm(x, axis=-1)
        arg = x @ self.k.T + self.ph
        sdf = r - self.r0 + self.amp * (np.sin(arg) @ self.a)
        grad = x / np.maximum(r, 1e-12)[..., None] + self.amp * ((np.cos(arg) * self.a) @ self.k)
        return sdf, grad


@dataclass
class RayBatch:
    sdf: torch.Tensor
    grad: torch.Tensor
    dists: torch.Tensor
    z: torch.Tensor
    o: torch.Tensor
    d: torch.Tensor
    refl: Optional[torch.Tensor]
    offset: Optional[torch.Tensor]
    n_valid: Optional[torch.Tensor]
    inv_s: torch.Tensor
    power: torch.Tensor
    radar: Dict[str, float]
    anneal: float = 1.0

    @property
    def R(self) -> int:
        return self.sdf.shape[0]

    @property
    def S(self) -> int:
        return self.sdf.shape[1]

    def to(self, device):
        f = lambda t: None if t is None else t.to(device)  # noqa: E731
        return RayBatch(f(self.sdf), f(self.grad), f(self.dists), f(self.z), f(self.o),
                        f(self.d), f(self.refl), f(self.offset), f(self.n_valid), f(self.inv_s),
                        f(self.power), self.radar, self.anneal)


def rays(R: int, S: int, K: int, seed: int = 0, n_ap: int = 64, pitch: float = 1.95e-3,
         center=(0.0, 0.0, 0.30), radius: float = 0.3, inv_s: float = 64.0,
         power: float = 1.0, with_refl: bool = True, with_offset: bool = False,
         uneven: Optional[tuple] = None, anneal: float = 1.0) -> RayBatch:
    rng = np.random.Generator(np.random.PCG64(seed))
    center = np.asarray(center)
    ap = planar_grid(n_ap, pitch)
    o = ap[rng.integers(0, len(ap), R)]
    q = rng.standard_normal((R, 3))
    q *= (rng.uniform(0, 1, R) ** (1 / 3) / np.linalg.norm(q, axis=1))[:, None]
    tgt = center + radius * q
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # entry/exit of the ray through the sphere |x - center| = radius
    oc = o - center
    b = np.sum(oc * d, axis=1)
    cc = np.sum(oc * oc, axis=1) - radius ** 2
    disc = np.sqrt(np.maximum(b * b - cc, 0.0))
    t0, t1 = -b - disc, -b + disc
    u = (np.arange(S)[None, :] + rng.uniform(0, 1, (R, S))) / S
    z = t0[:, None] + u * (t1 - t0)[:, None]
    dz = np.diff(z, axis=1)
    dists = np.concatenate([dz, dz[:, -1:]], axis=1) / radius
    xn = (o[:, None, :] + z[..., None] * d[:, None, :] - center) / radius
    field = NoisySphereSDF(rng)
    sdf, grad = field(xn)
    refl = rng.uniform(0.2, 1.0, (R, S)) if with_refl else None
    offset = rng.uniform(-0.02, 0.02, (R, S)) if with_offset else None
    nv = rng.integers(uneven[0], uneven[1] + 1, R) if uneven else None
    t = lambda x, dt=torch.float32: None if x is None else torch.tensor(x, dtype=dt)  # noqa: E731
    return RayBatch(t(sdf), t(grad), t(dists), t(z), t(o), t(d), t(refl), t(offset),
                    t(nv, torch.int32), torch.tensor(inv_s, dtype=torch.float32),
                    torch.tensor(power, dtype=torch.float32), radar_for_k(K), anneal)


def config3(seed: int = 0) -> RayBatch:
    """4,096 rays x 128 stratified samples from the 64x64 aperture into the unit
    sphere, noisy-sphere SDF (r = 0.5, 0.02 smooth noise), inv_s = 64,
    reflectance U(0.2, 1), power 1.0, K = 256."""
    return rays(4096, 128, 256, seed=seed)


def ap_training_case(dev, A_total: int = 4096, bank_size: int = 2, rays: int = 804,
                     samples: int = 64, K: int = 64, seed: int = 0):
    """Inputs of the reference's [A,P] signal tracer at its training shape:
    one antenna bank of a 64x64 MIMO aperture (A = A_total / bank_size =
    2,048 pairs, tx/rx -/+1 mm), P = rays x target samples = 804 x 64 =
    51,456 points in a 20 cm region at 30 cm (90 % in bounds), normals facing
    the array with jitter, per-antenna sigmas U(0.1, 1), K = 64 at the stage
    configs' 1.25 MHz / 70.15 MHz/us chirp."""
    rng = np.random.default_rng(seed)
    p = planar_grid(64, 1.95e-3)[::bank_size][: A_total // bank_size]
    off = np.array([1.0e-3, 0.0, 0.0])
    tx, rx = (p - off).astype(np.float32), (p + off).astype(np.float32)
    A, P = tx.shape[0], rays * samples
    pts = (rng.uniform(-0.1, 0.1, (P, 3)) + [0.0, 0.0, 0.30]).astype(np.float32)
    n = -pts / np.linalg.norm(pts, axis=1, keepdims=True)
    n = (n + 0.3 * rng.standard_normal((P, 3))).astype(np.float32)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    sig = rng.uniform(0.1, 1.0, (A, P)).astype(np.float32)
    inb = rng.uniform(0, 1, P) < 0.9
    radar = radar_for_k(K)
    T = lambda x, dt=None: torch.tensor(np.ascontiguousarray(x), device=dev, dtype=dt)  # noqa
    return dict(normals=T(n), sigmas=T(sig), points=T(pts), tx=T(tx), rx=T(rx),
                inbounds=T(inb, torch.bool), A=A, P=P, K=K, radar=radar)
