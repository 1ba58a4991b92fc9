"""Writes tests/golden/rf_golden.npz: small seeded inputs and their FP64 oracle
outputs for each function on the path (matched filter forward/adjoint, [A,P]
tracer forward/backward, bilinear resampler forward/backward, fused lensless
renderer forward and gradients).

The reference ships no golden vectors or tests, so these vectors come from the
oracle (oracle/), itself pinned by tests/test_oracle.py; freezing them lets the
GPU tests check against fixed data and catches any later drift of the oracle.

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import torch

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "geraf-sens_amd")]

from oracle import lensless as L  # noqa: E402
from oracle import rf as orc  # noqa: E402
from rf_workloads import mf_scenes, tracer_scenes as ts  # noqa: E402

F32 = np.float32


def main():
    out = {}
    rng = np.random.default_rng(20260101)
    # ---- matched filter: A=7, K=45 (not a multiple of 32), N=70
    A, K, N = 7, 45, 70
    tx = rng.uniform(-0.02, 0.02, (A, 3)).astype(F32)
    tx[:, 2] = 0
    rx = (tx + F32([0.001, 0.0005, 0])).astype(F32)
    vox = (rng.uniform(-0.04, 0.04, (N, 3)) + [0, 0, 0.28]).astype(F32)
    s = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    g = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    fs, As = F32(1.25e6), F32(701500.0)
    prm = dict(f0=mf_scenes.MF_F0, slope=float(As), fs=float(fs), c=mf_scenes.SPEED_OF_LIGHT,
               bias=mf_scenes.PATH_BIAS)
    out.update(mf_tx=tx, mf_rx=rx, mf_vox=vox, mf_echo=s, mf_gout=g, mf_fs=fs, mf_as=As,
               mf_out=orc.mf_forward(s.real, s.imag, tx, rx, vox, **prm),
               mf_grad=orc.mf_backward(g, tx, rx, vox, K, norm=A, **prm))
    # ---- [A,P] tracer: A=4, P=300, K=40
    A, P, K = 4, 300, 40
    ttx = rng.uniform(-0.04, 0.04, (A, 3)).astype(F32)
    ttx[:, 2] = 0
    trx = (ttx + F32([0.002, 0, 0])).astype(F32)
    pts = (rng.uniform(-0.08, 0.08, (P, 3)) + [0, 0, 0.3]).astype(F32)
    nrm = (-pts / np.linalg.norm(pts, axis=1, keepdims=True) +
           0.3 * rng.standard_normal((P, 3))).astype(F32)
    sig = rng.uniform(0.1, 1.0, (A, P)).astype(F32)
    inb = rng.uniform(0, 1, P) > 0.15
    gk = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    tp = dict(f0=ts.FC_START, slope=ts.AS_SCI, fs=float(fs), c=ts.SPEED_OF_LIGHT,
              bias=ts.PATH_BIAS)
    gn, gs = orc.st_backward(gk, nrm, sig, pts, ttx, trx, inb, **tp)
    out.update(st_tx=ttx, st_rx=trx, st_pts=pts, st_nrm=nrm, st_sig=sig, st_inb=inb, st_gout=gk,
               st_out=orc.st_forward(nrm, sig, pts, ttx, trx, inb, K, **tp), st_gn=gn, st_gs=gs)
    # ---- bilinear: A=2, R=5, S=12, T=20
    A, R, S, T = 2, 5, 12, 20
    z = np.sort(rng.uniform(0.2, 0.6, (R, S)), axis=1).astype(F32)
    tz = np.sort(rng.uniform(0.15, 0.65, (R, T)), axis=1).astype(F32)
    bn = rng.standard_normal((R, S, 3)).astype(F32)
    bs = rng.uniform(0, 1, (A, R, S)).astype(F32)
    on, os_, binb = orc.bilinear_forward(bn, bs, z, tz)
    ggn = rng.standard_normal((R, T, 3)).astype(F32)
    ggs = rng.standard_normal((A, R, T)).astype(F32)
    rgn, rgs = orc.bilinear_backward(ggn, ggs, z, tz, S)
    out.update(bl_z=z, bl_tz=tz, bl_n=bn, bl_s=bs, bl_on=on, bl_os=os_, bl_inb=binb, bl_gn=ggn,
               bl_gs=ggs, bl_rgn=rgn, bl_rgs=rgs)
    # ---- lensless: 24 rays x 50 samples, K=48, offsets and uneven rays
    b = ts.rays(24, 50, 48, seed=77, with_offset=True, uneven=(10, 50), anneal=0.7)
    leaves = {k: getattr(b, k).double().clone().requires_grad_(True)
              for k in ("sdf", "grad", "refl", "inv_s", "power")}
    r = b.radar
    o = L.lensless_forward(leaves["sdf"], leaves["grad"], b.dists.double(),
                           L.points_f32(b.o, b.z, b.d), b.o.double(), b.d.double(),
                           leaves["inv_s"], leaves["power"], r["numADCSample"],
                           r["adcSampleRate"], r["As_sci"], r["fc_start"], ts.SPEED_OF_LIGHT,
                           ts.PATH_BIAS, refl=leaves["refl"], offset=b.offset.double(),
                           n_valid=b.n_valid.long(), anneal=b.anneal)
    gl = torch.tensor(rng.standard_normal(o.shape) + 1j * rng.standard_normal(o.shape))
    (o.real * gl.real + o.imag * gl.imag).sum().backward()
    out.update({f"ll_{k}": getattr(b, k).numpy() for k in
                ("sdf", "grad", "refl", "dists", "z", "o", "d", "offset", "n_valid", "inv_s",
                 "power")})
    out.update(ll_anneal=np.float32(b.anneal), ll_K=np.int64(r["numADCSample"]),
               ll_fs=np.float32(r["adcSampleRate"]), ll_gout=gl.numpy(),
               ll_out=o.detach().numpy(),
               **{f"ll_g_{k}": v.grad.numpy() for k, v in leaves.items()})
    path = os.path.join(HERE, "rf_golden.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


if __name__ == "__main__":
    main()
