"""The stage renderers' per-step RF operations on the GPU against the oracle:
the sphere hits and the transmittance offset of the calibration
(rf_rendering.py:455-495,611-626), and the matched-filter prediction
(rf_rendering.py:686-716)."""
import math

import numpy as np
import pytest
import torch

from oracle import lensless as L
from oracle import rf as orc
from rf_workloads import mf_scenes

pytestmark = pytest.mark.gpu


def test_intersect_sphere_kernel_branches(cuda):
    """The closed-form cases of each branch (outside, inside, miss, sphere
    behind, centre on the sphere) through the kernel."""
    from geraf.models.rendering.rf_stage_ops import intersect_sphere

    cases = [((0, 0, -3.0), (0, 0, 0.0), True, (0, 0, -1.0)),
             ((0, 0, 0.5), (0, 0, 0.9), True, (0, 0, 1.0)),
             ((0, 0, -3.0), (2, 0, -2.0), False, None),
             ((0, 0, -3.0), (0, 0, -4.0), False, None),
             ((0, 0, -1.0), (0, 0, 0.0), True, (0, 0, 1.0))]
    for c, p, ok, want in cases:
        cen = torch.tensor([c], device=cuda)
        pts = torch.tensor([[p]], device=cuda)
        valid, hit = intersect_sphere(cen, pts)
        assert bool(valid[0, 0, 0]) == ok, (c, p)
        assert torch.all(torch.isfinite(hit))
        if ok:
            np.testing.assert_allclose(hit[0, 0, 0].cpu().numpy(), want, rtol=0, atol=1e-6)
        ov, oh = L.intersect_sphere(np.array([c]), np.array([[p]]))
        assert bool(ov[0, 0, 0]) == ok


def _intersect_f32(c, p):
    """The reference formula (rf_rendering.py:455-495) in float32 numpy, one
    rounding per torch operation, for a given centre c [3] (float32)."""
    f = np.float32
    v = p - c
    ln = np.sqrt((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2])
    d = v / (ln + f(1e-8))[..., None]
    b = f(2) * ((c[0] * d[..., 0] + c[1] * d[..., 1]) + c[2] * d[..., 2])
    cc = ((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]) - f(1)
    disc = b * b - f(4) * cc
    sq = np.sqrt(np.maximum(disc, f(0)))
    t1 = (-b - sq) / f(2)
    t2 = (-b + sq) / f(2)
    has = (t1 > f(1e-5)) | (t2 > f(1e-5))
    t = np.where(t1 > f(1e-5), t1, t2)
    valid = (disc >= 0) & has
    hit = c + np.where(has, t, f(0))[..., None] * d
    return valid, hit, disc, t1, t2, b, sq


def test_intersect_sphere_kernel_vs_oracle(cuda):
    """Random antenna sets and samples (an aperture in front of the unit
    sphere of the SDF frame, samples everywhere around it): the kernel equals
    the reference formula evaluated in float32 (flags exactly, hits to a few
    ulp), and the FP64 oracle up to that formula's own FP32 conditioning
    (grazing rays: the rounding of the discriminant b^2 - 4 (|c|^2 - 1), about
    eps (b^2 + 4 |c|^2 + 4), divided by 2 sqrt(disc))."""
    from geraf.models.rendering.rf_stage_ops import intersect_sphere

    rng = np.random.default_rng(0)
    eps32 = float(np.finfo(np.float32).eps)
    for trial in range(4):
        ant = rng.uniform(-0.3, 0.3, (50, 3)) + [0.0, 0.0, -1.0 - 0.5 * trial]
        pts = rng.uniform(-1.5, 1.5, (300, 40, 3))
        ant32, pts32 = ant.astype(np.float32), pts.astype(np.float32)
        ant_t = torch.tensor(ant32, device=cuda)
        valid, hit = intersect_sphere(ant_t, torch.tensor(pts32, device=cuda))
        vg, hg = valid[0].cpu().numpy(), hit[0].cpu().numpy()
        c32 = ant_t.mean(0).cpu().numpy()  # the centre the op used
        v32, h32, disc, t1, t2, b, sq = _intersect_f32(c32, pts32)
        edge = (np.abs(disc) < 1e-6) | (np.abs(t1 - 1e-5) < 1e-6) | (np.abs(t2 - 1e-5) < 1e-6)
        assert np.array_equal(vg[~edge], v32[~edge])
        ulp = np.max(np.abs(hg - h32)[vg & v32] / np.maximum(np.abs(h32[vg & v32]), 1e-3))
        ov, oh = L.intersect_sphere(c32[None].astype(np.float64), pts32.astype(np.float64))
        safe = ~edge & (np.abs(disc) > 1e-4)
        assert np.array_equal(vg[safe], ov[0][safe])
        both = vg & ov[0] & safe
        cond = 4 * eps32 * ((b * b + 4.0 * (float(c32 @ c32) + 1.0)) / np.maximum(sq, 1e-30)
                            + np.abs(b) + 1.0)
        err = np.linalg.norm(hg - oh[0], axis=-1)
        print(f"trial {trial}: {both.sum()} hits, {(~vg & safe).sum()} misses; vs float32 "
              f"formula {ulp / eps32:.1f} ulp; vs FP64 max err {err[both].max():.2e}, max "
              f"err / conditioning {np.max(err[both] / cond[both]):.2f}", flush=True)
        assert both.sum() > 1000 and ulp <= 4 * eps32
        assert np.all(err[both] <= cond[both] + 1e-6)


def test_transmission_offset_vs_oracle(cuda):
    """The offset kernel on its own, and transmission_offset with an analytic
    SDF (sphere of radius 0.5 plus a ripple) standing in for the network,
    against the oracle's composition."""
    from geraf.models.rendering.rf_stage_ops import intersect_sphere, transmission_offset
    from geraf.ops import _native

    rng = np.random.default_rng(1)
    R, S = 257, 33
    sdf_hit = rng.uniform(-0.1, 0.1, (R, S)).astype(np.float32)
    sdf_hit[0, 0] = np.nan  # never read where not valid
    sdf_first = rng.uniform(-0.1, 0.1, R).astype(np.float32)
    valid = rng.uniform(0, 1, (R, S)) > 0.3
    valid[0, 0] = False
    inv_s = np.float32(64.0)
    T = lambda x, dt=None: torch.tensor(x, device=cuda, dtype=dt)  # noqa: E731
    off = torch.empty((R, S), device=cuda)
    # the operands stay referenced until the kernel has run
    ins = (T(sdf_hit), T(sdf_first), T(valid, torch.bool), T(np.array([inv_s])))
    _native.call("geraf_calibration_offset_f32", *map(_native.ptr, ins), _native.ptr(off), R, S,
                 _native.stream_of(off))
    torch.cuda.synchronize()
    ref = L.transmission_offset(sdf_hit, sdf_first, valid, float(inv_s))
    got = off.cpu().numpy()
    assert np.all(got[~valid] == 0.0)
    assert np.max(np.abs(got - ref)) <= 1e-6

    def sdf_fn(x):
        r = x.norm(dim=-1)
        return r - 0.5 + 0.02 * torch.sin(5 * x[..., 0]) * torch.cos(3 * x[..., 1])

    ant = torch.tensor(rng.uniform(-0.3, 0.3, (40, 3)) + [0, 0, -1.4], dtype=torch.float32,
                       device=cuda)
    pts = torch.tensor(rng.uniform(-0.9, 0.9, (R, S, 3)), dtype=torch.float32, device=cuda)
    inv = torch.tensor(50.0, device=cuda)
    got = transmission_offset(sdf_fn, inv, pts, ant).cpu().numpy()
    valid, hit = intersect_sphere(ant, pts)
    sh = sdf_fn(hit[0]).double().cpu().numpy()
    sf = sdf_fn(pts[:, 0, :]).double().cpu().numpy()
    ref = L.transmission_offset(sh, sf, valid[0].cpu().numpy(), 50.0)
    print(f"transmission offset: {valid.float().mean():.3f} valid, max err "
          f"{np.max(np.abs(got - ref)):.2e}", flush=True)
    assert np.max(np.abs(got - ref)) <= 1e-6 and np.any(got != 0)


def test_matched_filter_prediction_vs_oracle(cuda):
    """pred = |MF| of the mixed echo at the in-bound targets and gt =
    gt_sar_value[inbounds] / sub_adc (rf_rendering.py:686-716), against the
    CPU oracle; the all-points form (inbounds=None) agrees with a full mask."""
    from geraf.models.rendering.rf_stage_ops import matched_filter_prediction

    rng = np.random.default_rng(2)
    A, K, N = 24, 64, 700
    tx = rng.uniform(-0.05, 0.05, (A, 3)).astype(np.float32)
    tx[:, 2] = 0
    rx = (tx + np.float32([0.001, 0, 0])).astype(np.float32)
    s = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    pts = (rng.uniform(-0.1, 0.1, (N, 3)) + [0, 0, 0.3]).astype(np.float32)
    gtv = rng.uniform(0, 2, N).astype(np.float32)
    inb = rng.uniform(0, 1, N) > 0.25
    radar = dict(numADCSample=K, adcSampleRate=1.25e6, As_sci=701500.0, sub_adc=8)
    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    pred, gt = matched_filter_prediction(T(s.real), T(s.imag), T(tx), T(rx), T(pts), T(gtv),
                                         T(inb), radar)
    n = int(inb.sum())
    assert pred.shape == (n, 1) and gt.shape == (n, 1)
    prm = dict(f0=mf_scenes.MF_F0, slope=radar["As_sci"], fs=radar["adcSampleRate"],
               c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)
    ref = np.abs(orc.mf_forward(s.real, s.imag, tx, rx, pts[inb], **prm))
    err = float(np.linalg.norm(pred[:, 0].double().cpu().numpy() - ref) / np.linalg.norm(ref))
    assert err <= 1e-5, err
    np.testing.assert_array_equal(gt[:, 0].cpu().numpy(), gtv[inb] / np.float32(8))
    # without sub_adc the ground truth is passed through
    _, gt2 = matched_filter_prediction(T(s.real), T(s.imag), T(tx), T(rx), T(pts), T(gtv),
                                       T(inb), {k: v for k, v in radar.items() if k != "sub_adc"})
    np.testing.assert_array_equal(gt2[:, 0].cpu().numpy(), gtv[inb])
    pa, ga = matched_filter_prediction(T(s.real), T(s.imag), T(tx), T(rx), T(pts), T(gtv), None,
                                       radar)
    pm, gm = matched_filter_prediction(T(s.real), T(s.imag), T(tx), T(rx), T(pts), T(gtv),
                                       T(np.ones(N, bool)), radar)
    assert torch.equal(pa, pm) and torch.equal(ga, gm)
