"""Pinning the CPU oracle (no GPU): closed forms, special cases, adjoint
identities, finite differences and cross-checks between independent
restatements.  The reference ships no tests or golden vectors, so these are
what fixes the oracle."""
import math

import numpy as np
import pytest
import torch

from oracle import lensless as L
from oracle import rf as orc
from rf_workloads import mf_scenes, tracer_scenes as ts

F0, C, BIAS = mf_scenes.MF_F0, mf_scenes.SPEED_OF_LIGHT, mf_scenes.PATH_BIAS
FS, AS = 1.25e6, 701500.0
PRM = dict(f0=F0, slope=AS, fs=FS, c=C, bias=BIAS)


def phase(tx, rx, p, k):
    tau = (np.linalg.norm(tx - p) + np.linalg.norm(rx - p) + BIAS) / C
    return 2 * np.pi * (F0 + AS * k / FS) * tau


def test_mf_single_scatterer_focuses_exactly():
    # Echo of a point at voxel v_s synthesised in the matched convention:
    # out(v_s) = K * mean_a(amp_a), phase 0 (closed form).
    rng = np.random.default_rng(0)
    A, K = 9, 16
    tx = rng.uniform(-0.02, 0.02, (A, 3))
    tx[:, 2] = 0
    rx = tx + np.array([0.001, 0, 0])
    vs = np.array([0.003, -0.002, 0.25])
    amp = rng.uniform(0.5, 1.5, A)
    k = np.arange(K)
    s = np.stack([amp[a] * np.exp(1j * phase(tx[a], rx[a], vs, k)) for a in range(A)])
    out = orc.mf_forward(s.real, s.imag, tx, rx, vs[None], **PRM)
    assert abs(out[0] - K * amp.mean()) <= 1e-12 * K


def test_mf_one_term_is_one_complex_multiply():
    tx = np.array([[0.01, 0.02, 0.0]])
    rx = np.array([[0.012, 0.02, 0.0]])
    v = np.array([[0.0, 0.01, 0.3]])
    s = np.array([[0.7 - 0.2j]])
    out = orc.mf_forward(s.real, s.imag, tx, rx, v, **PRM)
    assert abs(out[0] - s[0, 0] * np.exp(-1j * phase(tx[0], rx[0], v[0], 0))) < 1e-14


def test_mf_linearity_and_tx_rx_symmetry():
    rng = np.random.default_rng(1)
    A, K, N = 5, 12, 7
    tx = rng.uniform(-0.03, 0.03, (A, 3))
    rx = rng.uniform(-0.03, 0.03, (A, 3))
    v = rng.uniform(-0.05, 0.05, (N, 3)) + [0, 0, 0.3]
    s1 = rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))
    s2 = rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))
    f = lambda s, t=tx, r=rx: orc.mf_forward(s.real, s.imag, t, r, v, **PRM)  # noqa: E731
    a, b = 0.3 - 1.1j, 2.0
    np.testing.assert_allclose(f(a * s1 + b * s2), a * f(s1) + b * f(s2), rtol=0, atol=1e-12)
    np.testing.assert_allclose(f(s1, rx, tx), f(s1), rtol=0, atol=1e-13)


def test_mf_adjoint_identity_fp64():
    rng = np.random.default_rng(2)
    A, K, N = 4, 20, 30
    tx = rng.uniform(-0.03, 0.03, (A, 3))
    rx = tx + 0.001
    v = rng.uniform(-0.05, 0.05, (N, 3)) + [0, 0, 0.3]
    s = rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))
    g = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    fwd = orc.mf_forward(s.real, s.imag, tx, rx, v, **PRM)
    adj = orc.mf_backward(g, tx, rx, v, K, norm=A, **PRM)
    lhs = np.sum(fwd.real * g.real + fwd.imag * g.imag)
    rhs = np.sum(s.real * adj.real + s.imag * adj.imag)
    assert abs(lhs - rhs) <= 1e-12 * max(1.0, abs(lhs))


def test_mf_ref32_close_to_fp64():
    sc = mf_scenes.config1(0)
    sel = np.arange(0, sc.N, 997)
    out = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels[sel], **sc.params())
    o32 = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels[sel], form="ref32",
                         **sc.params())
    e = np.linalg.norm(o32 - out) / np.linalg.norm(out)
    assert 1e-7 < e < 1e-3  # FP32 per-term phase error of the reference formula


def _st(normals, sig, pts, tx, rx, inb, K=1):
    return orc.st_forward(normals, sig, pts, tx, rx, inb, K, **PRM)


def test_tracer_specular_bisector_and_normal_incidence():
    tx = np.array([[0.0, 0.0, 0.0]])
    rx = np.array([[0.02, 0.0, 0.0]])
    p = np.array([[0.01, 0.0, 0.3]])
    # normal along the bisector of the two directions back to the antennas -> spec = 1
    u1 = (tx[0] - p[0]) / np.linalg.norm(tx[0] - p[0])
    u2 = (rx[0] - p[0]) / np.linalg.norm(rx[0] - p[0])
    n = ((u1 + u2) / np.linalg.norm(u1 + u2))[None]
    out = _st(n, np.array([[2.0]]), p, tx, rx, [True], K=3)
    d = np.linalg.norm(tx[0] - p[0]) + np.linalg.norm(rx[0] - p[0])
    k = np.arange(3)
    expect = 2.0 / d ** 2 * np.exp(1j * phase(tx[0], rx[0], p[0], k))
    np.testing.assert_allclose(out[0], expect, rtol=2e-6)
    # monostatic normal incidence: amp = sigma / (2R)^2
    R = 0.37
    out = _st(np.array([[0.0, 0.0, -1.0]]), np.array([[1.5]]), np.array([[0.0, 0.0, R]]),
              np.zeros((1, 3)), np.zeros((1, 3)), [True])
    # (exact up to the 1e-7 of the reference's v/(|v|+1e-7) normalisations)
    assert abs(abs(out[0, 0]) / (1.5 / (2 * R) ** 2) - 1.0) < 2e-6


def test_tracer_culling_and_inbounds():
    tx = rx = np.zeros((1, 3))
    p = np.array([[0.0, 0.0, 0.3]])
    # facing away (inc . n > 0) -> culled
    assert _st(np.array([[0.0, 0.0, 1.0]]), np.ones((1, 1)), p, tx, rx, [True])[0, 0] == 0
    # grazing (spec < 1e-6) -> culled
    assert _st(np.array([[1.0, 0.0, -1e-4]]), np.ones((1, 1)), p, tx, rx, [True])[0, 0] == 0
    # out of bounds -> nothing, and P still divides
    n = np.array([[0.0, 0.0, -1.0], [0.0, 0.0, -1.0]])
    pp = np.array([[0.0, 0.0, 0.3], [0.0, 0.0, 0.3]])
    one = _st(n, np.ones((1, 2)), pp, tx, rx, [True, False])
    both = _st(n, np.ones((1, 2)), pp, tx, rx, [True, True])
    np.testing.assert_allclose(2 * one, both, rtol=1e-14)


def test_tracer_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    A, P, K = 3, 5, 16
    tx = rng.uniform(-0.03, 0.03, (A, 3))
    tx[:, 2] = 0
    rx = tx + [0.002, 0, 0]
    pts = rng.uniform(-0.03, 0.03, (P, 3)) + [0, 0, 0.3]
    n = -pts / np.linalg.norm(pts, axis=1, keepdims=True) + 0.2 * rng.standard_normal((P, 3))
    sig = rng.uniform(0.2, 1, (A, P))
    inb = np.array([True, True, False, True, True])
    g = rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))
    loss = lambda nn, ss: np.sum(  # noqa: E731
        (lambda o: o.real * g.real + o.imag * g.imag)(
            orc.st_forward(nn, ss, pts, tx, rx, inb, K, **PRM)))
    gn, gs = orc.st_backward(g, n, sig, pts, tx, rx, inb, **PRM)
    h = 1e-6
    for q in range(P):
        for c in range(3):
            e = np.zeros_like(n)
            e[q, c] = h
            fd = (loss(n + e, sig) - loss(n - e, sig)) / (2 * h)
            # The reference's normal Jacobian, (I - x x^T)/(|v| + 1e-7) with the
            # eps-normalised x (SignalTracer_cuda_kernel.cu:108-155), is the
            # derivative of v/(|v| + 1e-7) only up to O(1e-7/|v|), amplified by
            # the projection's cancellation: a few 1e-6 relative.
            assert abs(fd - gn[q, c]) <= 1e-5 * max(1.0, abs(fd)), (q, c, fd, gn[q, c])
    for a in range(A):
        for q in range(P):
            e = np.zeros_like(sig)
            e[a, q] = h
            fd = (loss(n, sig + e) - loss(n, sig - e)) / (2 * h)
            assert abs(fd - gs[a, q]) <= 1e-7 * max(1.0, abs(fd)) + 1e-9
    assert np.all(gn[2] == 0) and np.all(gs[:, 2] == 0)


def _bilinear_numpy(nrm, sig, z, tz):
    A, R, S = sig.shape
    T = tz.shape[1]
    on = np.zeros((R, T, 3))
    os_ = np.zeros((A, R, T))
    inb = np.zeros((R, T), bool)
    for r in range(R):
        for t in range(T):
            x = tz[r, t]
            if x < z[r, 0] or x > z[r, -1]:
                continue
            lo = min(max(np.searchsorted(z[r], x, side="right") - 1, 0), S - 2)
            w = (x - z[r, lo]) / max(z[r, lo + 1] - z[r, lo], 1e-6)
            on[r, t] = (1 - w) * nrm[r, lo] + w * nrm[r, lo + 1]
            os_[:, r, t] = (1 - w) * sig[:, r, lo] + w * sig[:, r, lo + 1]
            inb[r, t] = True
    return on, os_, inb


def test_bilinear_brute_force_identity_and_transpose():
    rng = np.random.default_rng(4)
    A, R, S, T = 2, 6, 9, 13
    z = np.sort(rng.uniform(0, 1, (R, S)), axis=1)
    tz = np.sort(rng.uniform(-0.1, 1.1, (R, T)), axis=1)
    nrm = rng.standard_normal((R, S, 3))
    sig = rng.standard_normal((A, R, S))
    on, os_, inb = orc.bilinear_forward(nrm, sig, z, tz)
    bn, bs, binb = _bilinear_numpy(nrm, sig, z, tz)
    assert np.array_equal(inb, binb)
    np.testing.assert_allclose(on, bn, atol=1e-14)
    np.testing.assert_allclose(os_, bs, atol=1e-14)
    # identity when the targets are the source depths
    on, os_, inb = orc.bilinear_forward(nrm, sig, z, z)
    assert inb.all()
    np.testing.assert_allclose(on, nrm, atol=1e-14)
    np.testing.assert_allclose(os_, sig, atol=1e-14)
    # backward is the exact transpose
    on, os_, _ = orc.bilinear_forward(nrm, sig, z, tz)
    gn = rng.standard_normal(on.shape)
    gs = rng.standard_normal(os_.shape)
    rn, rs = orc.bilinear_backward(gn, gs, z, tz, S)
    lhs = np.sum(on * gn) + np.sum(os_ * gs)
    rhs = np.sum(nrm * rn) + np.sum(sig * rs)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def _lens_inputs(b):
    f64 = lambda t: None if t is None else t.double()  # noqa: E731
    return dict(dists=f64(b.dists), pts=L.points_f32(b.o, b.z, b.d), o=f64(b.o), d=f64(b.d),
                offset=f64(b.offset),
                n_valid=None if b.n_valid is None else b.n_valid.long(), anneal=b.anneal)


def _radar_kw(b):
    r = b.radar
    return dict(K=r["numADCSample"], fs=r["adcSampleRate"], As_sci=r["As_sci"], f0=r["fc_start"],
                c=ts.SPEED_OF_LIGHT, bias=ts.PATH_BIAS)


def test_lensless_oracle_equals_ap_tracer_per_ray():
    # Two independent restatements: the torch glue + per-ray signal against the
    # C [A,P] tracer run ray by ray with tx = rx = origin on the same sigmas.
    b = ts.rays(5, 40, 24, seed=8)
    inp = _lens_inputs(b)
    out, parts = L.lensless_forward(b.sdf.double(), b.grad.double(), inp["dists"], inp["pts"],
                                    inp["o"], inp["d"], b.inv_s.double(), b.power.double(),
                                    refl=b.refl.double(), return_parts=True, **_radar_kw(b))
    nrm = torch.nn.functional.normalize(b.grad.double(), dim=-1).numpy()
    r = b.radar
    for i in range(b.R):
        o = b.o[i].double().numpy()[None]
        ref = orc.st_forward(nrm[i], parts["sigma"][i].detach().numpy()[None],
                             inp["pts"][i].numpy(), o, o, np.ones(b.S, bool), r["numADCSample"],
                             r["fc_start"], r["As_sci"], r["adcSampleRate"], ts.SPEED_OF_LIGHT,
                             ts.PATH_BIAS)
        np.testing.assert_allclose(out[i].detach().numpy(), ref[0], rtol=1e-9,
                                   atol=1e-12 * np.abs(ref).max())


def test_lensless_transmittance_is_sequential_product():
    b = ts.rays(3, 17, 8, seed=9)
    inp = _lens_inputs(b)
    _, parts = L.lensless_forward(b.sdf.double(), b.grad.double(), inp["dists"], inp["pts"],
                                  inp["o"], inp["d"], b.inv_s.double(), b.power.double(),
                                  return_parts=True, **_radar_kw(b))
    a = parts["alpha"].detach().numpy()
    for r in range(3):
        t = 1.0
        for s in range(17):
            assert abs(parts["trans"][r, s].item() - t) <= 1e-15
            t *= 1.0 - a[r, s] + 1e-7


def test_lensless_alpha_and_slope_vs_high_precision():
    """alpha and d alpha / d sdf of the oracle against a 60-digit evaluation
    of the reference formula, from behind the surface (sdf * inv_s = -30) to
    far in front of it (+35), where both cdfs lie within 1e-15 of 1.  The
    literal float64 difference of the two cdfs (and autograd's y (1 - y)
    slope) is off by 1e-3 relative there, and fails this check."""
    import decimal
    from decimal import Decimal as Dec

    decimal.getcontext().prec = 60
    inv_s, dist = 64.0, 0.012
    sdf = torch.linspace(-30.0 / inv_s, 35.0 / inv_s, 53, dtype=torch.float64)
    S = sdf.numel()
    o = torch.zeros(1, 3, dtype=torch.float64)
    d = torch.tensor([[0.0, 0.0, 1.0]], dtype=torch.float64)
    grad = torch.tensor([0.0, 0.0, -1.0], dtype=torch.float64).expand(1, S, 3).clone()
    leaf = sdf[None].clone().requires_grad_(True)
    pts = torch.zeros(1, S, 3, dtype=torch.float64)
    pts[0, :, 2] = torch.linspace(1.0, 2.0, S, dtype=torch.float64)
    dists = torch.full((1, S), dist, dtype=torch.float64)
    _, parts = L.lensless_forward(leaf, grad, dists, pts, o, d, torch.tensor(inv_s),
                                  torch.tensor(1.0), 1, 1.0, 0.0, 1.0, 1.0, 0.0, anneal=1.0,
                                  return_parts=True)
    alpha = parts["alpha"]
    (galpha,) = torch.autograd.grad(alpha.sum(), leaf)
    alpha = alpha.detach()

    one, eps = Dec(1), Dec(L.EPS)

    def sig(t):
        return one / (one + (-t).exp())

    def slope(t):
        e = (-t).exp()
        return e / ((one + e) * (one + e))

    for s in range(S):
        # cos = -1 and anneal = 1: iter_cos = -1, prev = sdf + dist/2, next = sdf - dist/2
        tp = (Dec(float(sdf[s])) + Dec(dist) / 2) * Dec(inv_s)
        tn = (Dec(float(sdf[s])) - Dec(dist) / 2) * Dec(inv_s)
        p, n = sig(tp), sig(tn)
        den = p + eps
        a_ref = (p - n + eps) / den
        g_ref = Dec(inv_s) * (slope(tp) * n / (den * den) - slope(tn) / den)
        assert abs(float(alpha[0, s]) - float(a_ref)) <= 1e-14 * float(a_ref), s
        assert abs(float(galpha[0, s]) - float(g_ref)) <= 1e-12 * abs(float(g_ref)), s


@pytest.mark.parametrize("seed,uneven,offset", [(10, None, False), (11, (4, 40), True)])
def test_lensless_oracle_gradients_vs_finite_differences(seed, uneven, offset):
    b = ts.rays(6, 40, 32, seed=seed, uneven=uneven, with_offset=offset)
    inp = _lens_inputs(b)
    kw = _radar_kw(b)
    base = {k: getattr(b, k).double() for k in ("sdf", "grad", "refl", "inv_s", "power")}
    rng = np.random.default_rng(seed)
    gK = torch.tensor(rng.standard_normal((b.R, kw["K"])) + 1j * rng.standard_normal((b.R, kw["K"])))

    def loss(vals):
        o = L.lensless_forward(vals["sdf"], vals["grad"], inp["dists"], inp["pts"], inp["o"],
                               inp["d"], vals["inv_s"], vals["power"], refl=vals["refl"],
                               offset=inp["offset"], n_valid=inp["n_valid"], anneal=b.anneal,
                               **kw)
        return (o.real * gK.real + o.imag * gK.imag).sum()

    leaves = {k: v.clone().requires_grad_(True) for k, v in base.items()}
    loss(leaves).backward()
    h = 1e-4
    parts = L.masks(base["sdf"], base["grad"], inp["dists"], inp["pts"], inp["o"], inp["d"],
                    base["inv_s"], base["power"], refl=base["refl"], offset=inp["offset"],
                    n_valid=inp["n_valid"])
    checked = 0
    for name in ("inv_s", "power", "sdf", "refl", "grad"):
        flat = base[name].reshape(-1)
        picks = range(flat.numel()) if flat.numel() == 1 else rng.choice(flat.numel(), 12, False)
        for i in picks:
            def shifted(dv):
                v = dict(base)
                t = flat.clone()
                t[i] += dv
                v[name] = t.reshape(base[name].shape)
                return v
            # skip picks whose +-h step crosses a branch (culling, relu, clamp)
            mp = L.masks(shifted(h)["sdf"], shifted(h)["grad"], inp["dists"], inp["pts"],
                         inp["o"], inp["d"], shifted(h)["inv_s"], base["power"],
                         refl=base["refl"], offset=inp["offset"], n_valid=inp["n_valid"])
            mm = L.masks(shifted(-h)["sdf"], shifted(-h)["grad"], inp["dists"], inp["pts"],
                         inp["o"], inp["d"], shifted(-h)["inv_s"], base["power"],
                         refl=base["refl"], offset=inp["offset"], n_valid=inp["n_valid"])
            if not torch.equal(mp["live"], mm["live"]) or not torch.equal(
                    mp["true_cos"] < 0, mm["true_cos"] < 0):
                continue
            fd = (loss(shifted(h)) - loss(shifted(-h))).item() / (2 * h)
            ad = leaves[name].grad.reshape(-1)[i].item()
            scale = max(1e-3 * leaves[name].grad.abs().max().item(), abs(fd))
            assert abs(fd - ad) <= 1e-4 * scale, (name, i, fd, ad)
            checked += 1
    assert checked >= 30


def test_range_bin_mf_known_answer_and_adjoint():
    rng = np.random.default_rng(12)
    A, R, N = 6, 64, 40
    tx = rng.uniform(-0.02, 0.02, (A, 3))
    tx[:, 2] = 0
    rx = tx + [0.001, 0, 0]
    v = rng.uniform(-0.05, 0.05, (N, 3)) + [0, 0, 0.3]
    spacing = 0.02  # bins of 4 cm of round-trip path; cutoff spacing*R = 1.28 m
    prm = dict(f0=F0, c=C, bias=BIAS)
    # echo holding exp(+j*2*pi*f0*d/c) in voxel 0's bin focuses voxel 0 to exactly 1
    s = np.zeros((A, R), complex)
    for a in range(A):
        d = np.linalg.norm(tx[a] - v[0]) + np.linalg.norm(rx[a] - v[0]) + BIAS
        s[a, int(np.floor(d / spacing / 2))] = np.exp(2j * np.pi * F0 * d / C)
    out = orc.mf_fft_forward(s.real, s.imag, tx, rx, v, spacing, **prm)
    assert abs(out[0] - 1.0) < 1e-12
    # adjoint identity
    s = rng.standard_normal((A, R)) + 1j * rng.standard_normal((A, R))
    g = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    fwd = orc.mf_fft_forward(s.real, s.imag, tx, rx, v, spacing, **prm)
    adj = orc.mf_fft_backward(g, tx, rx, v, R, spacing, A, **prm)
    lhs = np.sum(fwd.real * g.real + fwd.imag * g.imag)
    rhs = np.sum(s.real * adj.real + s.imag * adj.imag)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
    # pairs beyond spacing * R are skipped
    far = orc.mf_fft_forward(s.real, s.imag, tx, rx, v + [0, 0, 5.0], spacing, **prm)
    assert np.all(far == 0)


def test_stage_step_torch_form_matches_oracle():
    # the all-FP32 autograd form of the training-step benchmark restates the same
    # composition; in float64 it must equal the oracle
    from rf_workloads.stage_step import torch_lensless

    b = ts.rays(9, 30, 16, seed=13, uneven=(5, 30))
    r = b.radar
    f64 = lambda t: t.double()  # noqa: E731
    pts = L.points_f32(b.o, b.z, b.d)
    ref = L.lensless_forward(f64(b.sdf), f64(b.grad), f64(b.dists), pts, f64(b.o), f64(b.d),
                             f64(b.inv_s), f64(b.power), r["numADCSample"], r["adcSampleRate"],
                             r["As_sci"], r["fc_start"], ts.SPEED_OF_LIGHT, ts.PATH_BIAS,
                             refl=f64(b.refl), n_valid=b.n_valid.long())
    o_r, o_i = torch_lensless(f64(b.sdf), f64(b.grad), f64(b.dists), f64(b.z), f64(b.o),
                              f64(b.d), f64(b.inv_s), f64(b.power), f64(b.refl),
                              b.n_valid.double(), r["numADCSample"], r["adcSampleRate"],
                              r["As_sci"], r["fc_start"])
    out = (o_r + 1j * o_i).numpy()
    # points are formed in float64 here and in float32 by the oracle
    assert np.linalg.norm(out - ref.numpy()) <= 1e-4 * np.linalg.norm(ref.numpy())


# ------------------------------------------------ transmittance calibration (sphere hits)

def test_intersect_sphere_branches_closed_form():
    """Each branch of t = t1 > 1e-5 ? t1 : t2 > 1e-5 ? t2 : inf
    (rf_rendering.py:484-490) on rays whose roots are known: from outside the
    unit sphere toward it (t1 = 2), from inside (t1 < 0, t2 = 1/2), missing it
    (negative discriminant), with the sphere behind the centre (both roots
    negative), and t1 just below the 1e-5 cut (t2 taken instead)."""
    # centre (0, 0, -3): towards +z the roots of t^2 - 6t + 8 are 2 and 4
    valid, hit = L.intersect_sphere(np.array([[0, 0, -3.0]]), np.array([[[0, 0, 0.0]]]))
    scale = 3.0 / (3.0 + L.DIR_EPS)  # dir = v / (|v| + 1e-8) is not quite unit
    b = 2.0 * -3.0 * scale
    t1 = (-b - math.sqrt(b * b - 32.0)) / 2.0
    assert valid[0, 0, 0] and abs(t1 - 2.0) < 1e-7
    np.testing.assert_allclose(hit[0, 0, 0], [0, 0, -3.0 + t1 * scale], rtol=0, atol=1e-15)
    # centre inside the sphere: t1 = -1.5 < 0, t2 = 0.5 -> the far root
    valid, hit = L.intersect_sphere(np.array([[0, 0, 0.5]]), np.array([[[0, 0, 0.9]]]))
    assert valid[0, 0, 0] and abs(hit[0, 0, 0, 2] - 1.0) < 1e-7
    # misses: discriminant < 0; the reference's hit there is still c + t1 dir
    valid, hit = L.intersect_sphere(np.array([[0, 0, -3.0]]), np.array([[[2, 0, -2.0]]]))
    assert not valid[0, 0, 0] and np.all(np.isfinite(hit))
    # sphere behind: roots -4 and -2 -> no t, invalid, hit not finite
    valid, hit = L.intersect_sphere(np.array([[0, 0, -3.0]]), np.array([[[0, 0, -4.0]]]))
    assert not valid[0, 0, 0] and not np.all(np.isfinite(hit))
    # centre on the sphere (t1 = 0 <= 1e-5): the other root, across the sphere
    valid, hit = L.intersect_sphere(np.array([[0, 0, -1.0]]), np.array([[[0, 0, 0.0]]]))
    assert valid[0, 0, 0] and abs(hit[0, 0, 0, 2] - 1.0) < 1e-7


def test_intersect_sphere_hits_lie_on_the_sphere_nearest_positive_root():
    """Brute force: for random centres and samples, a valid hit lies on the
    sphere, on the ray, and no smaller t > 1e-5 along the ray reaches the
    sphere (a fine scan of |c + t dir| finds the first crossing at the hit)."""
    rng = np.random.default_rng(3)
    for _ in range(20):
        c = rng.uniform(-2, 2, (5, 3))
        pts = rng.uniform(-1.2, 1.2, (4, 6, 3))
        valid, hit = L.intersect_sphere(c, pts)
        cen = c.mean(0)
        for r, s in zip(*np.nonzero(valid[0])):
            h = hit[0, r, s]
            v = pts[r, s] - cen
            d = v / (np.linalg.norm(v) + L.DIR_EPS)
            t = np.dot(h - cen, d) / np.dot(d, d)
            assert t > 1e-5 and np.linalg.norm(cen + t * d - h) < 1e-9
            # on the sphere up to the 1e-8 in the direction's normalisation
            # (the roots take the direction as a unit vector)
            assert abs(np.linalg.norm(h) - 1.0) <= 1e-9 + 4.0 * t * (1.0 - np.linalg.norm(d))
            ts_ = np.linspace(1e-5, t, 2000)[:-1]
            rad = np.linalg.norm(cen[None] + ts_[:, None] * d[None], axis=1)
            inside0 = np.linalg.norm(cen) < 1.0
            # no earlier crossing: the side of the sphere does not change before t
            assert np.all((rad < 1.0) == inside0) or np.all(np.abs(rad - 1.0) < 1e-3)


def test_transmission_offset_closed_form():
    sig = lambda x: 1.0 / (1.0 + math.exp(-x))  # noqa: E731
    sdf_hit = np.array([[0.1, -0.2, 5.0], [0.0, 0.3, -0.4]])
    sdf_first = np.array([0.05, -0.1])
    valid = np.array([[True, False, True], [False, True, True]])
    off = L.transmission_offset(sdf_hit, sdf_first, valid, 64.0)
    for r in range(2):
        for s in range(3):
            want = (sig(64 * sdf_hit[r, s]) - sig(64 * sdf_first[r])) if valid[r, s] else 0.0
            assert abs(off[r, s] - want) < 1e-15
    # invalid samples give exactly zero even where sdf_hit is not finite
    off = L.transmission_offset(np.array([[np.nan, 0.1]]), np.array([0.2]),
                                np.array([[False, True]]), 10.0)
    assert off[0, 0] == 0.0 and np.isfinite(off[0, 1])
