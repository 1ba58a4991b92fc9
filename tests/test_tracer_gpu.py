"""Signal-tracer parity on the GPU: [A,P] op, bilinear resampler and the fused
lensless tracer against the CPU oracle."""
import numpy as np
import pytest
import torch

from oracle import lensless as L
from oracle import rf as orc
from rf_workloads import tracer_scenes as ts

pytestmark = pytest.mark.gpu

TOL_FWD = 1e-5    # FP32 kernels vs FP64 oracle, rel-L2
TOL_GRAD = 1e-4   # FP32 adjoints vs FP64 oracle gradients, rel-L2
TOL_F64 = 1e-10

FC, AS, FS = ts.FC_START, ts.AS_SCI, ts.f32(1.25e6)
PRM = dict(f0=FC, slope=AS, fs=FS, c=ts.SPEED_OF_LIGHT, bias=ts.PATH_BIAS)


def rel(x, ref):
    """rel-L2 of x against ref; an all-zero reference requires an exactly zero x
    (and then counts as 0)."""
    x, ref = np.asarray(x), np.asarray(ref)
    den = np.linalg.norm(ref)
    if den == 0.0:
        assert not np.any(x), "reference is zero, result is not"
        return 0.0
    return float(np.linalg.norm(x - ref) / den)


def ap_case(A=6, P=1500, seed=0, frac_out=0.2):
    rng = np.random.default_rng(seed)
    tx = rng.uniform(-0.05, 0.05, (A, 3)).astype(np.float32)
    tx[:, 2] = 0
    rx = (tx + np.float32([0.002, 0.0, 0.0])).astype(np.float32)
    pts = rng.uniform(-0.1, 0.1, (P, 3)).astype(np.float32)
    pts[:, 2] += 0.3
    # normals pointing back toward the array, jittered: many live pairs
    n = -(pts - np.array([0, 0, 0])) / np.linalg.norm(pts, axis=1, keepdims=True)
    n = (n + 0.3 * rng.standard_normal((P, 3))).astype(np.float32)
    sig = rng.uniform(0.1, 1.0, (A, P)).astype(np.float32)
    inb = rng.uniform(0, 1, P) > frac_out
    return tx, rx, pts, n, sig, inb


def test_signal_tracer_forward_backward_vs_oracle(cuda):
    from geraf.ops.RFSignalTracer import signal_tracer_cuda

    A, P, K = 6, 1500, 64
    tx, rx, pts, n, sig, inb = ap_case(A, P)
    T = lambda x: torch.tensor(x, device=cuda)  # noqa: E731
    nt = T(n).requires_grad_(True)
    st = T(sig).requires_grad_(True)
    o_r, o_i = signal_tracer_cuda(nt, st, T(pts), T(tx), T(rx), T(inb), K, FS, FC, AS, A, P)
    out = o_r.detach().double().cpu().numpy() + 1j * o_i.detach().double().cpu().numpy()
    ref = orc.st_forward(n, sig, pts, tx, rx, inb, K, **PRM)
    err = rel(out, ref)
    print(f"[A,P] tracer fwd rel-L2 {err:.2e}")
    assert err <= TOL_FWD
    rng = np.random.default_rng(1)
    g = rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))
    (o_r * T(g.real.astype(np.float32)) + o_i * T(g.imag.astype(np.float32))).sum().backward()
    gn_ref, gs_ref = orc.st_backward(g.astype(np.complex64), n, sig, pts, tx, rx, inb, **PRM)
    assert rel(nt.grad.cpu().numpy(), gn_ref) <= TOL_FWD * 10
    assert rel(st.grad.cpu().numpy(), gs_ref) <= TOL_FWD * 10
    # out-of-bounds points get exact zeros
    assert torch.all(nt.grad[~T(inb)] == 0) and torch.all(st.grad[:, ~T(inb)] == 0)


def test_signal_tracer_f64(cuda):
    from geraf.ops.RFSignalTracer import signal_tracer_cuda

    A, P, K = 3, 400, 40
    tx, rx, pts, n, sig, inb = ap_case(A, P, seed=4)
    T = lambda x: torch.tensor(x, device=cuda, dtype=torch.float64)  # noqa: E731
    nt = T(n).requires_grad_(True)
    st = T(sig).requires_grad_(True)
    o_r, o_i = signal_tracer_cuda(nt, st, T(pts), T(tx), T(rx),
                                  torch.tensor(inb, device=cuda), K, FS, FC, AS, A, P)
    ref = orc.st_forward(n, sig, pts, tx, rx, inb, K, **PRM)
    assert rel(o_r.detach().cpu().numpy() + 1j * o_i.detach().cpu().numpy(), ref) <= TOL_F64
    g = np.ones((A, K)) + 0.5j * np.ones((A, K))
    (o_r * T(g.real) + o_i * T(g.imag)).sum().backward()
    gn_ref, gs_ref = orc.st_backward(g, n, sig, pts, tx, rx, inb, **PRM)
    assert rel(nt.grad.cpu().numpy(), gn_ref) <= 1e-9
    assert rel(st.grad.cpu().numpy(), gs_ref) <= 1e-9


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_bilinear_vs_oracle(cuda, dtype):
    from geraf.ops.RFSignalTracer import bilinear_raysample_cuda

    rng = np.random.default_rng(2)
    A, R, S, Tn = 2, 37, 32, 64
    z = np.sort(rng.uniform(0.2, 0.5, (R, S)), axis=1)
    tz = np.sort(rng.uniform(0.15, 0.55, (R, Tn)), axis=1)
    tz[0, :S] = z[0]  # exact hits on source depths
    nrm = rng.standard_normal((R, S, 3))
    sig = rng.uniform(0, 1, (A, R, S))
    T = lambda x: torch.tensor(x, device=cuda, dtype=dtype)  # noqa: E731
    nt = T(nrm).requires_grad_(True)
    st = T(sig).requires_grad_(True)
    on, os_, inb = bilinear_raysample_cuda(nt, st, T(z), T(tz), A, R, S, Tn)
    # the oracle sees the same (possibly float32-rounded) inputs
    cv = lambda t: t.detach().double().cpu().numpy()  # noqa: E731
    rn, rs, rinb = orc.bilinear_forward(cv(nt), cv(st), cv(T(z)), cv(T(tz)))
    tol = 1e-6 if dtype == torch.float32 else 1e-13
    assert np.array_equal(inb.cpu().numpy(), rinb)
    assert np.max(np.abs(cv(on) - rn)) <= tol
    assert np.max(np.abs(cv(os_) - rs)) <= tol
    gn = rng.standard_normal((R, Tn, 3))
    gs = rng.standard_normal((A, R, Tn))
    (on * T(gn)).sum().add((os_ * T(gs)).sum()).backward()
    rgn, rgs = orc.bilinear_backward(gn, gs, cv(T(z)), cv(T(tz)), S)
    assert np.max(np.abs(cv(nt.grad) - rgn)) <= tol * 10
    assert np.max(np.abs(cv(st.grad) - rgs)) <= tol * 10
    # deterministic: same result twice, bit for bit
    nt.grad = None
    st.grad = None
    on2, os2, _ = bilinear_raysample_cuda(nt, st, T(z), T(tz), A, R, S, Tn)
    (on2 * T(gn)).sum().add((os2 * T(gs)).sum()).backward()
    g1 = nt.grad.clone()
    nt.grad = None
    on3, os3, _ = bilinear_raysample_cuda(nt, st, T(z), T(tz), A, R, S, Tn)
    (on3 * T(gn)).sum().add((os3 * T(gs)).sum()).backward()
    assert torch.equal(g1, nt.grad)


def _lensless_run(b, dev, with_grads=True, seed=0):
    from geraf.ops.RFSignalTracer.lensless_tracer import lensless_signal_tracer

    bb = b.to(dev)
    leaves = {k: getattr(bb, k).clone().requires_grad_(True) for k in
              ("sdf", "grad", "refl", "inv_s", "power") if getattr(bb, k) is not None}
    o_r, o_i = lensless_signal_tracer(leaves["sdf"], leaves["grad"], bb.dists, bb.z, bb.o, bb.d,
                                      leaves["inv_s"], leaves["power"],
                                      reflectance=leaves.get("refl"), trans_offset=bb.offset,
                                      n_valid=bb.n_valid, cos_anneal_ratio=b.anneal, **b.radar)
    out = o_r.detach().double().cpu().numpy() + 1j * o_i.detach().double().cpu().numpy()
    grads = None
    if with_grads:
        rng = np.random.default_rng(seed)
        g = rng.standard_normal(out.shape) + 1j * rng.standard_normal(out.shape)
        gt_r = torch.tensor(g.real, dtype=torch.float32, device=dev)
        gt_i = torch.tensor(g.imag, dtype=torch.float32, device=dev)
        (o_r * gt_r + o_i * gt_i).sum().backward()
        grads = {k: v.grad.double().cpu() for k, v in leaves.items()}
        grads["_g"] = g
    return out, grads


def _lensless_oracle(b, g=None):
    """FP64 oracle output and gradients.  inv_s and power enter as per-sample
    copies so that, besides their gradient (the sum), the sum of absolute
    per-sample contributions is known: it scales the tolerance of these two
    ill-conditioned sums."""
    f64 = lambda t: None if t is None else t.double()  # noqa: E731
    leaves = {k: getattr(b, k).double().clone().requires_grad_(True) for k in
              ("sdf", "grad", "refl") if getattr(b, k) is not None}
    for k in ("inv_s", "power"):
        leaves[k] = getattr(b, k).double().reshape(1, 1).expand(b.R, b.S).clone() \
            .requires_grad_(True)
    pts = L.points_f32(b.o, b.z, b.d)
    r = b.radar
    out = L.lensless_forward(leaves["sdf"], leaves["grad"], f64(b.dists), pts, f64(b.o),
                             f64(b.d), leaves["inv_s"], leaves["power"], r["numADCSample"],
                             r["adcSampleRate"], r["As_sci"], r["fc_start"], ts.SPEED_OF_LIGHT,
                             ts.PATH_BIAS, refl=leaves.get("refl"), offset=f64(b.offset),
                             n_valid=None if b.n_valid is None else b.n_valid.long(),
                             anneal=b.anneal)
    grads = None
    if g is not None:
        gt = torch.tensor(g)
        (out.real * gt.real + out.imag * gt.imag).sum().backward()
        grads = {k: v.grad for k, v in leaves.items()}
        for k in ("inv_s", "power"):
            grads[k + "_abs"] = grads[k].abs().sum()
            grads[k] = grads[k].sum()
    return out.detach().numpy(), grads


@pytest.mark.parametrize("case", ["plain", "offset_uneven", "anneal_half_K96", "K1", "K17"])
def test_lensless_vs_oracle(cuda, case):
    if case == "plain":
        b = ts.rays(256, 128, 256, seed=3)
    elif case in ("K1", "K17"):
        # one and a few frequencies: most of the synthesis tile and the h_s
        # gradient row are padding
        b = ts.rays(96, 80, int(case[1:]), seed=11, uneven=(8, 80))
    elif case == "offset_uneven":
        b = ts.rays(200, 96, 64, seed=5, with_offset=True, uneven=(16, 96))
    else:
        b = ts.rays(130, 70, 96, seed=7, anneal=0.5)
    out, grads = _lensless_run(b, cuda)
    ref, rgrads = _lensless_oracle(b, grads["_g"])
    err = rel(out, ref)
    print(f"lensless {case}: fwd rel-L2 {err:.2e}")
    assert err <= TOL_FWD
    for k in ("sdf", "grad", "refl"):
        e = rel(grads[k].numpy().ravel(), rgrads[k].numpy().ravel())
        print(f"  grad {k}: rel-L2 {e:.2e}")
        assert e <= TOL_GRAD, k
    # scalar parameters: the gradient sums +/- per-sample terms over the batch
    # (condition number sum|c|/|sum c| of 10-30 here); measure the error
    # against sum|c|
    for k in ("inv_s", "power"):
        e = abs(grads[k].item() - rgrads[k].item()) / rgrads[k + "_abs"].item()
        print(f"  grad {k}: err/sum|c| {e:.2e} (cond "
              f"{rgrads[k + '_abs'].item() / abs(rgrads[k].item()):.1f})")
        assert e <= TOL_GRAD, k


def test_lensless_config3_subset(cuda):
    b = ts.config3(0)
    out, _ = _lensless_run(b, cuda, with_grads=False)
    sub = ts.RayBatch(*(None if x is None else (x[:384] if torch.is_tensor(x) and x.dim() > 0
                                                   else x) for x in
                        (b.sdf, b.grad, b.dists, b.z, b.o, b.d, b.refl, b.offset, b.n_valid,
                         b.inv_s, b.power)), b.radar, b.anneal)
    ref, _ = _lensless_oracle(sub)
    assert rel(out[:384], ref) <= TOL_FWD


@pytest.mark.parametrize("R,S,K", [(96, 128, 200), (64, 96, 300), (64, 256, 512),
                                   (48, 160, 1000), (32, 80, 2048), (3, 2048, 129),
                                   (1, 64, 257), (5, 40, 1025), (64, 384, 256)])
def test_lensless_matrix_tiles(cuda, R, S, K):
    """The matrix-core synthesis and its adjoint (K > 128): one to eight
    256-frequency tiles, partial tiles (K = 200, 300, 1000), the first K of
    each tile count (129, 257, 1025), rays whose emitting samples overflow
    the 64-entry batch (S = 256, 384, and the largest S = 2048), a single
    ray (at least one valid sample), ragged rays with none."""
    b = ts.rays(R, S, K, seed=K, uneven=(0 if R > 1 else 1, S))
    out, grads = _lensless_run(b, cuda)
    ref, rgrads = _lensless_oracle(b, grads["_g"])
    err = rel(out, ref)
    print(f"lensless R{R} S{S} K{K}: fwd rel-L2 {err:.2e}")
    assert err <= TOL_FWD
    empty = (b.n_valid == 0).numpy()
    assert np.all(out[empty] == 0)
    for k in ("sdf", "grad", "refl"):
        e = rel(grads[k].numpy().ravel(), rgrads[k].numpy().ravel())
        print(f"  grad {k}: rel-L2 {e:.2e}")
        assert e <= TOL_GRAD, k
        assert np.all(grads[k].numpy()[empty] == 0)


@pytest.mark.parametrize("K", [512, 64])
def test_lensless_gradients_vs_finite_differences(cuda, K):
    """The GPU adjoint against FP64 central differences (h = 1e-4) of the loss
    Re<out, g> through the oracle, on the matrix-core (K = 512) and vector
    (K = 64) paths: every checked entry of the sdf, gradient, reflectance,
    inv_s and power gradients within 1e-4 of the difference quotient (scale
    max(|fd|, 1e-3 max|grad|)); picks whose +-h step crosses a branch of the
    composition (culling, the iter_cos relu) are skipped, as in the oracle's
    own check."""
    b = ts.rays(6, 96, K, seed=20 + K, uneven=(8, 96))
    _, grads = _lensless_run(b, cuda, seed=K)
    gK = torch.tensor(grads["_g"])
    f64 = lambda t: None if t is None else t.double()  # noqa: E731
    pts = L.points_f32(b.o, b.z, b.d)
    r = b.radar
    kw = dict(K=r["numADCSample"], fs=r["adcSampleRate"], As_sci=r["As_sci"], f0=r["fc_start"],
              c=ts.SPEED_OF_LIGHT, bias=ts.PATH_BIAS)
    nv = None if b.n_valid is None else b.n_valid.long()
    base = {k: getattr(b, k).double() for k in ("sdf", "grad", "refl", "inv_s", "power")}

    def loss(v):
        o = L.lensless_forward(v["sdf"], v["grad"], f64(b.dists), pts, f64(b.o), f64(b.d),
                               v["inv_s"], v["power"], refl=v["refl"], offset=f64(b.offset),
                               n_valid=nv, anneal=b.anneal, **kw)
        return float((o.real * gK.real + o.imag * gK.imag).sum())

    def masks(v):
        return L.masks(v["sdf"], v["grad"], f64(b.dists), pts, f64(b.o), f64(b.d), v["inv_s"],
                       v["power"], refl=v["refl"], offset=f64(b.offset), n_valid=nv,
                       anneal=b.anneal)

    rng = np.random.default_rng(K)
    h = 1e-4
    checked = 0
    for name in ("inv_s", "power", "sdf", "refl", "grad"):
        flat = base[name].reshape(-1)
        gpu = grads[name].reshape(-1)
        scale0 = 1e-3 * float(gpu.abs().max())
        picks = [0] if flat.numel() == 1 else rng.choice(flat.numel(), 10, False)
        for i in picks:
            def shifted(dv):
                v = dict(base)
                t = flat.clone()
                t[i] += dv
                v[name] = t.reshape(base[name].shape)
                return v
            mp, mm = masks(shifted(h)), masks(shifted(-h))
            if not torch.equal(mp["live"], mm["live"]) or not torch.equal(
                    mp["true_cos"] < 0, mm["true_cos"] < 0):
                continue
            fd = (loss(shifted(h)) - loss(shifted(-h))) / (2 * h)
            ad = float(gpu[i])
            assert abs(fd - ad) <= 1e-4 * max(abs(fd), scale0), (name, int(i), fd, ad)
            checked += 1
    print(f"lensless K={K}: {checked} gradient entries within 1e-4 of central differences")
    assert checked >= 25


@pytest.mark.parametrize("R,S,K,offset", [(200, 96, 64, True), (64, 256, 512, False),
                                          (3, 2048, 129, True), (32, 80, 2048, False),
                                          (96, 80, 17, True)])
def test_lensless_backward_with_forward_carries(cuda, R, S, K, offset):
    """The backward given the forward's per-tile transmittance carries (one
    walk over the samples, h_s tile by tile) equals the backward that
    rebuilds them (two walks, h_s in 64-sample batches) bit for bit, on
    ragged rays (some empty), with and without offsets, over the Horner,
    matrix-core and largest-S forms."""
    from geraf.ops import _native
    from geraf.ops.RFSignalTracer import _C

    b = ts.rays(R, S, K, seed=K + S, with_offset=offset, uneven=(0 if R > 1 else 1, S)).to(cuda)
    P = _native.ptr
    r = b.radar
    prm = _C.tracer_params(r["fc_start"], r["As_sci"], r["adcSampleRate"])
    inv_s, power = b.inv_s.reshape(1).contiguous(), b.power.reshape(1).contiguous()
    nv = b.n_valid.to(torch.int32).contiguous()
    ins = [P(b.sdf), P(b.grad), P(b.refl), P(b.dists), P(b.z), P(b.o), P(b.d), P(b.offset),
           P(nv), P(inv_s), P(power)]
    o_r = torch.empty((R, K), device=cuda)
    o_i = torch.empty_like(o_r)
    carry = torch.full((R, (S + 63) // 64), float("nan"), device=cuda)
    st = _native.stream_of(o_r)
    _native.call("geraf_lensless_forward_f32", *ins, P(o_r), P(o_i), P(None), P(carry), R, S, K,
                 b.anneal, prm, st)
    o2_r, o2_i = torch.empty_like(o_r), torch.empty_like(o_r)
    _native.call("geraf_lensless_forward_f32", *ins, P(o2_r), P(o2_i), P(None), P(None), R, S, K,
                 b.anneal, prm, st)
    torch.cuda.synchronize()
    assert torch.equal(o_r, o2_r) and torch.equal(o_i, o2_i)
    # the carries of the tiles with valid samples: products of the first
    # 64 t samples, 1 at the first tile
    ntile = (nv.long() + 63) // 64
    written = torch.arange(carry.shape[1], device=cuda)[None, :] < ntile[:, None]
    assert torch.all(torch.isfinite(carry[written]))
    assert torch.all(carry[:, 0][ntile > 0] == 1.0)
    assert torch.all(torch.isnan(carry[~written]))

    gen = torch.Generator(device=cuda).manual_seed(K)
    g_r = torch.randn((R, K), device=cuda, generator=gen)
    g_i = torch.randn((R, K), device=cuda, generator=gen)
    outs = []
    for tc in (carry, None):
        o = dict(gs=torch.full((R, S), 7.0, device=cuda), gg=torch.full((R, S, 3), 7.0, device=cuda),
                 gr=torch.full((R, S), 7.0, device=cuda), gi=torch.empty(R, device=cuda),
                 gp=torch.empty(R, device=cuda))
        _native.call("geraf_lensless_backward_f32", P(g_r), P(g_i), *ins, P(o["gs"]), P(o["gg"]),
                     P(o["gr"]), P(o["gi"]), P(o["gp"]), P(tc), R, S, K, b.anneal, prm, st)
        outs.append(o)
    torch.cuda.synchronize()
    for k in outs[0]:
        assert torch.equal(outs[0][k], outs[1][k]), k
    assert bool(torch.any(outs[0]["gs"] != 0))


def test_lensless_backward_refuses_inputs_modified_in_place(cuda):
    """The op saves its inputs for the backward: modifying one in place after
    the forward makes autograd raise instead of giving wrong gradients."""
    from geraf.ops.RFSignalTracer.lensless_tracer import lensless_signal_tracer

    b = ts.rays(8, 32, 64, seed=1).to(cuda)
    sdf = b.sdf.clone().requires_grad_(True)
    refl = b.refl.clone()
    o_r, o_i = lensless_signal_tracer(sdf * 1.0, b.grad, b.dists, b.z, b.o, b.d, b.inv_s,
                                      b.power, reflectance=refl, **b.radar)
    refl.mul_(2.0)
    with pytest.raises(RuntimeError, match="modified by an inplace operation"):
        (o_r.sum() + o_i.sum()).backward()


@pytest.mark.parametrize("order", ["sorted", "shuffled"])
@pytest.mark.parametrize("S,Tn", [(256, 512), (1536, 700), (5000, 300)])
def test_bilinear_backward_target_orders(cuda, order, S, Tn):
    """The resampler's transpose with sorted target depths (contiguous
    bracket ranges found by binary search; targets below and above the source
    depths included) and with shuffled ones (the scan over all targets),
    against the oracle; and each path is bit-reproducible.  Rays of >= 1,024
    source samples share one bracket search per boundary through LDS; 5,000
    samples exceed the depth staging in LDS (global-memory searches)."""
    from geraf.ops import _native

    rng = np.random.default_rng(31)
    A, R = 2, 48
    z = np.sort(rng.uniform(0.2, 0.5, (R, S)), axis=1).astype(np.float32)
    tz = np.sort(rng.uniform(0.18, 0.52, (R, Tn)), axis=1).astype(np.float32)
    tz[:, 100:104] = z[:, 50:51]  # exact hits on a source depth
    if order == "shuffled":
        tz = np.ascontiguousarray(tz[:, rng.permutation(Tn)])
    gn = rng.standard_normal((R, Tn, 3)).astype(np.float32)
    gs = rng.standard_normal((A, R, Tn)).astype(np.float32)
    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    ins = [T(gn), T(gs), T(z), T(tz)]
    outs = []
    for _ in range(2):
        o_n = torch.empty((R, S, 3), device=cuda)
        o_s = torch.empty((A, R, S), device=cuda)
        _native.call("geraf_bilinear_backward_f32", *map(_native.ptr, ins + [o_n, o_s]), A, R, S,
                     Tn, _native.stream_of(o_n))
        torch.cuda.synchronize()
        outs.append((o_n.cpu().numpy(), o_s.cpu().numpy()))
    rgn, rgs = orc.bilinear_backward(gn, gs, z, tz, S)
    assert np.max(np.abs(outs[0][0] - rgn)) <= 2e-5
    assert np.max(np.abs(outs[0][1] - rgs)) <= 2e-5
    assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
    # the forward at the same ray length (the ray's depths staged in LDS up to
    # 4,096 samples, searched in global memory beyond)
    nrm = rng.standard_normal((R, S, 3)).astype(np.float32)
    sig = rng.uniform(0, 1, (A, R, S)).astype(np.float32)
    f_n = torch.empty((R, Tn, 3), device=cuda)
    f_s = torch.empty((A, R, Tn), device=cuda)
    f_i = torch.empty((R, Tn), dtype=torch.bool, device=cuda)
    fins = [T(nrm), T(sig), T(z), T(tz), f_n, f_s, f_i]
    _native.call("geraf_bilinear_forward_f32", *map(_native.ptr, fins), A, R, S, Tn,
                 _native.stream_of(f_n))
    torch.cuda.synchronize()
    rn, rs, rinb = orc.bilinear_forward(nrm, sig, z, tz)
    assert np.array_equal(f_i.cpu().numpy(), rinb)
    assert np.max(np.abs(f_n.cpu().numpy() - rn)) <= 1e-6
    assert np.max(np.abs(f_s.cpu().numpy() - rs)) <= 1e-6
