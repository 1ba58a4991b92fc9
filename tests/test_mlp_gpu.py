"""The SDF network's BF16 matrix-core kernels on the GPU: the layer product
(every prologue and epilogue, ragged widths, broadcast rows) and the weight
gradient against torch on the same BF16-rounded operands; the hand-written
network (forward, normals, backward with the double-backward terms) against
the FP64 torch network, within the error of torch's own BF16 network."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

BETA, TH = 100.0, 20.0


def _bf(t):
    return t.to(torch.bfloat16)


def _sp(z):
    return torch.nn.functional.softplus(z, BETA, TH)


def _slope(z):
    return torch.where(z * BETA > TH, torch.ones_like(z), torch.sigmoid(z * BETA))


def _curv(z):
    s = torch.sigmoid(z * BETA)
    return torch.where(z * BETA > TH, torch.zeros_like(z), s * (1 - s))


def _code(z):
    """The slope code of z: copysign(sigmoid(-|beta z|), beta z), FP16."""
    t = z.double() * BETA
    return torch.copysign(torch.sigmoid(-t.abs()), t).to(torch.float16)


def _code_slope(q):
    m = q.float().abs()
    return torch.where(torch.signbit(q), m, 1.0 - m)


def _code_curv(q):
    m = q.float().abs()
    return m * (1.0 - m)


def test_slope_code_closed_forms():
    """The code's decoding: slope and curvature of z, exact limits (slope 1 /
    curvature 0 above the threshold, slope 0 for t -> -inf through -0)."""
    z = torch.tensor([-1.0, -0.05, -0.003, 0.0, 0.004, 0.05, 0.19, 0.21, 5.0])
    q = _code(z)
    s, c = _code_slope(q), _code_curv(q)
    assert torch.allclose(s, _slope(z), rtol=1e-3, atol=1e-7)
    assert torch.allclose(c, _curv(z), rtol=2e-3, atol=1e-7)
    assert float(s[-1]) == 1.0 and float(c[-1]) == 0.0 and float(s[0]) == 0.0
    assert torch.signbit(q[0]) and float(q[3]) == 0.5


def _rel(a, b):
    a, b = a.double(), b.double()
    return float((a - b).norm() / b.norm().clamp(min=1e-30))


@pytest.mark.parametrize("M,N,K", [(1000, 256, 256), (77, 217, 256), (513, 256, 217),
                                   (300, 257, 256), (65, 256, 257), (129, 256, 39),
                                   (96, 39, 256), (31, 1, 64)])
@pytest.mark.parametrize("pro", [0, 1])
def test_layer_product_epilogues(cuda, M, N, K, pro):
    from geraf.models.networks import fused_sdf as fs

    g = torch.Generator(device="cpu").manual_seed(M * 7 + N + K + pro)
    ldx = ((K + 7) // 8) * 8 + 8
    x = _bf(torch.randn(M, ldx, generator=g) * 0.5).to(cuda)
    x2 = _code(torch.randn(M, ldx, generator=g) * 0.02).to(cuda)
    w = (torch.randn(N, K, generator=g) / math.sqrt(K)).to(cuda)
    b = (torch.randn(N, generator=g) * 0.01).to(cuda)
    ldy = ((N + 7) // 8) * 8
    e0 = _code(torch.randn(M, ldy, generator=g) * 0.02).to(cuda)
    e1 = _bf(torch.randn(M, ldy, generator=g)).to(cuda)
    packed = fs.pack(w, False)
    xe = x[:, :K].float()
    if pro:
        xe = _bf(xe * _code_slope(x2[:, :K])).float()
    z = xe @ _bf(w).float().t() + b
    cases = {0: _bf(z), 1: _bf(_sp(z)), 3: _bf(z.clamp(min=0))}
    for epi, want in cases.items():
        y = torch.full((M, ldy), float("nan"), dtype=torch.bfloat16, device=cuda)
        fs.gemm(x, packed, N, K, y, bias=b, x2=x2 if pro else None, prologue=pro, epilogue=epi)
        err = _rel(y[:, :N].float(), want.float())
        assert err <= 8e-3, (epi, err)
    # training softplus: y = softplus(z), y2 = the slope code of z
    y = torch.empty((M, ldy), dtype=torch.bfloat16, device=cuda)
    y2 = torch.empty((M, ldy), dtype=torch.float16, device=cuda)
    fs.gemm(x, packed, N, K, y, bias=b, x2=x2 if pro else None, y2=y2, prologue=pro, epilogue=2)
    assert _rel(y[:, :N].float(), _sp(z)) <= 8e-3
    assert _rel(_code_slope(y2[:, :N]), _slope(z)) <= 2e-3
    assert _rel(_code_curv(y2[:, :N]), _curv(z)) <= 4e-3
    # adjoints (no bias): y = z, y2 = z e1 beta curv(e0); y = z s(e0) + e1
    y2 = torch.empty((M, ldy), dtype=torch.bfloat16, device=cuda)
    fs.gemm(x, packed, N, K, y, x2=x2 if pro else None, y2=y2, e0=e0, e1=e1, prologue=pro,
            epilogue=4)
    z0 = xe @ _bf(w).float().t()
    assert _rel(y[:, :N].float(), z0) <= 8e-3
    want = z0 * e1[:, :N].float() * BETA * _code_curv(e0[:, :N])
    assert _rel(y2[:, :N].float(), want) <= 1e-2
    fs.gemm(x, packed, N, K, y, x2=x2 if pro else None, e0=e0, e1=e1, prologue=pro, epilogue=5)
    want = z0 * _code_slope(e0[:, :N]) + e1[:, :N].float()
    assert _rel(y[:, :N].float(), want) <= 8e-3
    # slope-multiplied outputs: y = z s(e0) (6, no e1); with y2 = z e1 beta
    # (1 - s(e0)) (7) or z e1 beta s(e0) (1 - s(e0)) (8)
    zs = z0 * _code_slope(e0[:, :N])
    y.fill_(float("nan"))
    fs.gemm(x, packed, N, K, y, x2=x2 if pro else None, e0=e0, prologue=pro, epilogue=6)
    assert _rel(y[:, :N].float(), zs) <= 8e-3
    one_minus = 1.0 - _code_slope(e0[:, :N])
    for epi, fac in ((7, one_minus), (8, _code_curv(e0[:, :N]))):
        y.fill_(float("nan"))
        y2.fill_(float("nan"))
        fs.gemm(x, packed, N, K, y, x2=x2 if pro else None, y2=y2, e0=e0, e1=e1, prologue=pro,
                epilogue=epi)
        assert _rel(y[:, :N].float(), zs) <= 8e-3, epi
        want = z0 * e1[:, :N].float() * BETA * fac
        assert _rel(y2[:, :N].float(), want) <= 1e-2, epi


def test_layer_product_transposed_pack_and_broadcast(cuda):
    from geraf.models.networks import fused_sdf as fs

    g = torch.Generator(device="cpu").manual_seed(5)
    M, N, K = 700, 256, 217
    w = torch.randn(K, N, generator=g).to(cuda) / 16  # used as W^T
    row = _bf(torch.randn(1, 224, generator=g)).to(cuda)
    x = row.expand(M, 224)
    x2 = _code(torch.randn(M, 224, generator=g) * 0.03).to(cuda)
    y = torch.empty((M, N), dtype=torch.bfloat16, device=cuda)
    fs.gemm(x, fs.pack(w, True), N, K, y, x2=x2, prologue=1)
    xe = _bf(row[:, :K].float() * _code_slope(x2[:, :K])).float()
    assert _rel(y.float(), xe @ _bf(w).float()) <= 8e-3


@pytest.mark.parametrize("M,N,K,pro", [(5000, 256, 256, (1, 1)), (777, 257, 256, (0, 0)),
                                       (70001, 257, 256, (0, 0)), (4097, 272, 40, (0, 0)),
                                       (333, 217, 256, (1, 0)), (1024, 256, 39, (0, 1)),
                                       (64, 1, 256, (0, 1)), (100000, 256, 256, (0, 0))])
def test_weight_gradient(cuda, M, N, K, pro):
    from geraf.models.networks import fused_sdf as fs

    gen = torch.Generator(device="cpu").manual_seed(M + N + K)
    ldg, ldx = ((N + 7) // 8) * 8, ((K + 7) // 8) * 8
    G = _bf(torch.randn(M, ldg, generator=gen)).to(cuda)
    X = _bf(torch.randn(M, ldx, generator=gen)).to(cuda)
    G2 = _code(torch.randn(M, ldg, generator=gen) * 0.02).to(cuda)
    X2 = _code(torch.randn(M, ldx, generator=gen) * 0.02).to(cuda)
    dw = torch.full((N, K), 0.5, device=cuda)
    db = torch.full((N,), 0.25, device=cuda)
    fs.wgrad(G, X, N, K, dw, db=db, g2=G2 if pro[0] else None, x2=X2 if pro[1] else None,
             accumulate=True)
    Ge, Xe = G[:, :N].float(), X[:, :K].float()
    if pro[0]:
        Ge = _bf(Ge * _code_slope(G2[:, :N])).float()
    if pro[1]:
        Xe = _bf(Xe * _code_slope(X2[:, :K])).float()
    want = Ge.double().t() @ Xe.double() + 0.5
    # FP32 sums of the same BF16 operands (the slope is exact in the code)
    assert _rel(dw, want) <= 1e-5
    assert _rel(db, Ge.double().sum(0) + 0.25) <= 1e-5
    # deterministic
    dw2 = torch.full((N, K), 0.5, device=cuda)
    fs.wgrad(G, X, N, K, dw2, g2=G2 if pro[0] else None, x2=X2 if pro[1] else None,
             accumulate=True)
    assert torch.equal(dw, dw2)


def _nets(cuda, seed=0):
    from geraf.models.networks.sdf_network import SDFNetwork

    torch.manual_seed(seed)
    ref = SDFNetwork(3, 257, 256, 8, skip_in=(4,), multires=6).to(cuda)
    with torch.no_grad():  # move off the exact geometric init (features and skip weights nonzero)
        for p in ref.parameters():
            p.add_(0.02 * torch.randn_like(p))
    bf = SDFNetwork(3, 257, 256, 8, skip_in=(4,), multires=6,
                    compute_dtype=torch.bfloat16).to(cuda)
    bf.load_state_dict(ref.state_dict())
    return ref, bf


def test_fused_network_vs_fp64(cuda):
    """sdf, features, normals and the gradients of a loss through all three
    with respect to every layer's weights and bias (as the network uses them,
    after the weight normalisation) against the FP64 network; each error
    within 1.5x that of torch's BF16 network (split-K products, fused
    softplus) on the same points, i.e. within BF16 rounding."""
    import torch.nn.utils.parametrize as parametrize

    from geraf.models.networks.fused_sdf import fused_sdf

    ref, bfnet = _nets(cuda)
    M = 20000
    g = torch.Generator(device="cpu").manual_seed(1)
    x = (torch.rand(M, 3, generator=g) * 1.6 - 0.8).to(cuda)
    a = torch.randn(M, generator=g).to(cuda)
    bvec = torch.randn(M, 256, generator=g).to(cuda) * 0.05
    c = torch.randn(M, 3, generator=g).to(cuda)

    def run(kind):
        net = ref if kind == "f64" else bfnet
        if kind == "f64":
            net.double()
        with parametrize.cached():
            layers = [(getattr(net, f"lin{l}").weight, getattr(net, f"lin{l}").bias)
                      for l in range(9)]
            if kind == "f64":
                xx = x.double().clone().requires_grad_(True)
                out = net(xx)
                sdf, feat = out[:, 0], out[:, 1:]
                n = torch.autograd.grad(sdf.sum(), xx, create_graph=True)[0]
            elif kind == "torch_bf16":
                out, n = net.sdf_and_gradient(x.clone())
                sdf, feat = out[:, 0], out[:, 1:]
            else:
                sdf, feat, n = fused_sdf(net, x)
            L = (a.double() * sdf.double()).sum() + (bvec.double() * feat.double()).sum() + \
                (c.double() * n.double()).sum()
            grads = torch.autograd.grad(L, [t for wb in layers for t in wb])
        vals = (sdf.detach().double(), feat.detach().double(), n.detach().double())
        if kind == "f64":
            net.float()
        return vals, [t.double() for t in grads]

    v64, g64 = run("f64")
    vt, gt = run("torch_bf16")
    vf, gf = run("fused")
    for i, name in enumerate(["sdf", "features", "normals"]):
        et, ef = _rel(vt[i], v64[i]), _rel(vf[i], v64[i])
        print(f"{name}: fused {ef:.2e}, torch BF16 {et:.2e}", flush=True)
        assert ef <= 1.5 * et + 1e-4, name
    worst = 0.0
    for i, (a64, at, af) in enumerate(zip(g64, gt, gf)):
        name = f"lin{i // 2}.{'weight' if i % 2 == 0 else 'bias'}"
        et, ef = _rel(at, a64), _rel(af, a64)
        worst = max(worst, ef / et)
        print(f"grad {name}: fused {ef:.2e}, torch BF16 {et:.2e}", flush=True)
        assert ef <= 1.5 * et + 1e-4, name
    print(f"worst gradient error ratio fused / torch BF16: {worst:.2f}", flush=True)


def test_fused_sdf_value_and_no_normals(cuda):
    from geraf.models.networks.fused_sdf import fused_sdf, fused_sdf_value

    ref, bfnet = _nets(cuda, seed=3)
    x = (torch.rand(3001, 3, device=cuda) * 1.6 - 0.8)
    s1 = fused_sdf_value(bfnet, x)
    s2, _, n = fused_sdf(bfnet, x, with_normals=False)
    with torch.no_grad():
        s64 = ref.double()(x.double())[:, 0]
    ref.float()
    # the value path sums the sdf row in FP32; the layer form rounds it to BF16
    s2d = s2.detach()
    assert _rel(s1, s64) <= 1e-2
    assert float((s1 - s2d).abs().max()) <= 2.0 ** -8 * float(s2d.abs().max()) + 1e-6
    assert torch.all(n == 0)
    # no normals: the backward is the primal adjoint alone
    bfnet.zero_grad(set_to_none=True)
    s2.sum().backward()
    assert all(p.grad is not None and torch.isfinite(p.grad).all() for p in bfnet.parameters())


@pytest.mark.parametrize("M", [1, 300, 70001])
def test_chained_forward_matches_layers(cuda, M):
    """The chained hidden layers (one kernel, activations in registers)
    against the same forward run one layer product at a time: h_l and the
    slope codes q_l bit for bit (the skip's gamma columns and slope-one
    codes included), the output layer equal, and the value path's FP32 sdf
    within the layer form's BF16 rounding.  M = 70,001 is 274 tiles of 256
    points, more than one per workgroup."""
    from geraf.models.networks import fused_sdf as fs

    _, bfnet = _nets(cuda, seed=5)
    g = torch.Generator(device="cpu").manual_seed(M)
    x = (torch.rand(M, 3, generator=g) * 1.6 - 0.8).to(cuda)
    params = [t for l in range(9)
              for t in (getattr(bfnet, f"lin{l}").weight, getattr(bfnet, f"lin{l}").bias)]
    ws, bs = list(params[0::2]), fs._biases(params[1::2])
    P, _, hidden, _ = fs._packs(ws)
    _, q0, h0, out0 = fs._forward_layers(x, ws, bs, P, train=True)
    _, q1, h1, out1 = fs._forward(x, P, hidden, fs._bias_table(bs), bs[8])
    for l in range(8):
        assert torch.equal(h0[l].view(torch.int16), h1[l].view(torch.int16)), l
        assert torch.equal(q0[l].view(torch.int16), q1[l].view(torch.int16)), l
    assert torch.equal(out0[:, :257].view(torch.int16), out1[:, :257].view(torch.int16))
    v = fs.fused_sdf_value(bfnet, x)
    s = out0[:, 256].float()
    assert float((v - s).abs().max()) <= 2.0 ** -8 * float(s.abs().max()) + 1e-6


@pytest.mark.parametrize("M", [1, 300, 70001])
@pytest.mark.parametrize("mode", ["normal", "primal"])
def test_chained_sweep_matches_layers(cuda, M, mode):
    """The chained reverse sweep over layers 7 .. 1 (W_l^T, W_3^T padded with
    zero weights) against seven layer products, bit for bit: the normal pass
    (broadcast first row with the slope prologue, y = z s(q)) and the primal
    adjoint (y = z s(q) + t, also written in place over t)."""
    from geraf.models.networks import fused_sdf as fs

    _, bfnet = _nets(cuda, seed=7)
    ws = [getattr(bfnet, f"lin{l}").weight for l in range(9)]
    rev, Q = fs._reverse_pack(ws)
    g = torch.Generator(device="cpu").manual_seed(M + len(mode))
    q = _code(torch.randn(8, M, 256, generator=g) * 0.02).to(cuda)
    if mode == "normal":
        x = _bf(torch.randn(1, 256, generator=g)).to(cuda).expand(M, 256)
        x2, t = q[7], None
    else:
        x = _bf(torch.randn(M, 256, generator=g)).to(cuda)
        x2, t = None, _bf(torch.randn(7, M, 256, generator=g) * 0.1).to(cuda)
    y = torch.empty((7, M, 256), dtype=torch.bfloat16, device=cuda)
    fs._sweep(x, rev, q, t, y, x2=x2)
    cur = x
    for l in range(7, 0, -1):
        out = torch.empty((M, 256), dtype=torch.bfloat16, device=cuda)
        if mode == "normal":
            fs.gemm(cur, Q[l], 256, 256, out, x2=x2 if l == 7 else None, e0=q[l - 1],
                    prologue=1 if l == 7 else 0, epilogue=6)
        else:
            fs.gemm(cur, Q[l], 256, 256, out, e0=q[l - 1], e1=t[l - 1], epilogue=5)
        assert torch.equal(out.view(torch.int16), y[l - 1].view(torch.int16)), l
        cur = out
    if t is not None:
        fs._sweep(x, rev, q, t, t)
        assert torch.equal(t.view(torch.int16), y.view(torch.int16))


def test_fused_reflectance_vs_fp64(cuda):
    """The reflectance network (259 -> 256 -> 256 -> 256 -> 1, ReLU, sigmoid)
    on the matrix-core kernels: r and the gradients of a loss through it
    with respect to the features and every layer's weight and bias, against
    the FP64 network; each error within 1.5x that of torch's BF16 network.
    Features read in place from [features | sdf | padding] rows (the SDF
    network's output buffer) give the same results, bit for bit, as a
    contiguous copy."""
    import torch.nn.utils.parametrize as parametrize

    from geraf.models.networks.fused_reflectance import fused_reflectance
    from geraf.models.networks.sdf_network import ReflectivePowerNetwork

    torch.manual_seed(2)
    ref = ReflectivePowerNetwork(256).to(cuda)
    with torch.no_grad():
        for p in ref.parameters():
            p.add_(0.02 * torch.randn_like(p))
    bfn = ReflectivePowerNetwork(256, compute_dtype=torch.bfloat16).to(cuda)
    bfn.load_state_dict(ref.state_dict())
    M = 30001
    g = torch.Generator(device="cpu").manual_seed(3)
    pts = (torch.rand(M, 3, generator=g) * 1.6 - 0.8).to(cuda)
    rows = torch.zeros(M, 264, dtype=torch.bfloat16, device=cuda)
    rows[:, :256] = (torch.randn(M, 256, generator=g) * 0.3).to(cuda)
    rows[:, 256] = 7.0  # the sdf column: multiplied by a zero weight
    feat_view = rows[:, :256]
    feat = feat_view.contiguous()
    w = torch.randn(M, generator=g).to(cuda)

    def run(kind, f):
        net = ref if kind == "f64" else bfn
        if kind == "f64":
            net.double()
        lins = [net.refelective_predictor[i] for i in (0, 2, 4, 6)]
        with parametrize.cached():
            params = [t for l in lins for t in (l.weight, l.bias)]
            fe = f.detach().double().requires_grad_(True) if kind == "f64" else \
                f.detach().clone().requires_grad_(True) if kind == "torch_bf16" else \
                f.detach().requires_grad_(True)
            if kind == "fused":
                r = fused_reflectance(net, pts, fe)
            else:
                r = net.reflectance(pts.double() if kind == "f64" else pts, fe)
            L = (w.double() * r.double()).sum()
            grads = torch.autograd.grad(L, [fe] + params)
        if kind == "f64":
            net.float()
        return r.detach().double(), [t.double() for t in grads]

    r64, g64 = run("f64", feat)
    rt, gt = run("torch_bf16", feat)
    rf, gf = run("fused", feat)
    rv, gv = run("fused", feat_view)
    assert torch.equal(rf, rv) and all(torch.equal(a, b) for a, b in zip(gf, gv))
    et, ef = _rel(rt, r64), _rel(rf, r64)
    print(f"reflectance: fused {ef:.2e}, torch BF16 {et:.2e}", flush=True)
    assert ef <= 1.5 * et + 1e-4
    names = ["features"] + [f"lin{i}.{k}" for i in range(4) for k in ("weight", "bias")]
    for name, a64, at, af in zip(names, g64, gt, gf):
        et, ef = _rel(at, a64), _rel(af, a64)
        print(f"grad {name}: fused {ef:.2e}, torch BF16 {et:.2e}", flush=True)
        assert ef <= 1.5 * et + 1e-4, name
