"""The MLPs that feed the lensless tracer (CPU, float64): the split-K tall
products are exact restatements of the plain layer products, to any
derivative order."""
import pytest
import torch

from geraf.models.networks.sdf_network import (ReflectivePowerNetwork, SDFNetwork, col_sum,
                                                linear, tall_matmul, tall_tn)


@pytest.mark.parametrize("m", [64, 96, 37])
def test_tall_products_gradcheck(m):
    torch.manual_seed(m)
    d = torch.float64
    a = torch.randn(m, 5, dtype=d, requires_grad=True)
    one = torch.randn(m, 1, dtype=d, requires_grad=True)
    b = torch.randn(5, 3, dtype=d, requires_grad=True)
    g = torch.randn(m, 3, dtype=d, requires_grad=True)
    assert torch.autograd.gradcheck(tall_matmul, (a, b))
    assert torch.autograd.gradgradcheck(tall_matmul, (a, b))
    assert torch.autograd.gradcheck(tall_tn, (a, g))
    assert torch.autograd.gradgradcheck(tall_tn, (a, g))
    assert torch.autograd.gradgradcheck(tall_tn, (a, one))
    assert torch.autograd.gradgradcheck(tall_tn, (one, g))
    lin = torch.nn.Linear(5, 3).double()
    assert torch.autograd.gradgradcheck(lambda x: linear(x, lin, None), (a,))
    # split-K sum equals the plain product; blocked column sums equal sum(0)
    assert torch.allclose(tall_tn(a, g), a.t() @ g, rtol=1e-13, atol=1e-13)
    assert torch.allclose(col_sum(g), g.sum(0), rtol=1e-13, atol=1e-13)
    assert torch.autograd.gradgradcheck(col_sum, (g,))


def _eikonal_step(net, x):
    out = net.sdf_hidden_appearance(x)
    grad = torch.autograd.grad(out[:, :1], x, torch.ones_like(out[:, :1]), create_graph=True)[0]
    loss = out[:, 0].abs().mean() + ((grad.norm(dim=-1) - 1.0) ** 2).mean() + out[:, 1:].mean()
    params = list(net.parameters())
    gp = torch.autograd.grad(loss, params)
    return out.detach(), grad.detach(), gp


def test_sdf_network_tall_matches_plain():
    torch.manual_seed(0)
    kw = dict(d_in=3, d_out=17, d_hidden=64, n_layers=8, skip_in=(4,), multires=6)
    tall = SDFNetwork(**kw, tall_products=True).double()
    plain = SDFNetwork(**kw, tall_products=False).double()
    plain.load_state_dict(tall.state_dict())
    x = (torch.rand(160, 3, dtype=torch.float64) * 2 - 1).requires_grad_(True)
    o1, g1, p1 = _eikonal_step(tall, x)
    x2 = x.detach().clone().requires_grad_(True)
    o2, g2, p2 = _eikonal_step(plain, x2)
    assert torch.allclose(o1, o2, rtol=1e-12, atol=1e-12)
    assert torch.allclose(g1, g2, rtol=1e-12, atol=1e-12)
    for a, b in zip(p1, p2):
        assert torch.allclose(a, b, rtol=1e-10, atol=1e-12)


def test_reflectance_network_tall_matches_plain():
    torch.manual_seed(1)
    tall = ReflectivePowerNetwork(16, tall_products=True).double()
    plain = ReflectivePowerNetwork(16, tall_products=False).double()
    plain.load_state_dict(tall.state_dict())
    pts = torch.randn(50, 3, dtype=torch.float64)
    feat = torch.randn(50, 16, dtype=torch.float64, requires_grad=True)
    r1 = tall.reflectance(pts, feat)
    r2 = plain.reflectance(pts, feat)
    assert torch.allclose(r1, r2, rtol=1e-12, atol=1e-12)
    g1 = torch.autograd.grad(r1.sum(), list(tall.refelective_predictor.parameters()) + [feat])
    g2 = torch.autograd.grad(r2.sum(), list(plain.refelective_predictor.parameters()) + [feat])
    for a, b in zip(g1, g2):
        assert torch.allclose(a, b, rtol=1e-10, atol=1e-12)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [4096 * 256, 1000 * 39 + 3, -(256 * 64)])
def test_softplus_bf16_kernels(cuda, n):
    """The fused BF16 softplus (forward, backward, double backward) against
    FP64 autograd on the same BF16 inputs: every output within one BF16
    rounding (2^-8 relative) of the FP64 value, and never farther from it than
    torch's own BF16 chain plus one rounding.  Inputs straddle the linear
    threshold (beta x in [-60, 60]); n = 39003 covers the unaligned tail; a
    negative n feeds a transposed (non-contiguous) view, whose second-order
    gradients must still reach it.
    (FP32 autograd is not the reference: its 1 - sigmoid cancels to 0 for
    beta x in about [16.6, 20], where the curvature is ~beta exp(-beta x).)"""
    from torch.nn import functional as F

    from geraf.models.networks.sdf_network import SoftplusBF16

    torch.manual_seed(0)
    if n > 0:
        x = ((torch.rand(n, device=cuda) - 0.5) * 1.2).bfloat16().requires_grad_(True)
        shape = (n,)
    else:
        base = ((torch.rand(256, -n // 256, device=cuda) - 0.5) * 1.2).bfloat16()
        x = base.requires_grad_(True).t()
        assert not x.is_contiguous()
        shape = x.shape
    g = torch.randn(shape, device=cuda).bfloat16().requires_grad_(True)
    gg = torch.randn(shape, device=cuda).bfloat16()

    def chain(act, xx, gg_in, g_in):
        y = act(xx)
        (gx,) = torch.autograd.grad(y, xx, g_in, create_graph=True)
        gx_x, gx_g = torch.autograd.grad(gx, (xx, g_in), gg_in)
        return [t.detach().double() for t in (y, gx, gx_x, gx_g)]

    ours = chain(SoftplusBF16(beta=100), x, gg, g)
    tb = chain(torch.nn.Softplus(beta=100), x, gg, g)
    xf = x.detach().double().requires_grad_(True)
    gf = g.detach().double().requires_grad_(True)
    ref = chain(lambda t: F.softplus(t, beta=100), xf, gg.double(), gf)
    for name, o, t, r in zip(("y", "dy", "d2y/dx", "d2y/dg"), ours, tb, ref):
        tol = 2.0 ** -8 * r.abs() + 1e-30
        err = (o - r).abs()
        print(f"softplus {name}: max err/|ref| ours {float((err / (r.abs() + 1e-30)).max()):.2e}, "
              f"torch bf16 {float(((t - r).abs() / (r.abs() + 1e-30)).max()):.2e}")
        assert bool((err <= tol).all()), name
        assert bool((err <= (t - r).abs() + tol).all()), name


def test_positional_encoding_order_and_derivatives():
    """The all-bands encoding equals the reference's per-band Embedder form
    (sdf_network.py:19-75) bit for bit, and its first and second derivatives
    match that form's in FP64."""
    from geraf.models.networks.sdf_network import positional_encoding

    def per_band(x, multires):
        feats = [x]
        for i in range(multires):
            f = float(2.0 ** i)
            feats.append(torch.sin(x * f))
            feats.append(torch.cos(x * f))
        return torch.cat(feats, dim=-1)

    g = torch.Generator().manual_seed(0)
    x = torch.rand(5, 7, 3, generator=g) * 2 - 1
    for m in (1, 6, 10):
        assert torch.equal(positional_encoding(x, m), per_band(x, m))
    assert torch.equal(positional_encoding(x, 0), x)
    xd = x.double().requires_grad_(True)
    w = torch.randn(5, 7, 39, dtype=torch.float64, generator=g)
    grads = []
    for fn in (positional_encoding, per_band):
        (gx,) = torch.autograd.grad((fn(xd, 6) * w).sum(), xd, create_graph=True)
        (ggx,) = torch.autograd.grad((gx * gx).sum(), xd)
        grads.append((gx.detach(), ggx))
    assert torch.allclose(grads[0][0], grads[1][0], rtol=1e-12, atol=1e-12)
    assert torch.allclose(grads[0][1], grads[1][1], rtol=1e-12, atol=1e-12)
