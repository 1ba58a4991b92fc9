"""Networks that produce the lensless tracer's inputs (drop-in for the classes
of geraf/models/networks/sdf_network.py and sdf_network_debug.py used by the RF
stages).

* ``SDFNetwork`` (sdf_network.py:78-199): positional encoding (input + sin/cos
  at 2^0..2^(multires-1)), ``n_layers`` hidden layers of ``d_hidden`` with
  Softplus(beta=100), input re-injected at ``skip_in`` (scaled by 1/sqrt 2),
  NeuS geometric initialisation, weight normalisation; ``gradient(x)`` is the
  autograd gradient of the SDF w.r.t. x with ``create_graph=True`` (the
  normals, differentiable again for the training step).
* ``SingleVarianceNetwork`` (sdf_network.py:339-359): inv_s = exp(10 * variance).
* ``ReflectivePowerNetwork`` (sdf_network.py:450-485): reflectance from a
  4-layer, 256-wide predictor on [feature, point] (field.py:21-59), times
  exp(light_power) times the transmittance, clamped at 1e-6.
* ``FixedPowerNetwork`` (sdf_network_debug.py:10-26).

MI355X: ``compute_dtype=torch.bfloat16`` runs the layers in BF16 (GEMMs on
the MFMA units with FP32 accumulation, activations stored in BF16); the
parameters stay FP32 and are cast once per forward.  The reference runs these
networks under FP16 autocast (rf_rendering.py:740).  Every layer product goes
through ``tall_matmul``: the batch of points is tall ([M, 256] with M ~ 5e5)
and the weights are small, and the weight gradient A^T G of such a product is
a GEMM with only (256/64)^2 output tiles when handed to the BLAS as one call.
``tall_matmul`` forms it as a split-K batched product instead, and does so
again in the double backward the SDF gradient needs.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import torch
import torch.nn as nn

_SPLITS = 32


def _splits(m: int) -> int:
    s = _SPLITS
    while s > 1 and m % s:
        s //= 2
    return s


def col_sum(x: torch.Tensor) -> torch.Tensor:
    """Column sums of a tall [M, q] tensor, accumulated in at least FP32:
    partial sums over up to 512 row blocks, then over the blocks.  (A one-pass
    BF16 sum over dim 0 takes milliseconds when q is not a multiple of 8.)"""
    m, q = x.shape
    s = 512
    while s > 1 and m % s:
        s //= 2
    acc = torch.float64 if x.dtype == torch.float64 else torch.float32
    return x.reshape(s, m // s, q).sum(1, dtype=acc).sum(0).to(x.dtype)


def _tn(a: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """a^T g for tall a [M, p], g [M, q]: split-K over M, then a sum.  A single
    column on either side is a broadcast product and a column sum (the BLAS
    takes milliseconds over such a BF16 product)."""
    if g.shape[1] == 1:
        return col_sum(a * g).unsqueeze(1)
    if a.shape[1] == 1:
        return col_sum(a * g).unsqueeze(0)
    m = a.shape[0]
    s = _splits(m)
    if s == 1:
        return a.t() @ g
    return torch.bmm(a.reshape(s, m // s, a.shape[1]).transpose(1, 2),
                     g.reshape(s, m // s, g.shape[1])).sum(0)


class _TallMatmul(torch.autograd.Function):
    """C = A @ B, A [M, p] tall, B [p, q] small."""

    @staticmethod
    def forward(ctx, a, b):
        ctx.save_for_backward(a, b)
        return a @ b

    @staticmethod
    def backward(ctx, gc):
        a, b = ctx.saved_tensors
        ga = tall_matmul(gc, b.t()) if ctx.needs_input_grad[0] else None
        gb = tall_tn(a, gc) if ctx.needs_input_grad[1] else None
        return ga, gb


class _TallTN(torch.autograd.Function):
    """C = A^T G, A [M, p] and G [M, q] tall; C [p, q] small."""

    @staticmethod
    def forward(ctx, a, g):
        ctx.save_for_backward(a, g)
        return _tn(a, g)

    @staticmethod
    def backward(ctx, gc):
        a, g = ctx.saved_tensors
        ga = tall_matmul(g, gc.t()) if ctx.needs_input_grad[0] else None
        gg = tall_matmul(a, gc) if ctx.needs_input_grad[1] else None
        return ga, gg


def tall_matmul(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """A @ B for a tall A; differentiable to any order."""
    return _TallMatmul.apply(a, b)


def tall_tn(a: torch.Tensor, g: torch.Tensor) -> torch.Tensor:
    """A^T G for tall A and G; differentiable to any order."""
    return _TallTN.apply(a, g)


def linear(x: torch.Tensor, lin: nn.Linear, dtype: Optional[torch.dtype]) -> torch.Tensor:
    """x @ W^T + b in `dtype` (None: the parameters' dtype) for a batch of
    points x [M, in]."""
    w = lin.weight if dtype is None else lin.weight.to(dtype)
    if lin.bias is None:
        return tall_matmul(x, w.t())
    return _TallLinear.apply(x, w.t(), lin.bias if dtype is None else lin.bias.to(dtype))


class _TallLinear(torch.autograd.Function):
    """x @ B + b for a tall x, with the bias added in the GEMM's epilogue
    (one pass over the output instead of a product and a separate add); the
    backward is the tall products and a blocked column sum, differentiable to
    any order."""

    @staticmethod
    def forward(ctx, x, b_mat, bias):
        ctx.save_for_backward(x, b_mat)
        return torch.addmm(bias, x, b_mat)

    @staticmethod
    def backward(ctx, gy):
        x, b_mat = ctx.saved_tensors
        gx = tall_matmul(gy, b_mat.t()) if ctx.needs_input_grad[0] else None
        gb_mat = tall_tn(x, gy) if ctx.needs_input_grad[1] else None
        gbias = col_sum(gy) if ctx.needs_input_grad[2] else None
        return gx, gb_mat, gbias


class _SoftplusBF16(torch.autograd.Function):
    """Softplus of a CUDA BF16 tensor by the library's kernel (one pass, FP32
    arithmetic); its backward is differentiable once more."""

    @staticmethod
    def forward(ctx, x, beta, threshold):
        from geraf.ops import _native

        xc = x.contiguous()
        y = torch.empty_like(xc)
        _native.call("geraf_softplus_bf16", _native.ptr(xc), _native.ptr(y), xc.numel(), beta,
                     threshold, _native.stream_of(xc))
        # the input itself (not a copy): the double backward reaches it through
        # _SoftplusGradBF16
        ctx.save_for_backward(x)
        ctx.beta, ctx.threshold = beta, threshold
        return y

    @staticmethod
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        return _SoftplusGradBF16.apply(gy, x, ctx.beta, ctx.threshold), None, None


class _SoftplusGradBF16(torch.autograd.Function):
    """gx = g * s(x), s = dsoftplus/dx; its adjoint (both inputs) is one fused
    kernel.  A third derivative is not provided."""

    @staticmethod
    def forward(ctx, g, x, beta, threshold):
        from geraf.ops import _native

        gc = g.to(torch.bfloat16).contiguous()
        xc = x.contiguous()
        gx = torch.empty_like(xc)
        _native.call("geraf_softplus_grad_bf16", _native.ptr(gc), _native.ptr(xc),
                     _native.ptr(gx), xc.numel(), beta, threshold, _native.stream_of(xc))
        ctx.save_for_backward(gc, xc)
        ctx.beta, ctx.threshold = beta, threshold
        return gx.view(x.shape)

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, ggx):
        from geraf.ops import _native

        g, x = ctx.saved_tensors
        ggx = ggx.to(torch.bfloat16).contiguous()
        out_g = torch.empty_like(x)
        out_x = torch.empty_like(x)
        _native.call("geraf_softplus_grad2_bf16", _native.ptr(ggx), _native.ptr(g), _native.ptr(x),
                     _native.ptr(out_g), _native.ptr(out_x), x.numel(), ctx.beta, ctx.threshold,
                     _native.stream_of(x))
        return out_g, out_x, None, None


class SoftplusBF16(nn.Module):
    """nn.Softplus(beta, threshold); CUDA BF16 inputs go through the fused
    kernels (forward, backward and double backward in one pass each), other
    inputs through torch."""

    def __init__(self, beta: float = 1.0, threshold: float = 20.0):
        super().__init__()
        self.beta, self.threshold = float(beta), float(threshold)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        if x.is_cuda and x.dtype == torch.bfloat16:
            return _SoftplusBF16.apply(x, self.beta, self.threshold)
        return nn.functional.softplus(x, self.beta, self.threshold)


def positional_encoding(x: torch.Tensor, multires: int) -> torch.Tensor:
    """[x, sin(2^i x), cos(2^i x) for i < multires] along the last dim, the
    order of the reference's Embedder (sdf_network.py:19-75: include_input,
    log-sampled bands 2^0 .. 2^(multires-1), sin then cos per band).

    All bands at once: one product with the [multires, 1] band column, one
    sine, one cosine and one concatenation, instead of two products, two
    transcendentals and a gradient accumulation per band.  The products by
    powers of two are exact, so the features equal the per-band form bit for
    bit."""
    if multires <= 0:
        return x
    bands = 2.0 ** torch.arange(multires, dtype=x.dtype, device=x.device)
    xf = x.unsqueeze(-2) * bands.unsqueeze(-1)                     # [..., L, d]
    enc = torch.stack((torch.sin(xf), torch.cos(xf)), dim=-2)     # [..., L, 2, d]
    return torch.cat((x, enc.flatten(-3)), dim=-1)


class SDFNetwork(nn.Module):
    def __init__(self, d_in: int, d_out: int, d_hidden: int, n_layers: int,
                 skip_in: Sequence[int] = (4,), multires: int = 0, bias: float = 0.5,
                 scale: float = 1.0, geometric_init: bool = True, weight_norm: bool = True,
                 inside_outside: bool = False, sdf_activation: str = "none",
                 layer_activation: str = "softplus",
                 compute_dtype: Optional[torch.dtype] = None, tall_products: bool = True):
        super().__init__()
        self.multires = multires
        self.tall_products = tall_products
        in_ch = d_in * (1 + 2 * multires) if multires > 0 else d_in
        dims = [in_ch] + [d_hidden] * n_layers + [d_out]
        self.num_layers = len(dims)
        self.skip_in = tuple(skip_in)
        self.scale = scale
        self.compute_dtype = compute_dtype
        for l in range(self.num_layers - 1):
            out_dim = dims[l + 1] - dims[0] if (l + 1) in self.skip_in else dims[l + 1]
            lin = nn.Linear(dims[l], out_dim)
            if geometric_init:
                self._geometric_init(lin, l, dims, out_dim, bias, inside_outside, multires)
            if weight_norm:
                lin = nn.utils.parametrizations.weight_norm(lin)
            setattr(self, f"lin{l}", lin)
        if layer_activation == "softplus":
            self.activation = SoftplusBF16(beta=100)
        elif layer_activation == "relu":
            self.activation = nn.ReLU()
        else:
            raise NotImplementedError(layer_activation)

    def _geometric_init(self, lin, l, dims, out_dim, bias, inside_outside, multires):
        with torch.no_grad():
            if l == self.num_layers - 2:
                sign = -1.0 if inside_outside else 1.0
                nn.init.normal_(lin.weight, mean=sign * math.sqrt(math.pi) / math.sqrt(dims[l]),
                                std=1e-4)
                nn.init.constant_(lin.bias, sign * -bias)
            elif multires > 0 and l == 0:
                nn.init.constant_(lin.bias, 0.0)
                nn.init.constant_(lin.weight[:, 3:], 0.0)
                nn.init.normal_(lin.weight[:, :3], 0.0, math.sqrt(2) / math.sqrt(out_dim))
            elif multires > 0 and l in self.skip_in:
                nn.init.constant_(lin.bias, 0.0)
                nn.init.normal_(lin.weight, 0.0, math.sqrt(2) / math.sqrt(out_dim))
                nn.init.constant_(lin.weight[:, -(dims[0] - 3):], 0.0)
            else:
                nn.init.constant_(lin.bias, 0.0)
                nn.init.normal_(lin.weight, 0.0, math.sqrt(2) / math.sqrt(out_dim))

    def forward(self, inputs: torch.Tensor, float_out: bool = True) -> torch.Tensor:
        """[.., d_out] = (sdf, features); in FP32, or with float_out=False in
        the compute dtype (a BF16 network then skips the cast of its whole
        output)."""
        inputs = positional_encoding(inputs * self.scale, self.multires)
        dt = self.compute_dtype if self.compute_dtype not in (None, torch.float32) else None
        shape = inputs.shape[:-1]
        xin = inputs.reshape(-1, inputs.shape[-1])
        if dt is not None:
            xin = xin.to(dt)
        x = xin
        for l in range(self.num_layers - 1):
            if l in self.skip_in:
                x = torch.cat([x, xin], dim=-1) / math.sqrt(2.0)
            lin = getattr(self, f"lin{l}")
            x = linear(x, lin, dt) if self.tall_products else lin(x)
            if l < self.num_layers - 2:
                x = self.activation(x)
        return (x.float() if float_out else x).reshape(*shape, -1)

    def sdf(self, x: torch.Tensor) -> torch.Tensor:
        return self.forward(x)[..., :1]

    def sdf_value(self, x: torch.Tensor) -> torch.Tensor:
        """The SDF alone, [...] in FP32: the output stays in the compute dtype
        and only its first column is converted (no FP32 copy of the features)."""
        return self.forward(x, float_out=False)[..., 0].float()

    def sdf_hidden_appearance(self, x: torch.Tensor) -> torch.Tensor:
        return self.forward(x)

    def gradient(self, x: torch.Tensor) -> torch.Tensor:
        x.requires_grad_(True)
        with torch.enable_grad():
            y = self.sdf(x)
        return torch.autograd.grad(y, x, torch.ones_like(y), create_graph=True,
                                   retain_graph=True, only_inputs=True)[0]

    def sdf_and_gradient(self, x: torch.Tensor, float_out: bool = True):
        """(full output [.., d_out], gradient of the SDF) from one forward.
        With float_out=False the output stays in the compute dtype; the SDF
        column is still differentiated in FP32."""
        x.requires_grad_(True)
        with torch.enable_grad():
            out = self.forward(x, float_out=float_out)
            y = out[..., :1].float()
        g = torch.autograd.grad(y, x, torch.ones_like(y), create_graph=True, retain_graph=True,
                                only_inputs=True)[0]
        return out, g


class SingleVarianceNetwork(nn.Module):
    def __init__(self, init_val: float, activation: str = "exp"):
        super().__init__()
        self.act = activation
        self.variance = nn.Parameter(torch.tensor(float(init_val)))

    def inv_s(self) -> torch.Tensor:
        if self.act == "exp":
            return torch.exp(self.variance * 10.0)
        if self.act == "linear":
            return self.variance * 10.0
        if self.act == "square":
            return (self.variance * 10.0) ** 2
        raise NotImplementedError(self.act)

    def forward(self, x: torch.Tensor) -> torch.Tensor:
        return torch.ones([*x.shape[:-1], 1], device=x.device) * self.inv_s()


def make_predictor(feats_dim: int, output_dim: int, weight_norm: bool = True,
                   activation: str = "sigmoid") -> nn.Sequential:
    """Three 256-wide ReLU layers and an output layer (field.py:21-59)."""
    act = {"sigmoid": nn.Sigmoid(), "softplus": nn.Softplus(), "relu": nn.ReLU(),
           "none": nn.Identity()}[activation]
    wn = nn.utils.parametrizations.weight_norm if weight_norm else (lambda m: m)
    return nn.Sequential(wn(nn.Linear(feats_dim, 256)), nn.ReLU(), wn(nn.Linear(256, 256)),
                         nn.ReLU(), wn(nn.Linear(256, 256)), nn.ReLU(),
                         wn(nn.Linear(256, output_dim)), act)


class ReflectivePowerNetwork(nn.Module):
    def __init__(self, feats_dim: int = 256, light_power: float = 0.0,
                 refelective_act: str = "sigmoid", compute_dtype: Optional[torch.dtype] = None,
                 tall_products: bool = True):
        super().__init__()
        self.refelective_predictor = make_predictor(feats_dim + 3, 1, activation=refelective_act)
        self.light_power = nn.Parameter(torch.tensor(float(light_power)))
        self.compute_dtype = compute_dtype
        self.tall_products = tall_products

    def reflectance(self, points: torch.Tensor, features: torch.Tensor) -> torch.Tensor:
        dt = self.compute_dtype if self.compute_dtype not in (None, torch.float32) else None
        # features may already be in the compute dtype (no cast of them then)
        x = torch.cat([features, points.to(features.dtype)], dim=-1)
        shape = x.shape[:-1]
        x = x.reshape(-1, x.shape[-1])
        if dt is not None:
            x = x.to(dt)
        for m in self.refelective_predictor:
            x = linear(x, m, dt) if self.tall_products and isinstance(m, nn.Linear) else m(x)
        return x.float().reshape(shape)

    def power(self) -> torch.Tensor:
        return torch.exp(self.light_power)

    def forward(self, points, feature_vectors, trans_prob):
        color = self.reflectance(points, feature_vectors) * self.power() * trans_prob
        return torch.clamp(color, min=1e-6)


class FixedPowerNetwork(nn.Module):
    def __init__(self, light_power: float = 0.0):
        super().__init__()
        self.light_power = nn.Parameter(torch.tensor(float(light_power)))

    def power(self) -> torch.Tensor:
        return torch.exp(self.light_power)

    def forward(self, points, feature_vectors, trans_prob):
        return torch.clamp(self.power() * trans_prob, min=1e-6)
