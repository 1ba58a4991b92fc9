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

MI355X: ``compute_dtype=torch.bfloat16`` runs the linear layers under BF16
autocast, i.e. hipBLASLt GEMMs on the MFMA units with FP32 accumulation; the
parameters stay FP32.  The reference runs these networks under FP16 autocast
(rf_rendering.py:740).
"""
from __future__ import annotations

import contextlib
import math
from typing import Optional, Sequence

import torch
import torch.nn as nn


def positional_encoding(x: torch.Tensor, multires: int) -> torch.Tensor:
    """[x, sin(2^i x), cos(2^i x) for i < multires] along the last dim."""
    if multires <= 0:
        return x
    feats = [x]
    for i in range(multires):
        f = float(2.0 ** i)
        feats.append(torch.sin(x * f))
        feats.append(torch.cos(x * f))
    return torch.cat(feats, dim=-1)


def _autocast(dtype: Optional[torch.dtype], device: torch.device):
    if dtype is None or dtype == torch.float32 or device.type != "cuda":
        return contextlib.nullcontext()
    return torch.autocast(device_type="cuda", dtype=dtype)


class SDFNetwork(nn.Module):
    def __init__(self, d_in: int, d_out: int, d_hidden: int, n_layers: int,
                 skip_in: Sequence[int] = (4,), multires: int = 0, bias: float = 0.5,
                 scale: float = 1.0, geometric_init: bool = True, weight_norm: bool = True,
                 inside_outside: bool = False, sdf_activation: str = "none",
                 layer_activation: str = "softplus",
                 compute_dtype: Optional[torch.dtype] = None):
        super().__init__()
        self.multires = multires
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
            self.activation = nn.Softplus(beta=100)
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

    def forward(self, inputs: torch.Tensor) -> torch.Tensor:
        inputs = positional_encoding(inputs * self.scale, self.multires)
        with _autocast(self.compute_dtype, inputs.device):
            x = inputs
            for l in range(self.num_layers - 1):
                if l in self.skip_in:
                    x = torch.cat([x, inputs.to(x.dtype)], dim=-1) / math.sqrt(2.0)
                x = getattr(self, f"lin{l}")(x)
                if l < self.num_layers - 2:
                    x = self.activation(x)
        return x.float()

    def sdf(self, x: torch.Tensor) -> torch.Tensor:
        return self.forward(x)[..., :1]

    def sdf_hidden_appearance(self, x: torch.Tensor) -> torch.Tensor:
        return self.forward(x)

    def gradient(self, x: torch.Tensor) -> torch.Tensor:
        x.requires_grad_(True)
        with torch.enable_grad():
            y = self.sdf(x)
        return torch.autograd.grad(y, x, torch.ones_like(y), create_graph=True,
                                   retain_graph=True, only_inputs=True)[0]

    def sdf_and_gradient(self, x: torch.Tensor):
        """(full output [.., d_out], gradient of the SDF) from one forward."""
        x.requires_grad_(True)
        with torch.enable_grad():
            out = self.forward(x)
            y = out[..., :1]
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
                 refelective_act: str = "sigmoid", compute_dtype: Optional[torch.dtype] = None):
        super().__init__()
        self.refelective_predictor = make_predictor(feats_dim + 3, 1, activation=refelective_act)
        self.light_power = nn.Parameter(torch.tensor(float(light_power)))
        self.compute_dtype = compute_dtype

    def reflectance(self, points: torch.Tensor, features: torch.Tensor) -> torch.Tensor:
        x = torch.cat([features, points], dim=-1)
        with _autocast(self.compute_dtype, x.device):
            r = self.refelective_predictor(x)
        return r.float().squeeze(-1)

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
