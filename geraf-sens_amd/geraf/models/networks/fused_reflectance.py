"""The reflectance network's forward and backward on the library's BF16
matrix-core kernels (csrc/mlp_gemm_kernels.hip).

Network (sdf_network.py:450-485 with field.py:21-59, the RF stages'
ReflectivePowerNetwork): r = sigmoid(W_D a_C + b_D), a_C = relu(W_C a_B +
b_C), a_B = relu(W_B a_A + b_B), a_A = relu(W_A x_0 + b_A), x_0 = [features
(256), points (3)].

Forward: three layer products with the ReLU epilogue and the one-row output
product (its sigmoid in FP32).  The input is read in place from rows of at
least 264 elements: the features in columns 0..255 and the points in 257..259,
with column 256 multiplied by a zero weight.  The SDF network's output
buffer has exactly that layout ([features | sdf | padding]), so features
coming from it need no concatenation: the points are written into its
padding.

Backward, for d r: g_D = d r * r (1 - r); g_C = (g_D w_D) * [a_C > 0];
g_B = (g_C W_C) * [a_B > 0]; g_A = (g_B W_B) * [a_A > 0]; d features =
(g_A W_A)[:, :256] -- one chained sweep (the ReLU masks read from the
stored activations, g_C formed and stored by its prologue); the weight
gradients dW_C = g_C^T a_B, dW_B = g_B^T a_A, dW_A = g_A^T x_0, dw_D =
g_D^T a_C with their bias sums.  Gradients reach the weight-norm
parameters through torch autograd of W = g v / |v|.
"""
from __future__ import annotations

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from geraf.models.networks import fused_sdf as fs
from geraf.ops import _native

BF = torch.bfloat16
PTS = 257   # first column of the points (the input rows: features, one
            # unused column, the points, padding to 264)


def _layers(net):
    seq = net.refelective_predictor
    lins = [seq[0], seq[2], seq[4], seq[6]]
    shapes = [(256, 259), (256, 256), (256, 256), (1, 256)]
    if len(seq) != 8 or any(tuple(l.weight.shape) != s for l, s in zip(lins, shapes)) or \
            not isinstance(seq[7], torch.nn.Sigmoid):
        raise ValueError("the fused reflectance network needs 259 -> 256 -> 256 -> 256 -> 1 "
                         "layers with ReLU and a sigmoid output")
    return lins


def _input_rows(points: torch.Tensor, features: torch.Tensor) -> torch.Tensor:
    """[M, 264] BF16 rows [features | - | points | -]: the features' own
    buffer when they are the first 256 columns of wider rows, else a copy."""
    M = points.shape[0]
    L = fs.wide_rows(features, M)
    if L:
        x0 = torch.as_strided(features.detach(), (M, 264), (L, 1))
    else:
        x0 = torch.zeros((M, fs.OUT_ROW), dtype=BF, device=points.device)[:, :264]
        x0[:, :256] = features.detach()
    x0[:, PTS:PTS + 3] = points.detach().to(BF)
    return x0


class _FusedReflectance(Function):
    @staticmethod
    def forward(ctx, points, features, *params):
        ws, bs = [w.detach().float() for w in params[0::2]], \
            [b.detach().float().contiguous() for b in params[1::2]]
        M = points.shape[0]
        dev = points.device
        x0 = _input_rows(points, features)
        # W_A with its point columns moved to 257..259 (column 256 zero)
        wa = torch.zeros((256, 260), dtype=torch.float32, device=dev)
        wa[:, :256] = ws[0][:, :256]
        wa[:, PTS:PTS + 3] = ws[0][:, 256:]
        P = [fs.pack(wa, False), fs.pack(ws[1], False), fs.pack(ws[2], False),
             fs.pack(ws[3], False)]
        a = torch.empty((3, M, 256), dtype=BF, device=dev)
        fs.gemm(x0, P[0], 256, 260, a[0], bias=bs[0], epilogue=fs._EPI_RELU, M=M)
        fs.gemm(a[0], P[1], 256, 256, a[1], bias=bs[1], epilogue=fs._EPI_RELU)
        fs.gemm(a[1], P[2], 256, 256, a[2], bias=bs[2], epilogue=fs._EPI_RELU)
        z = torch.empty((M, 8), dtype=BF, device=dev)
        fs.gemm(a[2], P[3], 1, 256, z, bias=bs[3])
        r = torch.sigmoid(z[:, 0].float())
        ctx.save_for_backward(*params)
        ctx.state = (x0, a, r, ws)
        return r

    @staticmethod
    @once_differentiable
    def backward(ctx, d_r):
        x0, a, r, ws = ctx.state
        ctx.state = None
        M = x0.shape[0]
        dev = x0.device
        g_d = (d_r.float() * r * (1.0 - r)).contiguous()
        # W_C^T, W_B^T and W_A^T's feature rows, back to back
        n = int(_native.load().geraf_mlp_pack_bytes(256, 256))
        rev = torch.empty(3 * n, dtype=torch.uint8, device=dev)
        for s, w in enumerate((ws[2], ws[1], ws[0][:, :256].contiguous())):
            fs.pack(w, True, out=rev[s * n:(s + 1) * n])
        w_d = ws[3].to(BF).expand(M, 256)
        g_c = torch.empty((M, 256), dtype=BF, device=dev)
        # g_B, g_A, d features in rows of OUT_ROW: the SDF network's backward
        # takes d features as the first columns of its output-layer gradient
        g = torch.empty((3, M, fs.OUT_ROW), dtype=BF, device=dev)[:, :, :256]
        if M > 0:
            fs.sweep(w_d, rev, 3, g[0], M * fs.OUT_ROW, e0=a[1], e0_step=-M * 256, x2=a[2],
                     x_scale=g_d, relu=True, mask_steps=2, x_out=g_c)
        dws = [torch.zeros_like(w) for w in ws]
        dbs = [torch.zeros(w.shape[0], dtype=torch.float32, device=dev) for w in ws]
        dwa = torch.zeros((256, 260), dtype=torch.float32, device=dev)
        fs.wgrad(g_c, a[1], 256, 256, dws[2], db=dbs[2])
        fs.wgrad(g[0], a[0], 256, 256, dws[1], db=dbs[1])
        fs.wgrad(g[1], x0, 256, 260, dwa, db=dbs[0])
        dws[0][:, :256] = dwa[:, :256]
        dws[0][:, 256:] = dwa[:, PTS:PTS + 3]
        gd8 = torch.zeros((M, 8), dtype=BF, device=dev)
        gd8[:, 0] = g_d.to(BF)
        fs.wgrad(gd8, a[2], 1, 256, dws[3])
        dbs[3][0] = g_d.sum()
        grads = []
        for dw, db in zip(dws, dbs):
            grads += [dw, db]
        return (None, g[2]) + tuple(grads)


def fused_reflectance(net, points: torch.Tensor, features: torch.Tensor) -> torch.Tensor:
    """r [M] FP32 of a ReflectivePowerNetwork's reflectance predictor for
    points [M,3] and features [M,256] (as net.reflectance computes it)."""
    params = []
    for lin in _layers(net):
        params += [lin.weight, lin.bias]
    return _FusedReflectance.apply(points.reshape(-1, 3), features, *params)
