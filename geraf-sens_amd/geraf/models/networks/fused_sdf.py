"""The SDF network's forward, its analytic gradient (the normals) and the
hand-derived backward of both, on the library's BF16 matrix-core layer
products (csrc/mlp_gemm_kernels.hip).

Network (sdf_network.py:78-199, the configuration the RF stages use): the
positional encoding gamma(x) (39 features), hidden layers l = 0..7 with
z_l = W_l a_l + b_l, h_l = softplus_100(z_l), a_0 = gamma, a_4 =
[h_3, gamma] / sqrt 2 (the skip; h_3 has 217 features), a_l = h_{l-1}
otherwise, and the output layer [sdf, features] = W_8 h_7 + b_8.

Normals, reverse mode by hand: u_7 = W_8[0], u_{l-1} = (u_l * s_l) W_l for
l = 7..1 (s_l the softplus slope at z_l; layer 4's product splits into u_3
and the skip's share of d sdf / d gamma), d sdf / d gamma = (u_0 * s_0) W_0 +
the skip share, then through the encoding's Jacobian.  One product per layer;
each stores g_{l-1} = u_{l-1} * s_{l-1} (its epilogue applies the slope of the
layer below), the form every later use needs.

Backward, for incoming gradients (d sdf, d features, d normals):
  * the adjoint of the normal pass (layer 0 upward): d gamma = J^T d n;
    vbar_0 = d gamma W_0^T, and for l = 0..6 vbar_{l+1} = (vbar_l * s_l)
    W_{l+1}^T, with dW_{l+1} += (u_{l+1} s_{l+1})^T (vbar_l s_l), dW_0 +=
    (u_0 s_0)^T d gamma, dW_8[0] += sum (vbar_7 s_7); each layer leaves the
    second-order term t_l = vbar_l * u_l * s'_l for the primal adjoint.  The
    products store vbar_l s_l (and t_l = vbar_l (u_l s_l) beta (1 - s_l)), so
    the weight gradients read two tensors and no slope codes;
  * the primal adjoint (layer 8 downward): zbar_8 = (d sdf, d features),
    zbar_{l-1} = (zbar_l W_l) * s_{l-1} + t_{l-1}, dW_l += zbar_l^T a_l,
    db_l = sum zbar_l.
Each step is one product with its element-wise work fused in (the kernels'
prologues and epilogues).  Activations are stored in BF16 as the torch BF16
network stores them; in place of each pre-activation z_l the forward stores
its softplus slope code q_l = copysign(sigmoid(-|beta z_l|), beta z_l) in
FP16, computed from the FP32 accumulator (the backward needs z_l only
through the slope s_l and the curvature s_l (1 - s_l), both exact functions
of q_l).  The skip's 1/sqrt 2 is folded into W_4.  Gradients reach the
weight-norm parameters through torch autograd of W_l = g v / |v|.
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from geraf.ops import _native

BF = torch.bfloat16
CODE = torch.float16  # softplus slope codes
BETA, THRESHOLD = 100.0, 20.0
_PRO_NONE, _PRO_SLOPE = 0, 1
(_EPI_BIAS, _EPI_SOFTPLUS, _EPI_SOFTPLUS_TRAIN, _EPI_RELU, _EPI_GRAD_ADJ, _EPI_LAYER_ADJ,
 _EPI_MUL_SLOPE, _EPI_GRAD_ADJ_S, _EPI_GRAD_ADJ_SU) = range(9)
_SLOPE_ONE = 0.0  # the slope code of slope 1, curvature 0


def _ptr(t: Optional[torch.Tensor]):
    return _native.ptr(t)


def _ld(t: Optional[torch.Tensor]) -> int:
    return 0 if t is None else (0 if t.stride(0) == 0 else t.stride(0))


def pack(w: torch.Tensor, transpose: bool, scale: float = 1.0,
         out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Operand-order BF16 copy of W [N,K] (or of W^T with transpose), into
    ``out`` (a uint8 slice of the packed size) when given."""
    w = w.detach().float().contiguous()
    N, K = (w.shape[1], w.shape[0]) if transpose else (w.shape[0], w.shape[1])
    size = int(_native.load().geraf_mlp_pack_bytes(N, K))
    if size < 0:
        raise ValueError(f"layer of {N} x {K} is beyond the kernels' limits")
    if out is None:
        out = torch.empty(size, dtype=torch.uint8, device=w.device)
    elif out.numel() != size:
        raise ValueError("pack: out has the wrong size")
    _native.call("geraf_mlp_pack_weights", _ptr(w), w.stride(0), int(transpose), float(scale),
                 N, K, _ptr(out), _native.stream_of(w))
    return out


def gemm(x, w_packed, N, K, y, bias=None, x2=None, y2=None, e0=None, e1=None,
         prologue=_PRO_NONE, epilogue=_EPI_BIAS, M=None) -> None:
    """y = epilogue(prologue(x) W^T + bias) into the (possibly strided) y."""
    M = x.shape[0] if M is None else M
    _native.call("geraf_mlp_gemm_bf16", _ptr(x), _ld(x), _ptr(x2), _ld(x2), _ptr(w_packed),
                 _ptr(bias), _ptr(y), _ld(y), _ptr(y2), _ld(y2), _ptr(e0), _ld(e0), _ptr(e1),
                 _ld(e1), M, N, K, prologue, epilogue, BETA, THRESHOLD, _native.stream_of(y))


class _Workspace:
    def __init__(self):
        self.buf = None

    def get(self, device, nbytes: int) -> torch.Tensor:
        if self.buf is None or self.buf.numel() < nbytes or self.buf.device != device:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=device)
        return self.buf


_WS = _Workspace()


def wgrad(g, x, N, K, dw, db=None, g2=None, x2=None, accumulate=False, M=None) -> None:
    """dw[N,K] (+)= prologue(g)^T prologue(x); db (+)= column sums."""
    M = (g.shape[0] if g.stride(0) else x.shape[0]) if M is None else M
    nbytes = int(_native.load().geraf_mlp_wgrad_workspace_bytes(M, N, K))
    ws = _WS.get(dw.device, nbytes)
    _native.call("geraf_mlp_wgrad_bf16", _ptr(g), _ld(g), _ptr(g2), _ld(g2),
                 _PRO_SLOPE if g2 is not None else _PRO_NONE, _ptr(x), _ld(x), _ptr(x2), _ld(x2),
                 _PRO_SLOPE if x2 is not None else _PRO_NONE, _ptr(dw), dw.stride(0), _ptr(db),
                 int(accumulate), M, N, K, BETA, THRESHOLD, _ptr(ws), ws.numel(),
                 _native.stream_of(dw))


# layer l: (fan-in, fan-out) of the RF stages' SDF network
IN = [39, 256, 256, 256, 256, 256, 256, 256, 256]
OUT = [256, 256, 256, 217, 256, 256, 256, 256, 257]
SKIP = 4
NL = 9
# row pitch of the output layer's [features | sdf | padding] rows and of its
# gradient: 640 bytes, a multiple of 128, so that the products' 64-byte row
# segments stay aligned (rows of 264, 528 bytes, made them straddle lines)
OUT_ROW = 320
L_BANDS = 6
SQRT2 = math.sqrt(2.0)


def _check_network(ws: List[torch.Tensor]) -> None:
    for l, w in enumerate(ws):
        if tuple(w.shape) != (OUT[l], IN[l]):
            raise ValueError(f"layer {l}: weight {tuple(w.shape)}, the fused network needs "
                             f"{(OUT[l], IN[l])}")


def _out_rows(t: torch.Tensor) -> torch.Tensor:
    """The output layer's rows in the kernels' order: the 256 feature rows,
    then the sdf row (the reference's row 0).  The layer's output and its
    gradient are laid out [features | sdf] so that the feature block starts
    16-byte aligned."""
    return torch.cat([t[1:], t[:1]])


def _hidden_pack(ws):
    """The hidden layers' packed W_0 .. W_7 back to back (the chained
    forward's operand) and the slice of each (the layer products')."""
    lib = _native.load()
    total = int(lib.geraf_mlp_sdf_chain_pack_bytes(L_BANDS))
    buf = torch.empty(total, dtype=torch.uint8, device=ws[0].device)
    views, o = [], 0
    for l in range(NL - 1):
        n = int(lib.geraf_mlp_pack_bytes(OUT[l], IN[l]))
        views.append(pack(ws[l], False, 1.0 / SQRT2 if l == SKIP else 1.0, out=buf[o:o + n]))
        o += n
    if o != total:
        raise RuntimeError("hidden-layer pack sizes disagree with the chained forward's")
    return buf, views


def _reverse_pack(ws):
    """W_7^T .. W_1^T packed back to back (the chained reverse sweeps'
    operand; W_3^T's 217 inputs padded to 256 with zero weights, which adds
    exact zeros) and the slice of each."""
    n = int(_native.load().geraf_mlp_pack_bytes(256, 256))
    buf = torch.empty(7 * n, dtype=torch.uint8, device=ws[0].device)
    views = [None] * (NL - 1)
    for s, l in enumerate(range(NL - 2, 0, -1)):
        w = ws[l].detach().float()
        if OUT[l] < 256:
            w = torch.cat([w, w.new_zeros((256 - OUT[l], IN[l]))])
        views[l] = pack(w, True, 1.0 / SQRT2 if l == SKIP else 1.0, out=buf[s * n:(s + 1) * n])
    return buf, views


def _packs(ws):
    """P_l packs W_l ([out, in]; the forward and the normal-pass adjoint),
    Q_l packs W_l^T ([in, out]; the normal pass and the primal adjoint);
    layer 4 carries the skip's 1/sqrt 2, layer 8 its rows in _out_rows
    order.  Also returns the buffers holding P_0 .. P_7 and Q_7 .. Q_1."""
    hidden, P = _hidden_pack(ws)
    P.append(pack(_out_rows(ws[NL - 1].detach()), False))
    rev, Q = _reverse_pack(ws)
    Q.append(pack(_out_rows(ws[NL - 1].detach()), True))
    return P, Q, hidden, rev


def sweep(x, wpacked, steps, y, y_step, e0=None, e0_step=0, e1=None, e1_step=0, x2=None,
          x_scale=None, relu=False, mask_steps=None, x_out=None) -> None:
    """Chained layer products (geraf_mlp_chain_adjoint_bf16): y, e0 and e1
    are the first step's [M, 256] views, each step moving by *_step
    elements; x (with the prologue slopes x2, optionally scaled per point by
    x_scale) the first operand."""
    M = y.shape[0]
    _native.call("geraf_mlp_chain_adjoint_bf16", _ptr(x), _ld(x), _ptr(x2), _ld(x2),
                 _ptr(x_scale), int(relu), _ptr(wpacked), steps,
                 steps if mask_steps is None else mask_steps, _ptr(e0), _ld(e0), e0_step,
                 _ptr(e1), _ld(e1), e1_step, _ptr(y), _ld(y), y_step, _ptr(x_out), _ld(x_out),
                 M, _native.stream_of(y))


def _sweep(x, rev, e0, e1, y, x2=None) -> None:
    """The chained reverse sweep over layers 7 .. 1: e0, e1 (or None) and y
    are [7, M, 256] tensors indexed by the output's layer (step s writes
    layer 6 - s); x (and the prologue codes x2) the first operand."""
    M = y.shape[1]
    step = -M * 256
    sweep(x, rev, NL - 2, y[6], step, e0=e0[6], e0_step=step,
          e1=None if e1 is None else e1[6], e1_step=step, x2=x2)


def _biases(params) -> List[torch.Tensor]:
    bs = [b.detach().float().contiguous() for b in params]
    bs[NL - 1] = _out_rows(bs[NL - 1]).contiguous()
    return bs


def _bias_table(bs) -> torch.Tensor:
    """The hidden layers' biases as [8, 256] FP32 (zero past a layer's width)."""
    t = torch.zeros((NL - 1, 256), dtype=torch.float32, device=bs[0].device)
    for l in range(NL - 1):
        t[l, :OUT[l]] = bs[l]
    return t


def _sdf_row(params) -> torch.Tensor:
    """[W_8[0], b_8[0]] (257 FP32): the output layer's sdf row and bias."""
    w8, b8 = params[2 * (NL - 1)], params[2 * (NL - 1) + 1]
    return torch.cat([w8.detach()[0].float(), b8.detach()[:1].float()]).contiguous()


def _chain(x, hidden, table, w_sdf=None, sdf=None, h=None, q=None) -> None:
    M = x.shape[0]
    _native.call("geraf_mlp_sdf_chain_bf16", _ptr(x), M, L_BANDS, 1.0, _ptr(hidden),
                 _ptr(table), _ptr(w_sdf), _ptr(sdf), _ptr(h), 256, 0 if h is None else M * 256,
                 _ptr(q), 256, 0 if q is None else M * 256, BETA, THRESHOLD,
                 _native.stream_of(x))


def _forward(x, P, hidden, table, b8):
    """The training forward: the hidden layers chained in one kernel (h_l and
    the slope codes q_l stored), then the output layer.  Returns (gam, the
    slope codes [8, M, 256], h list, out8 [M, OUT_ROW] = [features | sdf |
    padding])."""
    M = x.shape[0]
    dev = x.device
    gam = torch.empty((M, 64), dtype=BF, device=dev)
    _native.call("geraf_mlp_positional_encoding", _ptr(x), M, L_BANDS, 1.0, _ptr(gam), 64, 0, 64,
                 _native.stream_of(x))
    h = torch.empty((NL - 1, M, 256), dtype=BF, device=dev)
    q = torch.empty((NL - 1, M, 256), dtype=CODE, device=dev)
    _chain(x, hidden, table, h=h, q=q)
    out8 = torch.empty((M, OUT_ROW), dtype=BF, device=dev)
    gemm(h[7], P[8], OUT[8], IN[8], out8, bias=b8)
    return gam, q, list(h.unbind(0)), out8


def _forward_layers(x, ws, bs, P, train: bool):
    """The same forward, one layer product per layer (the chained kernel's
    reference in the tests).  Returns (gam, slope-code list, h list, out8)."""
    M = x.shape[0]
    dev = x.device
    st = _native.stream_of(x)
    gam = torch.empty((M, 64), dtype=BF, device=dev)
    _native.call("geraf_mlp_positional_encoding", _ptr(x), M, L_BANDS, 1.0, _ptr(gam), 64, 0, 64,
                 st)
    qs, hs = [], []
    a = gam
    for l in range(8):
        h = torch.empty((M, 256), dtype=BF, device=dev)
        z = torch.empty((M, 256), dtype=CODE, device=dev) if train else None
        gemm(a, P[l], OUT[l], IN[l], h, bias=bs[l], y2=z,
             epilogue=_EPI_SOFTPLUS_TRAIN if train else _EPI_SOFTPLUS)
        if l == 3:
            # the skip input [h_3, gamma] shares h_3's buffer; q_3's last
            # columns hold the code of slope 1 (the backward passes the skip's
            # gradients through them).  Both after the product, which writes
            # its 217 columns in groups of 4.
            _native.call("geraf_mlp_positional_encoding", _ptr(x), M, L_BANDS, 1.0,
                         _ptr(h[:, 217:]), 256, 0, 39, st)
            if train:
                z[:, 217:].fill_(_SLOPE_ONE)
        qs.append(z)
        hs.append(h)
        a = h
    out8 = torch.empty((M, OUT_ROW), dtype=BF, device=dev)
    gemm(a, P[8], OUT[8], IN[8], out8, bias=bs[8])
    return gam, qs, hs, out8


def _normal_pass(x, ws, q, rev):
    """g_l = u_l s_l for l = 0..6 as one [7, M, 256] tensor (g_3 also carries
    the skip's d sdf / d gamma in columns 217..255, where q_3 holds slope 1),
    from the chained sweep over layers 7 .. 1; u_7 = W_8[0] (one broadcast
    row, not multiplied); the normals."""
    M = x.shape[0]
    dev = x.device
    u7 = torch.zeros((1, 256), dtype=BF, device=dev)
    u7[0] = ws[8].detach()[0].to(BF)
    u7 = u7.expand(M, 256)
    g = torch.empty((NL - 2, M, 256), dtype=BF, device=dev)
    _sweep(u7, rev, q, None, g, x2=q[7])
    us = list(g.unbind(0)) + [u7]
    g0 = torch.empty((M, 40), dtype=BF, device=dev)
    # layer 0's product with W_0^T: the packed W_0 read as [in, out]
    gemm(us[0], pack(ws[0], True), IN[0], OUT[0], g0)
    n = torch.empty((M, 3), dtype=torch.float32, device=dev)
    skip = us[3][:, 217:]
    _native.call("geraf_mlp_pe_gradient", _ptr(x), M, L_BANDS, 1.0, _ptr(g0), 40, _ptr(skip), 256,
                 _ptr(n), _native.stream_of(x))
    return g, us, n


def wide_rows(t: torch.Tensor, M: int) -> int:
    """The row pitch L when t [M,256] BF16 is the first 256 columns of rows
    of L >= 264 elements (a multiple of 8), else 0."""
    if t.dtype != BF or t.dim() != 2 or tuple(t.shape) != (M, 256) or t.stride(1) != 1:
        return 0
    L = t.stride(0)
    ok = L >= 264 and L % 8 == 0 and \
        (t.storage_offset() + M * L) * 2 <= t.untyped_storage().nbytes()
    return L if ok else 0


class _FusedSDF(Function):
    @staticmethod
    def forward(ctx, x, with_normals: bool, *params):
        ws, bs = list(params[0::2]), _biases(params[1::2])
        _check_network(ws)
        x = x.detach().float().contiguous()
        P, Q, hidden, rev = _packs(ws)
        gam, q, hs, out8 = _forward(x, P, hidden, _bias_table(bs), bs[NL - 1])
        sdf = out8[:, 256].float()
        feat = out8[:, :256]
        if with_normals:
            g, us, n = _normal_pass(x, ws, q, rev)
        else:
            g, us = None, None
            n = torch.zeros((x.shape[0], 3), dtype=torch.float32, device=x.device)
        ctx.save_for_backward(*params)
        ctx.state = (x, gam, q, hs, g, us, P, Q, rev, with_normals)
        ctx.mark_non_differentiable(*([] if with_normals else [n]))
        return sdf, feat, n

    @staticmethod
    @once_differentiable
    def backward(ctx, d_sdf, d_feat, d_n):
        params = ctx.saved_tensors
        ws = list(params[0::2])
        x, gam, q, hs, g, us, P, Q, rev, with_normals = ctx.state
        ctx.state = None
        qs = list(q.unbind(0))
        M = x.shape[0]
        dev = x.device
        dws = [torch.zeros((OUT[l], IN[l]), dtype=torch.float32, device=dev) for l in range(NL)]
        dbs = [torch.zeros(OUT[l], dtype=torch.float32, device=dev) for l in range(NL)]
        ts = [None] * 8
        zero_row = torch.zeros((1, 256), dtype=BF, device=dev).expand(M, 256)
        normal_terms = with_normals and d_n is not None
        if normal_terms:
            dn = d_n.float().contiguous()
            dgam = torch.empty((M, 40), dtype=BF, device=dev)
            _native.call("geraf_mlp_pe_gradient_adjoint", _ptr(x), M, L_BANDS, 1.0, _ptr(dn),
                         _ptr(dgam), 40, 0, 40, _native.stream_of(x))
            # dW_0 += (u_0 s_0)^T d gamma (before g_0 = u_0 s_0 is overwritten
            # by t_0); vb holds vbar_l s_l from here on
            wgrad(us[0], dgam, OUT[0], IN[0], dws[0])
            vb = torch.empty((M, 256), dtype=BF, device=dev)
            gemm(dgam, P[0], OUT[0], IN[0], vb, y2=us[0], e0=qs[0], e1=us[0],
                 epilogue=_EPI_GRAD_ADJ_S)
            ts[0] = us[0]
            for l in range(7):
                nxt = l + 1
                if l == 3:
                    # the skip's share: vbar_3's last columns carry d gamma
                    _native.call("geraf_mlp_pe_gradient_adjoint", _ptr(x), M, L_BANDS, 1.0,
                                 _ptr(dn), _ptr(vb[:, 217:]), 256, 0, 39,
                                 _native.stream_of(x))
                u_n = us[nxt]  # g_nxt = u_nxt s_nxt, or the raw broadcast u_7
                vb_n = torch.empty((M, 256), dtype=BF, device=dev)
                if nxt == 7:
                    wgrad(u_n, vb, OUT[nxt], IN[nxt], dws[nxt], g2=qs[nxt], accumulate=True)
                    ts[7] = torch.empty((M, 256), dtype=BF, device=dev)
                    gemm(vb, P[nxt], OUT[nxt], IN[nxt], vb_n, y2=ts[7], e0=qs[nxt], e1=u_n,
                         epilogue=_EPI_GRAD_ADJ_SU)
                else:
                    wgrad(u_n, vb, OUT[nxt], IN[nxt], dws[nxt], accumulate=True)
                    gemm(vb, P[nxt], OUT[nxt], IN[nxt], vb_n, y2=u_n, e0=qs[nxt], e1=u_n,
                         epilogue=_EPI_GRAD_ADJ_S)
                    ts[nxt] = u_n
                vb = vb_n
            # dW_8[0] += sum over the points of vbar_7 s_7
            ones = torch.ones((1, 8), dtype=BF, device=dev).expand(M, 8)
            row = torch.zeros((1, 256), dtype=torch.float32, device=dev)
            wgrad(ones, vb, 1, 256, row)
            dws[8][0] += row[0]
            del vb
        # primal adjoint; the output layer's gradient [d features | d sdf],
        # in place when d features already are the first columns of such rows
        # (the fused reflectance network's backward writes them so)
        L = wide_rows(d_feat, M) if d_feat is not None else 0
        if L:
            zb = torch.as_strided(d_feat, (M, 264), (L, 1))
        else:
            zb = torch.empty((M, OUT_ROW), dtype=BF, device=dev)[:, :264]
            if d_feat is not None:
                zb[:, :256].copy_(d_feat)
            else:
                zb[:, :256].zero_()
        if d_sdf is not None:
            zb[:, 256].copy_(d_sdf)
        else:
            zb[:, 256].zero_()
        zb[:, 257:].zero_()
        # dW_8: the 256 feature rows, then the sdf row
        dw_f = torch.zeros((256, 256), dtype=torch.float32, device=dev)
        db_f = torch.zeros(256, dtype=torch.float32, device=dev)
        wgrad(zb, hs[7], 256, 256, dw_f, db=db_f)
        dw_s = torch.zeros((1, 256), dtype=torch.float32, device=dev)
        db_s = torch.zeros(1, dtype=torch.float32, device=dev)
        wgrad(zb[:, 256:], hs[7], 1, 256, dw_s, db=db_s)
        dws[8][1:] += dw_f
        dws[8][0] += dw_s[0]
        dbs[8][1:] += db_f
        dbs[8][0] += db_s[0]
        # zbar_7 = (zbar_8 W_8) * s_7 + t_7 over t_7, then the chained sweep
        # over layers 7 .. 1: zbar_l over t_l (which holds g_l's buffer), or
        # into a new tensor without the normal terms
        if normal_terms:
            zb7, t7, tail, t_tail = ts[7], ts[7], g, g
        else:
            zb7, t7 = torch.empty((M, 256), dtype=BF, device=dev), zero_row
            tail, t_tail = torch.empty((NL - 2, M, 256), dtype=BF, device=dev), None
        gemm(zb, Q[8], IN[8], OUT[8], zb7, e0=qs[7], e1=t7, epilogue=_EPI_LAYER_ADJ)
        _sweep(zb7, rev, q, t_tail, tail)
        zbs = list(tail.unbind(0)) + [zb7]
        for prev in range(NL - 2, -1, -1):
            a_prev = hs[prev - 1] if prev >= 1 else gam  # layer 4: [h_3, gamma]
            wgrad(zbs[prev], a_prev, OUT[prev], IN[prev], dws[prev], db=dbs[prev],
                  accumulate=True)
        dws[SKIP] *= 1.0 / SQRT2
        grads = []
        for l in range(NL):
            grads += [dws[l], dbs[l]]
        return (None, None) + tuple(grads)


def fused_sdf(net, x: torch.Tensor, with_normals: bool = True
              ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """(sdf [M] FP32, features [M,256] BF16, normals d sdf / d x [M,3] FP32) of
    an SDFNetwork (the RF stages' configuration) for points x [M,3]."""
    params = []
    for l in range(NL):
        lin = getattr(net, f"lin{l}")
        params += [lin.weight, lin.bias]
    return _FusedSDF.apply(x, with_normals, *params)


@torch.no_grad()
def fused_sdf_value(net, x: torch.Tensor) -> torch.Tensor:
    """The SDF alone [M] FP32, no gradients (the calibration's evaluations):
    the chained hidden layers and the sdf row, nothing stored per layer."""
    params = []
    for l in range(NL):
        lin = getattr(net, f"lin{l}")
        params += [lin.weight, lin.bias]
    ws = list(params[0::2])
    _check_network(ws)
    x = x.detach().float().reshape(-1, 3).contiguous()
    hidden, _ = _hidden_pack(ws)
    table = _bias_table([b.detach().float() for b in params[1::2]])
    sdf = torch.empty(x.shape[0], dtype=torch.float32, device=x.device)
    _chain(x, hidden, table, w_sdf=_sdf_row(params), sdf=sdf)
    return sdf
