"""Fast matched filter by per-tile expansions (no counterpart in the
reference; the operator is the reference's matched filter,
MatchedFilter_cuda_kernel.cu:65-115, evaluated to a stated truncation bound).

``matched_filter_tiled`` takes the arguments of ``matched_filter_cuda`` plus a
tolerance.  It orders the voxels along a Morton curve so that every 512
consecutive voxels form a compact tile (exact 8x8x8 bricks when the voxels
form a regular grid), measures each tile's radius, and picks the number of
expansion terms N_T from the largest tile:

    Z = pi (K-1) (slope/fs) 2 r_max / c,   Z^N_T / N_T! e^Z <= tol,

which bounds the truncation error of every antenna's contribution by
tol * sum_k |s[a,k]|.  When N_T would exceed 32 (tiles too large against the
range resolution) or K < 2 it runs the direct kernel instead.  The backward is
the adjoint by the same expansion (same bound) when the forward ran tiled and
K <= 2048, else the direct adjoint.
"""
from __future__ import annotations

import math
from typing import Tuple

import torch
from torch.autograd import Function
from torch.autograd.function import once_differentiable

from . import _C
from .matched_filter_cuda import matched_filter_cuda
from .. import _native

TILE = 512
MAX_TERMS = 32
WORKSPACE_CAP = 2 << 30


def _spread3(x: torch.Tensor) -> torch.Tensor:
    """Spread the low 21 bits of x so that two zero bits follow each bit."""
    x = x & 0x1FFFFF
    x = (x | (x << 32)) & 0x1F00000000FFFF
    x = (x | (x << 16)) & 0x1F0000FF0000FF
    x = (x | (x << 8)) & 0x100F00F00F00F00F
    x = (x | (x << 4)) & 0x10C30C30C30C30C3
    x = (x | (x << 2)) & 0x1249249249249249
    return x


def morton_order(voxel: torch.Tensor) -> torch.Tensor:
    """Permutation of the voxels [N, 3] along a Morton (Z-order) curve.  The
    integer cell of a voxel is its coordinate rank when the voxels form a
    regular grid (every (x, y, z) combination of the unique coordinates
    present), otherwise its cell in a grid with one cell per voxel on
    average."""
    n = voxel.shape[0]
    ranks = []
    uniq = [torch.unique(voxel[:, i], return_inverse=True) for i in range(3)]
    if math.prod(len(u[0]) for u in uniq) == n:
        ranks = [u[1] for u in uniq]
    else:
        lo = voxel.min(0).values
        ext = (voxel.max(0).values - lo).clamp(min=1e-12)
        cell = float((ext.prod() / max(n, 1)) ** (1.0 / 3.0))
        cell = max(cell, float(ext.max()) / (1 << 20))
        ranks = [((voxel[:, i] - lo[i]) / cell).floor().long().clamp(0, (1 << 21) - 1)
                 for i in range(3)]
    code = _spread3(ranks[0]) | (_spread3(ranks[1]) << 1) | (_spread3(ranks[2]) << 2)
    return torch.argsort(code)


def tile_bounds(vox_sorted: torch.Tensor, tile: int = TILE) -> torch.Tensor:
    """[ceil(N/tile), 4] float32: each tile's bounding-box centre and the
    largest voxel distance from it (slightly inflated so that FP32 rounding
    cannot put a voxel outside)."""
    n = vox_sorted.shape[0]
    nt = (n + tile - 1) // tile
    pad = nt * tile - n
    v = vox_sorted.double()
    if pad:
        v = torch.cat([v, v[-1:].expand(pad, 3)])
    v = v.reshape(nt, tile, 3)
    c = (v.amin(1) + v.amax(1)) * 0.5
    r = (v - c[:, None, :]).norm(dim=-1).amax(1) * (1.0 + 1e-6) + 1e-9
    return torch.cat([c, r[:, None]], dim=1).float().contiguous()


def expansion_terms(Z: float, tol: float) -> int:
    """Smallest N with Z^N / N! * e^Z <= tol (at least 1)."""
    n, term = 1, Z
    while term * math.exp(Z) > tol:
        n += 1
        term *= Z / n
        if n > 64:
            break
    return n


class TilePlan:
    """Morton order, tile bounds and term count of one voxel set."""

    def __init__(self, voxel_nx3: torch.Tensor, K: int, prm, tol: float):
        self.N = voxel_nx3.shape[0]
        self.perm = morton_order(voxel_nx3)
        vs = voxel_nx3[self.perm].contiguous()
        self.info = tile_bounds(vs)
        self.vt = vs.t().contiguous()
        r_max = float(self.info[:, 3].max())
        self.Z = math.pi * (K - 1) * (prm.slope / prm.fs) * 2.0 * r_max / prm.c
        self.terms = expansion_terms(self.Z, tol)
        self.usable = K >= 2 and self.terms <= MAX_TERMS

    def workspace(self, A: int, K: int, device) -> Tuple[torch.Tensor, int]:
        lib = _native.load()
        need = int(lib.geraf_mf_tiled_workspace_bytes(A, K, self.N, self.terms))
        per_one = int(lib.geraf_mf_tiled_workspace_bytes(1, K, self.N, self.terms))
        size = max(per_one, min(need, WORKSPACE_CAP))
        return torch.empty(size, dtype=torch.uint8, device=device), size


def tiled_forward(real, image, tpos, rpos, voxel_nx3, numADCSample: int, adcSampleRate: float,
                  As_sci: float, tol: float = 1e-6, plan: "TilePlan" = None):
    """(out_real, out_imag, info) for float32 inputs on one device; info says
    which path ran and with how many terms (and carries the plan)."""
    A, K = real.shape
    N = voxel_nx3.shape[0]
    out_r = torch.empty(N, dtype=torch.float32, device=real.device)
    out_i = torch.empty_like(out_r)
    if N == 0:
        return out_r, out_i, {"path": "tiled", "terms": 0, "Z": 0.0, "plan": None}
    prm = _C._params(adcSampleRate, As_sci)
    plan = plan or TilePlan(voxel_nx3, K, prm, tol)
    if not plan.usable:
        o_r, o_i = matched_filter_cuda(real, image, tpos, rpos, voxel_nx3, A, K, adcSampleRate,
                                       As_sci, N)
        return o_r, o_i, {"path": "direct", "terms": plan.terms, "Z": plan.Z, "plan": plan}
    ws, size = plan.workspace(A, K, real.device)
    s_r = torch.empty(N, dtype=torch.float32, device=real.device)
    s_i = torch.empty_like(s_r)
    P = _native.ptr
    _native.call("geraf_mf_forward_tiled_f32", P(real), P(image), P(tpos), P(rpos), P(s_r),
                 P(s_i), P(plan.vt), P(plan.info), A, K, N, plan.terms, P(ws), size, prm,
                 _native.stream_of(real))
    out_r[plan.perm] = s_r
    out_i[plan.perm] = s_i
    return out_r, out_i, {"path": "tiled", "terms": plan.terms, "Z": plan.Z, "plan": plan}


def tiled_adjoint(g_r, g_i, tpos, rpos, plan: "TilePlan", A: int, K: int, norm_count: int,
                  adcSampleRate: float, As_sci: float):
    """grad [A, K] (real, imag) of the matched filter by the tile expansion;
    g_r/g_i [N] in the caller's voxel order."""
    prm = _C._params(adcSampleRate, As_sci)
    gr = g_r.float()[plan.perm].contiguous()
    gi = g_i.float()[plan.perm].contiguous()
    out_r = torch.empty((A, K), dtype=torch.float32, device=gr.device)
    out_i = torch.empty_like(out_r)
    ws, size = plan.workspace(A, K, gr.device)
    P = _native.ptr
    _native.call("geraf_mf_backward_tiled_f32", P(gr), P(gi), P(tpos), P(rpos), P(out_r),
                 P(out_i), P(plan.vt), P(plan.info), A, K, plan.N, norm_count, plan.terms,
                 P(ws), size, prm, _native.stream_of(gr))
    return out_r, out_i


class _TiledMatchedFilter(Function):
    @staticmethod
    def forward(ctx, real, image, tpos, rpos, voxel, total_ant_size: int, numADCSample: int,
                adcSampleRate: float, As_sci: float, N: int, tol: float):
        assert real.dim() == 2 and image.dim() == 2 and voxel.dim() == 2
        assert real.shape == (total_ant_size, numADCSample), (
            f"Size mismatch for real tensor: expected ({total_ant_size}, {numADCSample}), "
            f"but got: {tuple(real.shape)}")
        assert voxel.shape == (N, 3), f"expected voxel ({N}, 3), got {tuple(voxel.shape)}"
        args = [t.float().contiguous() for t in (real, image, tpos, rpos, voxel)]
        o_r, o_i, info = tiled_forward(*args, numADCSample, adcSampleRate, As_sci, tol)
        ctx.save_for_backward(*args)
        ctx.cfg = (total_ant_size, numADCSample, adcSampleRate, As_sci, N)
        ctx.plan = info["plan"] if info["path"] == "tiled" and numADCSample <= 2048 else None
        ctx.dtype = real.dtype
        return o_r.to(real.dtype), o_i.to(real.dtype)

    @staticmethod
    @once_differentiable
    def backward(ctx, g_r, g_i):
        real, image, tpos, rpos, voxel = ctx.saved_tensors
        A, K, fs, As_sci, N = ctx.cfg
        grad_real = grad_image = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            if ctx.plan is not None:
                grad_real, grad_image = tiled_adjoint(g_r, g_i, tpos, rpos, ctx.plan, A, K, A,
                                                      fs, As_sci)
            else:
                grad_real = torch.empty_like(real)
                grad_image = torch.empty_like(image)
                _C.MatchedFilter_backward_realimag(
                    g_r.float().contiguous(), g_i.float().contiguous(), real, image, tpos, rpos,
                    grad_real, grad_image, voxel.t().contiguous(), A, K, fs, As_sci, N)
            grad_real, grad_image = grad_real.to(ctx.dtype), grad_image.to(ctx.dtype)
        return (grad_real, grad_image) + (None,) * 9


def matched_filter_tiled(real: torch.Tensor, image: torch.Tensor, tpos: torch.Tensor,
                         rpos: torch.Tensor, voxel: torch.Tensor, total_ant_size: int,
                         numADCSample: int, adcSampleRate: float, As_sci: float, N: int,
                         tol: float = 1e-6) -> Tuple[torch.Tensor, torch.Tensor]:
    """``matched_filter_cuda`` evaluated by per-tile expansions (see the module
    docstring for the error bound)."""
    assert real.is_cuda and image.is_cuda and tpos.is_cuda and rpos.is_cuda and voxel.is_cuda, \
        "All inputs must be CUDA tensors."
    return _TiledMatchedFilter.apply(real, image, tpos, rpos, voxel, total_ant_size,
                                     numADCSample, adcSampleRate, As_sci, N, tol)
