"""Fast matched filter by per-tile expansions (no counterpart in the
reference; the operator is the reference's matched filter,
MatchedFilter_cuda_kernel.cu:65-115, evaluated to a stated truncation bound).

``matched_filter_tiled`` takes the arguments of ``matched_filter_cuda`` plus a
tolerance.  It orders the voxels along a Morton curve so that every 512 (or
128) consecutive voxels form a compact tile (exact 8x8x8 bricks when the
voxels form a regular grid), measures each tile's radius r_t, and gives each
tile the smallest instantiated number of expansion terms N_T with

    Z_t = pi (K-1) (slope/fs) 2 r_t / c,   Z_t^N_T / N_T! e^Z_t <= tol,

which bounds the truncation error of every antenna's contribution to the
tile's voxels by tol * sum_k |s[a,k]|.  Tiles of one term count run in one
launch sequence; tiles that would need more than 32 terms (too large against
the range resolution) go through the direct kernel, as does everything when
K < 2.  The backward is the adjoint by the same expansion (same bound, plus
the direct adjoint for the direct voxels) when the forward ran tiled and
K <= 2048, else the direct adjoint.

Voxels of 512-voxel tiles that would need more than 32 terms can be tiled
again in 128-voxel tiles (smaller radius, fewer terms).  A 128-voxel tile
carries the same per-tile work as a 512-voxel one (per antenna: the forward
moments and adjoint evaluation, K (18 + 4 N_T), and the adjoint moments'
cross-lane sums and centre delay, ~1024 N_T + 1920 lane operations), against
(110 + 5 N_T) per tiled voxel and ~10 K + 140 per direct voxel for forward
plus adjoint.  The leftovers are re-tiled only when that count, with a 1.5x
margin for the extra launches, beats the direct kernels: at K = 512 by ~6x,
at K = 64 not at all.
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
TILE_SIZES = (512, 128)
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


# Term counts the kernels are instantiated for.
TERM_SETS = (4, 6, 8, 10, 12, 14, 16, 20, 24, 32)


def _z_limit(nterms: int, tol: float) -> float:
    """Largest Z with Z^nterms / nterms! * e^Z <= tol (bisection)."""
    lo, hi = 0.0, 64.0
    for _ in range(80):
        mid = 0.5 * (lo + hi)
        if mid ** nterms / math.factorial(nterms) * math.exp(mid) <= tol:
            lo = mid
        else:
            hi = mid
    return lo


class _Bucket:
    """Tiles sharing one term count: slot i of the bucket holds voxel idx[i]
    (padding slots repeat a voxel and are masked out)."""

    def __init__(self, terms, idx, valid, voxel_nx3, info, tile=TILE):
        self.terms = terms
        self.tile = tile
        self.idx = idx
        self.valid = valid
        self.vt = voxel_nx3[idx].t().contiguous()
        self.info = info.contiguous()
        self.N = idx.numel()


class TilePlan:
    """Morton order, tiles, and the tiles grouped by the term count each needs
    (the smallest instantiated N_T with Z_t^N_T / N_T! e^Z_t <= tol).
    ``tile`` None: 512-voxel tiles, their leftovers (tiles that would need
    more than 32 terms) in 128-voxel tiles, the rest to the direct kernel;
    512 or 128: that size only."""

    def __init__(self, voxel_nx3: torch.Tensor, K: int, prm, tol: float, tile=None):
        if tile is not None and tile not in TILE_SIZES:
            raise ValueError(f"tile must be one of {TILE_SIZES}")
        dev = voxel_nx3.device
        self.N = n = voxel_nx3.shape[0]
        zfac = math.pi * (K - 1) * (prm.slope / prm.fs) * 2.0 / prm.c
        self.buckets = []
        self.Z = 0.0
        left = morton_order(voxel_nx3)  # voxels not yet in a tiled bucket, in Morton order
        for tv in ((tile,) if tile is not None else TILE_SIZES):
            if left.numel() == 0 or K < 2:
                break
            if tile is None and tv != TILE and not self._retile_pays(voxel_nx3, left, tv, zfac,
                                                                    K, tol):
                break
            left = self._tile(voxel_nx3, left, tv, zfac, tol)
        self.direct_idx = left
        self.tile = tile if tile is not None else TILE
        self.terms = expansion_terms(self.Z, tol)
        self.usable = len(self.buckets) > 0
        self.tiled_fraction = 1.0 - self.direct_idx.numel() / max(n, 1)

    @staticmethod
    def _retile_pays(voxel_nx3, order, tv, zfac, K, tol) -> bool:
        """Work count per antenna of tiling `order` in tv-voxel tiles against
        leaving it to the direct kernels (forward + adjoint)."""
        m = order.numel()
        nt = (m + tv - 1) // tv
        pad = nt * tv - m
        slots = torch.cat([order, order[-1:].expand(pad)]) if pad else order
        z = tile_bounds(voxel_nx3[slots], tv)[:, 3].double() * zfac
        tiled = 0.0
        direct_vox = float(m)
        lower = -1.0
        for terms in TERM_SETS:
            lim = _z_limit(terms, tol)
            cnt = int(((z > lower) & (z <= lim)).sum())
            lower = lim
            tiled += cnt * (K * (18 + 4 * terms) + 1024 * terms + 1920 + tv * (110 + 5 * terms))
            direct_vox -= cnt * tv
        direct_cost = (10 * K + 140)
        return 1.5 * (tiled + max(direct_vox, 0.0) * direct_cost) < m * direct_cost

    def _tile(self, voxel_nx3, order, tv, zfac, tol):
        """Tiles of tv consecutive voxels of `order` into buckets; returns the
        voxels of tiles that would need more than MAX_TERMS terms."""
        dev = voxel_nx3.device
        m = order.numel()
        nt = (m + tv - 1) // tv
        pad = nt * tv - m
        slots = torch.cat([order, order[-1:].expand(pad)]) if pad else order
        valid = torch.ones(nt * tv, dtype=torch.bool, device=dev)
        if pad:
            valid[m:] = False
        info = tile_bounds(voxel_nx3[slots], tv)
        z = info[:, 3].double() * zfac
        direct = torch.ones(nt, dtype=torch.bool, device=dev)
        lower = -1.0
        for terms in TERM_SETS:
            lim = _z_limit(terms, tol)
            sel = (z > lower) & (z <= lim)
            lower = lim
            tiles = torch.nonzero(sel).reshape(-1)
            if tiles.numel() == 0:
                continue
            direct[tiles] = False
            self.Z = max(self.Z, float(z[tiles].max()))
            pos = (tiles[:, None] * tv + torch.arange(tv, device=dev)[None, :]).reshape(-1)
            self.buckets.append(_Bucket(terms, slots[pos], valid[pos], voxel_nx3, info[tiles],
                                        tv))
        dpos = (torch.nonzero(direct).reshape(-1)[:, None] * tv +
                torch.arange(tv, device=dev)[None, :]).reshape(-1)
        return slots[dpos][valid[dpos]]

    @staticmethod
    def workspace(A: int, K: int, b: "_Bucket", device) -> Tuple[torch.Tensor, int]:
        lib = _native.load()
        need = int(lib.geraf_mf_tiled_workspace_bytes(A, K, b.N, b.terms, b.tile))
        per_one = int(lib.geraf_mf_tiled_workspace_bytes(1, K, b.N, b.terms, b.tile))
        size = max(per_one, min(need, WORKSPACE_CAP))
        return torch.empty(size, dtype=torch.uint8, device=device), size


def _info(plan: "TilePlan", path: str):
    return {"path": path, "terms": plan.terms, "Z": plan.Z, "plan": plan,
            "tiled_fraction": plan.tiled_fraction,
            "bucket_terms": [b.terms for b in plan.buckets]}


def tiled_forward(real, image, tpos, rpos, voxel_nx3, numADCSample: int, adcSampleRate: float,
                  As_sci: float, tol: float = 1e-6, plan: "TilePlan" = None):
    """(out_real, out_imag, info) for float32 inputs on one device; info says
    which path ran, how many terms the largest tile needs, and which share of
    the voxels went through the expansion (the rest through the direct
    kernel)."""
    A, K = real.shape
    N = voxel_nx3.shape[0]
    out_r = torch.empty(N, dtype=torch.float32, device=real.device)
    out_i = torch.empty_like(out_r)
    if N == 0:
        return out_r, out_i, {"path": "tiled", "terms": 0, "Z": 0.0, "plan": None,
                              "tiled_fraction": 1.0, "bucket_terms": []}
    prm = _C._params(adcSampleRate, As_sci)
    plan = plan or TilePlan(voxel_nx3, K, prm, tol)
    if not plan.usable:
        o_r, o_i = matched_filter_cuda(real, image, tpos, rpos, voxel_nx3, A, K, adcSampleRate,
                                       As_sci, N)
        return o_r, o_i, _info(plan, "direct")
    P = _native.ptr
    st = _native.stream_of(real)
    for b in plan.buckets:
        ws, size = TilePlan.workspace(A, K, b, real.device)
        s_r = torch.empty(b.N, dtype=torch.float32, device=real.device)
        s_i = torch.empty_like(s_r)
        _native.call("geraf_mf_forward_tiled_f32", P(real), P(image), P(tpos), P(rpos), P(s_r),
                     P(s_i), P(b.vt), P(b.info), A, K, b.N, b.terms, b.tile, P(ws), size, prm,
                     st)
        out_r[b.idx[b.valid]] = s_r[b.valid]
        out_i[b.idx[b.valid]] = s_i[b.valid]
    if plan.direct_idx.numel():
        v = voxel_nx3[plan.direct_idx]
        d_r, d_i = matched_filter_cuda(real, image, tpos, rpos, v, A, K, adcSampleRate, As_sci,
                                       v.shape[0])
        out_r[plan.direct_idx] = d_r
        out_i[plan.direct_idx] = d_i
    return out_r, out_i, _info(plan, "tiled")


def tiled_adjoint(g_r, g_i, tpos, rpos, plan: "TilePlan", A: int, K: int, norm_count: int,
                  adcSampleRate: float, As_sci: float, voxel_nx3: torch.Tensor = None):
    """grad [A, K] (real, imag) of the matched filter through the plan's
    buckets (and the direct adjoint for its direct voxels, which needs
    voxel_nx3); g_r/g_i [N] in the caller's voxel order."""
    prm = _C._params(adcSampleRate, As_sci)
    gr, gi = g_r.float(), g_i.float()
    out_r = torch.zeros((A, K), dtype=torch.float32, device=gr.device)
    out_i = torch.zeros_like(out_r)
    P = _native.ptr
    st = _native.stream_of(gr)
    for b in plan.buckets:
        br = (gr[b.idx] * b.valid).contiguous()
        bi = (gi[b.idx] * b.valid).contiguous()
        ws, size = TilePlan.workspace(A, K, b, gr.device)
        p_r = torch.empty((A, K), dtype=torch.float32, device=gr.device)
        p_i = torch.empty_like(p_r)
        _native.call("geraf_mf_backward_tiled_f32", P(br), P(bi), P(tpos), P(rpos), P(p_r),
                     P(p_i), P(b.vt), P(b.info), A, K, b.N, norm_count, b.terms, b.tile, P(ws),
                     size, prm, st)
        out_r += p_r
        out_i += p_i
    if plan.direct_idx.numel():
        if voxel_nx3 is None:
            raise ValueError("the plan has direct voxels: pass voxel_nx3")
        v = voxel_nx3[plan.direct_idx].t().contiguous()
        dr = gr[plan.direct_idx].contiguous()
        di = gi[plan.direct_idx].contiguous()
        p_r = torch.empty((A, K), dtype=torch.float32, device=gr.device)
        p_i = torch.empty_like(p_r)
        _native.call("geraf_mf_backward_f32", P(dr), P(di), P(tpos), P(rpos), P(p_r), P(p_i),
                     P(v), A, K, v.shape[1], norm_count,
                     _native.make_params(prm.f0, prm.slope, prm.fs, prm.c, prm.path_bias), st)
        out_r += p_r
        out_i += p_i
    return out_r, out_i


class _TiledMatchedFilter(Function):
    @staticmethod
    def forward(ctx, real, image, tpos, rpos, voxel, total_ant_size: int, numADCSample: int,
                adcSampleRate: float, As_sci: float, N: int, tol: float, plan):
        assert real.dim() == 2 and image.dim() == 2 and voxel.dim() == 2
        assert real.shape == (total_ant_size, numADCSample), (
            f"Size mismatch for real tensor: expected ({total_ant_size}, {numADCSample}), "
            f"but got: {tuple(real.shape)}")
        assert voxel.shape == (N, 3), f"expected voxel ({N}, 3), got {tuple(voxel.shape)}"
        args = [t.float().contiguous() for t in (real, image, tpos, rpos, voxel)]
        o_r, o_i, info = tiled_forward(*args, numADCSample, adcSampleRate, As_sci, tol, plan)
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
                                                      fs, As_sci, voxel)
            else:
                grad_real = torch.empty_like(real)
                grad_image = torch.empty_like(image)
                _C.MatchedFilter_backward_realimag(
                    g_r.float().contiguous(), g_i.float().contiguous(), real, image, tpos, rpos,
                    grad_real, grad_image, voxel.t().contiguous(), A, K, fs, As_sci, N)
            grad_real, grad_image = grad_real.to(ctx.dtype), grad_image.to(ctx.dtype)
        return (grad_real, grad_image) + (None,) * 10


def matched_filter_tiled(real: torch.Tensor, image: torch.Tensor, tpos: torch.Tensor,
                         rpos: torch.Tensor, voxel: torch.Tensor, total_ant_size: int,
                         numADCSample: int, adcSampleRate: float, As_sci: float, N: int,
                         tol: float = 1e-6, plan: "TilePlan" = None
                         ) -> Tuple[torch.Tensor, torch.Tensor]:
    """``matched_filter_cuda`` evaluated by per-tile expansions (see the module
    docstring for the error bound).  ``plan``: a ``TilePlan`` of the same voxel
    set (and tolerance), for callers that evaluate one voxel set many times."""
    assert real.is_cuda and image.is_cuda and tpos.is_cuda and rpos.is_cuda and voxel.is_cuda, \
        "All inputs must be CUDA tensors."
    return _TiledMatchedFilter.apply(real, image, tpos, rpos, voxel, total_ant_size,
                                     numADCSample, adcSampleRate, As_sci, N, tol, plan)


def make_plan(voxel: torch.Tensor, numADCSample: int, adcSampleRate: float, As_sci: float,
              tol: float = 1e-6) -> "TilePlan":
    """The tile plan of a voxel set [N, 3] for ``matched_filter_tiled``."""
    return TilePlan(voxel.float().contiguous(), numADCSample, _C._params(adcSampleRate, As_sci),
                    tol)
