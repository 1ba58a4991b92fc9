"""The per-tile expansion behind the fast matched filter (CPU, FP64): the
expansion equals the direct k-sum to within its truncation bound, and the
Morton tiling of a regular grid yields exact bricks."""
import math

import numpy as np
import torch

from geraf.ops.RFMatchedFilter.tiled_matched_filter import (expansion_terms, morton_order,
                                                            tile_bounds)
from rf_workloads import mf_scenes


def _direct(sc, v):
    k = np.arange(sc.K)
    out = np.zeros(len(v), complex)
    for a in range(sc.A):
        tau = (np.linalg.norm(sc.tx[a] - v, axis=1) + np.linalg.norm(sc.rx[a] - v, axis=1) +
               sc.bias) / sc.c
        out += (sc.echo[a][None, :] *
                np.exp(-2j * np.pi * (sc.f0 + k[None, :] * sc.As_sci / sc.fs) * tau[:, None])).sum(1)
    return out


def _expanded(sc, v, nterms):
    """sum_k s_k e^{-j2pi(f0+kd)tau} = e^{-j2pi(f0+k_c d)tau} sum_n C_n u^n,
    u = (tau - tau_c)/H, C_n = (-j)^n Z^n/n! sum_k s_k e^{-j2pi(k-k_c) d tau_c} xi_k^n."""
    d, K = sc.As_sci / sc.fs, sc.K
    c = (v.min(0) + v.max(0)) / 2
    r = np.linalg.norm(v - c, axis=1).max()
    H = 2 * r / sc.c
    kc = (K - 1) / 2
    Z = math.pi * (K - 1) * d * H
    k = np.arange(K)
    xi = (k - kc) / kc
    out = np.zeros(len(v), complex)
    bound = 0.0
    for a in range(sc.A):
        tac = (np.linalg.norm(sc.tx[a] - c) + np.linalg.norm(sc.rx[a] - c) + sc.bias) / sc.c
        B = sc.echo[a] * np.exp(-2j * np.pi * (k - kc) * d * tac)
        C = [(-1j) ** n * Z ** n / math.factorial(n) * (B * xi ** n).sum() for n in range(nterms)]
        tau = (np.linalg.norm(sc.tx[a] - v, axis=1) + np.linalg.norm(sc.rx[a] - v, axis=1) +
               sc.bias) / sc.c
        u = (tau - tac) / H
        assert np.all(np.abs(u) <= 1.0 + 1e-12)
        out += np.exp(-2j * np.pi * (sc.f0 + kc * d) * tau) * np.polyval(C[::-1], u)
        bound += np.abs(sc.echo[a]).sum() * Z ** nterms / math.factorial(nterms) * math.exp(Z)
    return out, Z, bound


def test_expansion_within_truncation_bound():
    sc = mf_scenes.config1(0)
    sc.echo = sc.echo.astype(np.complex128)
    g = sc.voxels.astype(np.float64).reshape(32, 32, 32, 3)
    for (i, j, l) in ((0, 0, 0), (3, 5, 2), (7, 7, 7)):
        v = g[4 * i:4 * i + 4, 4 * j:4 * j + 4, 4 * l:4 * l + 4].reshape(-1, 3)
        ref = _direct(sc, v)
        for n in (4, 6, 9):
            got, Z, bound = _expanded(sc, v, n)
            err = np.abs(got - ref).max()
            assert err <= bound, (n, err, bound)
        assert np.abs(got - ref).max() <= 1e-8 * np.abs(ref).max()


def test_expansion_terms_rule():
    for Z in (0.0, 0.3, 0.8, 2.0, 5.0):
        n = expansion_terms(Z, 1e-6)
        assert Z ** n / math.factorial(n) * math.exp(Z) <= 1e-6
        if n > 1:
            m = n - 1
            assert Z ** m / math.factorial(m) * math.exp(Z) > 1e-6


def test_morton_tiles_of_a_grid_are_bricks():
    ax = torch.linspace(-0.1, 0.1, 32, dtype=torch.float64)
    X, Y, Z = torch.meshgrid(ax, ax, ax, indexing="ij")
    vox = torch.stack([X.reshape(-1), Y.reshape(-1), Z.reshape(-1)], 1).float()
    perm = morton_order(vox)
    assert torch.equal(torch.sort(perm).values, torch.arange(vox.shape[0]))
    info = tile_bounds(vox[perm])
    assert info.shape == (64, 4)
    pitch = 0.2 / 31
    brick = math.sqrt(3) * 3.5 * pitch  # centre-to-corner of an 8x8x8 brick
    assert torch.allclose(info[:, 3], torch.full((64,), brick), rtol=1e-4)
    # scattered points: every voxel lies within its tile's radius
    pts = torch.rand(3000, 3)
    p = morton_order(pts)
    inf = tile_bounds(pts[p])
    t = torch.arange(3000) // 512
    dist = (pts[p].double() - inf[t, :3].double()).norm(dim=1)
    assert torch.all(dist <= inf[t, 3].double())


def test_plan_tile_sizes_and_coverage():
    """A dense grid at K = 512 is all 512-voxel tiles.  Scattered ray samples
    leave voxels of large 512-voxel tiles: at K = 512 they are re-tiled in
    128-voxel tiles (which take part of them back from the direct kernel), at
    K = 64 they stay direct (the re-tiling would not pay).  Every voxel
    appears exactly once, in a bucket slot or in the direct list, for each
    tile setting."""
    from geraf.ops.RFMatchedFilter import _C as mfC
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import TilePlan
    from rf_workloads import tracer_scenes as ts

    lin = torch.linspace(-0.05, 0.05, 32)
    gx, gy, gz = torch.meshgrid(lin, lin, lin + 0.3, indexing="ij")
    grid = torch.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)], 1)
    plan = TilePlan(grid, 512, mfC._params(10e6, 701500.0), 1e-6)
    assert {bk.tile for bk in plan.buckets} == {512} and plan.tiled_fraction == 1.0

    b = ts.rays(1024, 64, 64, seed=0, uneven=(16, 64))
    pts = b.o[:, None, :] + b.z[..., None] * b.d[:, None, :]
    valid = torch.arange(64)[None, :] < b.n_valid[:, None].long()
    tgt = pts[valid].contiguous()
    r = b.radar
    prm = mfC._params(r["adcSampleRate"], r["As_sci"])
    plans = {t: TilePlan(tgt, 64, prm, 1e-6, tile=t) for t in (None, 512, 128)}
    for p in plans.values():
        seen = torch.cat([bk.idx[bk.valid] for bk in p.buckets] + [p.direct_idx])
        assert torch.equal(torch.sort(seen).values, torch.arange(tgt.shape[0]))
    assert {bk.tile for bk in plans[None].buckets} == {512}
    assert plans[128].tiled_fraction > plans[512].tiled_fraction
    prm512 = mfC._params(r["adcSampleRate"] * 8, r["As_sci"])  # the same band at K = 512
    p512 = TilePlan(tgt, 512, prm512, 1e-6)
    seen = torch.cat([bk.idx[bk.valid] for bk in p512.buckets] + [p512.direct_idx])
    assert torch.equal(torch.sort(seen).values, torch.arange(tgt.shape[0]))
    assert {bk.tile for bk in p512.buckets} == {512, 128}
    assert p512.tiled_fraction > TilePlan(tgt, 512, prm512, 1e-6, tile=512).tiled_fraction
