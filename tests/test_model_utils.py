"""Host-side model utilities (CPU): the device ADC bank's mix order and the
matched-filter volume format/loader."""
import math

import numpy as np
import pytest
import torch

from geraf.models.model_utils import mf_volume as mv
from geraf.models.model_utils.adc_bank import ADCBank, generate_series


def _reference_mix(cache_r, cache_i, groups, cur, sim_r, sim_i, t_pos, r_pos):
    """The reference's list-of-groups bookkeeping (rf_rendering.py:652-684,
    797-800), restated: the current group's cache entry is replaced, the other
    groups are concatenated in ascending order after the fresh signals."""
    cache_r[cur], cache_i[cur] = sim_r.clone(), sim_i.clone()
    others = [i for i in range(len(groups)) if i != cur]
    mix_r = torch.cat([sim_r] + [cache_r[i] for i in others])
    mix_i = torch.cat([sim_i] + [cache_i[i] for i in others])
    mix_t = torch.cat([t_pos[groups[cur]]] + [t_pos[groups[i]] for i in others])
    mix_p = torch.cat([r_pos[groups[cur]]] + [r_pos[groups[i]] for i in others])
    return mix_r, mix_i, mix_t, mix_p


@pytest.mark.parametrize("n_ant,bank", [(10, 3), (12, 4), (7, 7), (5, 2)])
def test_adc_bank_mix_matches_reference_bookkeeping(n_ant, bank):
    torch.manual_seed(n_ant * 10 + bank)
    K = 6
    groups = generate_series(n_ant, bank)
    assert torch.equal(torch.sort(torch.cat(groups)).values, torch.arange(n_ant))
    assert all(torch.equal(g, torch.arange(i, n_ant, bank)) for i, g in enumerate(groups))
    t_pos, r_pos = torch.randn(n_ant, 3), torch.randn(n_ant, 3)
    b = ADCBank(n_ant, K, bank, torch.device("cpu"))
    cache_r = [torch.zeros(len(g), K) for g in groups]
    cache_i = [torch.zeros(len(g), K) for g in groups]
    for step in range(3 * bank):
        cur = b.chunk_id
        sim_r, sim_i = torch.randn(len(groups[cur]), K), torch.randn(len(groups[cur]), K)
        got = b.mix(sim_r, sim_i, t_pos, r_pos)
        ref = _reference_mix(cache_r, cache_i, groups, cur, sim_r, sim_i, t_pos, r_pos)
        for g, r in zip(got, ref):
            assert torch.equal(g, r), step
        b.advance()
    with pytest.raises(ValueError):
        ADCBank(4, 2, 5, torch.device("cpu"))


def test_grid_num_and_shape():
    # the stage-1 scene bounds at 1 mm give the 601^3 grid
    assert mv.grid_num(1.100, 0.500, 0.001) == 601
    assert mv.grid_shape([0.5, 1.1, -0.37, 0.23, -0.15, 0.45], [0.001] * 3) == (601, 601, 601)
    with pytest.raises(ValueError):
        mv.grid_num(1.0, 0.0, 0.0)
    ax = mv.grid_axes([0.0, 1.0, 0.0, 0.5, -1.0, 1.0], [0.25, 0.25, 0.5])
    assert [len(a) for a in ax] == [5, 3, 5] and ax[2][0] == -1.0 and ax[2][-1] == 1.0


def test_sarimage_roundtrip(tmp_path):
    bound, res = [0.0, 0.3, 0.0, 0.2, 0.0, 0.4], [0.1, 0.1, 0.1]
    shape = mv.grid_shape(bound, res)
    vol = np.random.default_rng(0).random(shape).astype(np.float32)
    p = str(tmp_path / "sarimage.bin")
    mv.write_sarimage(p, torch.tensor(vol))
    assert (tmp_path / "sarimage.bin").stat().st_size == 4 * vol.size
    assert np.array_equal(mv.read_sarimage(p, bound, res), vol)
    assert np.allclose(mv.read_sarimage(p, bound, res, trans_power=4.0), vol / 4.0)
    with pytest.raises(ValueError):
        mv.read_sarimage(p, bound, [0.05, 0.1, 0.1])
    acc = mv.accumulate([vol, 2 * vol, vol])
    assert np.allclose(acc, 4 * vol)


def _trilinear(vol, pts):
    """Brute-force trilinear lookup, point (x, y, z) in [-1, 1]^3 mapped to
    volume index ((x+1)/2 (Nx-1), (y+1)/2 (Ny-1), (z+1)/2 (Nz-1))."""
    n = np.array(vol.shape)
    out = []
    for p in pts:
        u = (p + 1.0) / 2.0 * (n - 1)
        i0 = np.clip(np.floor(u).astype(int), 0, n - 2)
        f = u - i0
        acc = 0.0
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = ((f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) *
                         (f[2] if dz else 1 - f[2]))
                    acc += w * vol[i0[0] + dx, i0[1] + dy, i0[2] + dz]
        out.append(acc)
    return np.array(out)


def test_sample_volume_is_trilinear_in_xyz():
    rng = np.random.default_rng(1)
    vol = rng.random((5, 7, 9)).astype(np.float32)
    pts = rng.uniform(-1, 1, (200, 3))
    got = mv.sample_volume(vol, pts).numpy()
    assert np.allclose(got, _trilinear(vol.astype(np.float64), pts), atol=1e-5)
    # corners hit the corner voxels exactly
    c = mv.sample_volume(vol, np.array([[-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]])).numpy()
    assert np.allclose(c, [vol[0, 0, 0], vol[-1, -1, -1]])


@pytest.mark.parametrize("R,positions", [(8192, 4096), (300, 7), (5, 5)])
def test_stage_antenna_groups_cover_every_ray_once(R, positions):
    """Rays from one aperture position form one antenna; the fixed-order row
    sums equal a scatter-add of the rays' signals, and every ray is in
    exactly one group."""
    from rf_workloads.stage_step import antenna_groups, merge_rays

    rng = np.random.default_rng(R)
    ap = torch.tensor(rng.uniform(-0.06, 0.06, (positions, 3)), dtype=torch.float32)
    o = ap[torch.tensor(rng.integers(0, positions, R))]
    pos, idx, of_ray = antenna_groups(o)
    inv = torch.cdist(o.double(), pos.double()).argmin(1)
    assert torch.equal(pos[inv], o) and torch.equal(of_ray, inv)
    assert len({tuple(p) for p in pos.tolist()}) == pos.shape[0]
    members = idx[idx < R]
    assert torch.equal(members.sort().values, torch.arange(R))
    for u in range(pos.shape[0]):
        assert torch.all(inv[idx[u][idx[u] < R]] == u)
    sig = torch.tensor(rng.standard_normal((R, 6)), requires_grad=True)
    ref = torch.zeros(pos.shape[0], 6, dtype=sig.dtype).index_add_(0, inv, sig)
    got = merge_rays(sig, idx, of_ray)
    assert torch.allclose(got, ref, rtol=0, atol=1e-12)
    # the adjoint: every ray receives its antenna's gradient row
    g = torch.tensor(rng.standard_normal((pos.shape[0], 6)))
    (gs,) = torch.autograd.grad(got, sig, g)
    (gr,) = torch.autograd.grad(ref, sig, g)
    assert torch.equal(gs, gr)


def test_compute_mf_loss_modes_closed_form():
    """compute_mf_loss (rf_rendering.py:196-209) on [n,1] predictions: each
    mode against its formula written out per element, and an unknown mode
    raises NotImplementedError."""
    from geraf.models.rendering.rf_stage_ops import compute_mf_loss

    pred = torch.tensor([[0.1], [0.9], [2.0], [0.5]], dtype=torch.float64)
    gt = torch.tensor([[0.3], [0.2], [2.1], [0.5]], dtype=torch.float64)
    d = [p - g for p, g in zip(pred[:, 0].tolist(), gt[:, 0].tolist())]
    want = {
        "l2": [x * x for x in d],
        "l1": [abs(x) for x in d],
        # smooth L1 with beta = 0.25: 0.5 x^2 / beta below beta, |x| - beta / 2 above
        "smooth_l1": [0.5 * x * x / 0.25 if abs(x) < 0.25 else abs(x) - 0.125 for x in d],
        "charbonier": [math.sqrt(x * x + 1e-3) for x in d],
    }
    for mode, w in want.items():
        got = compute_mf_loss(pred, gt, mode)
        assert got.shape == (4,)
        np.testing.assert_allclose(got.numpy(), w, rtol=1e-15, atol=1e-15, err_msg=mode)
    with pytest.raises(NotImplementedError):
        compute_mf_loss(pred, gt, "sum")
