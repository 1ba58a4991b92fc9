"""The dispatch helpers and the rows built on the ops, each against the CPU
oracle directly: the selected/unselected split matched filter and the
device ADC bank (values and selected-row gradients), the volume producer,
and the model_utils helpers called with radar_cfg dicts as the renderers
call them."""
import numpy as np
import pytest
import torch

from oracle import lensless as L
from oracle import rf as orc
from rf_workloads import mf_scenes
from rf_workloads import tracer_scenes as ts
from test_mf_gpu import rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _geometry(A, N, seed):
    rng = np.random.default_rng(seed)
    tx = rng.uniform(-0.05, 0.05, (A, 3)).astype(np.float32)
    tx[:, 2] = 0
    rx = (tx + np.float32([0.001, 0.0005, 0])).astype(np.float32)
    vox = (rng.uniform(-0.1, 0.1, (N, 3)) + [0, 0, 0.3]).astype(np.float32)
    return rng, tx, rx, vox


def _prm(K):
    fs = 1.25e6 * K / 64
    return fs, dict(f0=mf_scenes.MF_F0, slope=701500.0, fs=fs, c=mf_scenes.SPEED_OF_LIGHT,
                    bias=mf_scenes.PATH_BIAS)


@pytest.mark.parametrize("K", [64, 300])
def test_faster_matched_filter_vs_oracle(cuda, K):
    """Forward = the matched filter of the concatenated banks (÷ total); the
    selected rows' gradient = the adjoint of those rows ÷ total
    (faster_matched_filter_cuda.py:177-238), against the oracle."""
    from geraf.ops.RFMatchedFilter import faster_matched_filter_cuda

    A, sel, N = 11, 4, 900
    rng, tx, rx, vox = _geometry(A, N, K)
    fs, prm = _prm(K)
    s = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    g = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    re = T(s.real[:sel]).requires_grad_(True)
    im = T(s.imag[:sel]).requires_grad_(True)
    o_r, o_i = faster_matched_filter_cuda(re, im, T(s.real[sel:]), T(s.imag[sel:]), T(tx[:sel]),
                                          T(rx[:sel]), T(tx[sel:]), T(rx[sel:]), T(vox), A, sel,
                                          K, fs, prm["slope"], N)
    (o_r * T(g.real) + o_i * T(g.imag)).sum().backward()
    out = (o_r.detach().double() + 1j * o_i.detach().double()).cpu().numpy()
    ref = orc.mf_forward(s.real, s.imag, tx, rx, vox, **prm)
    assert rel_l2(out, ref) <= TOL
    grad = (re.grad.double() + 1j * im.grad.double()).cpu().numpy()
    gref = orc.mf_backward(g, tx[:sel], rx[:sel], vox, K, norm=A, **prm)
    err = rel_l2(grad, gref)
    print(f"faster MF K={K}: forward {rel_l2(out, ref):.2e}, selected-row gradient {err:.2e}")
    assert err <= TOL


def test_adc_bank_selected_rows_vs_oracle(cuda):
    """ADCBank.matched_filter (device cache, selected-rows adjoint) on the
    16x16 K=128 aperture: the value is the oracle's matched filter of the
    mixed bank, the selected rows' gradients the oracle adjoint ÷ A."""
    from geraf.models.model_utils.adc_bank import ADCBank

    sc = mf_scenes.config1(0)
    A, K = sc.A, sc.K
    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    pts = sc.voxels[::29]
    radar = dict(numADCSample=K, adcSampleRate=sc.fs, As_sci=sc.As_sci)
    bank = ADCBank(A, K, 3, cuda)
    rng = np.random.default_rng(5)
    cache = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    for grp in range(3):
        idx = bank.groups[grp].cpu().numpy()
        bank.store(grp, T(cache.real[idx]), T(cache.imag[idx]))
    bank.advance()
    selected = bank.selected.cpu().numpy()
    fresh = (rng.standard_normal((selected.size, K)) +
             1j * rng.standard_normal((selected.size, K))).astype(np.complex64)
    re = T(fresh.real).requires_grad_(True)
    im = T(fresh.imag).requires_grad_(True)
    yr, yi = bank.matched_filter(re, im, T(sc.tx), T(sc.rx), T(pts), radar,
                                 selected_rows_only=True)
    g = (rng.standard_normal(len(pts)) + 1j * rng.standard_normal(len(pts))).astype(np.complex64)
    (yr * T(g.real) + yi * T(g.imag)).sum().backward()
    mixed = cache.copy()
    mixed[selected] = fresh
    ref = orc.mf_forward(mixed.real, mixed.imag, sc.tx, sc.rx, pts, **sc.params())
    out = (yr.detach().double() + 1j * yi.detach().double()).cpu().numpy()
    gref = orc.mf_backward(g, sc.tx[selected], sc.rx[selected], pts, K, norm=A, **sc.params())
    grad = (re.grad.double() + 1j * im.grad.double()).cpu().numpy()
    print(f"ADC bank: forward {rel_l2(out, ref):.2e}, selected rows {rel_l2(grad, gref):.2e}")
    assert rel_l2(out, ref) <= TOL and rel_l2(grad, gref) <= TOL


@pytest.mark.parametrize("algo", ["tiled", "direct"])
def test_volume_producer_vs_oracle(cuda, tmp_path, algo):
    """backproject / produce_frame: |MF| on the grid of `bound` / `mf_res`
    (grid_num points per axis, both ends) against |oracle| on the same grid,
    and the sarimage.bin the producer writes reads back as that cube."""
    from geraf.models.model_utils import mf_volume as mv

    sc = mf_scenes.config1(0)
    radar = dict(numADCSample=sc.K, adcSampleRate=sc.fs, As_sci=sc.As_sci)
    bound, res = [-0.02, 0.02, -0.015, 0.015, 0.23, 0.27], [0.002, 0.002, 0.002]
    d = tmp_path / "0_0_0"
    d.mkdir()
    for name, arr in (("adcr.npy", sc.echo.real), ("adci.npy", sc.echo.imag),
                      ("tpos.npy", sc.tx), ("rpos.npy", sc.rx)):
        np.save(d / name, np.asarray(arr, dtype=np.float32))
    cube = mv.produce_frame(str(d), bound, res, radar, device=str(cuda), algo=algo)
    shape = mv.grid_shape(bound, res)
    assert cube.shape == shape == (21, 16, 21)
    ax = [a.astype(np.float32) for a in mv.grid_axes(bound, res)]
    gx, gy, gz = np.meshgrid(*ax, indexing="ij")
    grid = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1)
    ref = np.abs(orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, grid, **sc.params()))
    err = rel_l2(cube.ravel(), ref)
    print(f"volume producer ({algo}): rel-L2 vs |oracle| {err:.2e}")
    assert err <= TOL
    assert np.array_equal(mv.read_sarimage(str(d / "sarimage.bin"), bound, res), cube)


def test_dispatch_helpers_with_radar_cfg(cuda):
    """base_matched_filter, lensless_ray_tracer ([A,P] op, with its backward),
    bilinear_ray_sampler and lensless_volume_tracer called with radar_cfg
    dicts, against the oracle."""
    from geraf.models.model_utils.matched_filter import base_matched_filter
    from geraf.models.model_utils.ray_sampler import bilinear_ray_sampler
    from geraf.models.model_utils.ray_tracer import lensless_ray_tracer, lensless_volume_tracer

    T = lambda x, dt=None: torch.tensor(np.ascontiguousarray(x), device=cuda, dtype=dt)  # noqa
    # matched filter
    A, K, N = 7, 96, 500
    rng, tx, rx, vox = _geometry(A, N, 1)
    fs, prm = _prm(K)
    s = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    radar = dict(numADCSample=K, adcSampleRate=fs, As_sci=prm["slope"], fc_start=ts.FC_START)
    o_r, o_i = base_matched_filter(T(s.real), T(s.imag), T(tx), T(rx), T(vox), radar)
    ref = orc.mf_forward(s.real, s.imag, tx, rx, vox, **prm)
    assert rel_l2((o_r.double() + 1j * o_i.double()).cpu().numpy(), ref) <= TOL
    # [A,P] tracer, forward and backward
    P = 800
    pts = (rng.uniform(-0.1, 0.1, (P, 3)) + [0, 0, 0.3]).astype(np.float32)
    n = (-pts / np.linalg.norm(pts, axis=1, keepdims=True)
         + 0.3 * rng.standard_normal((P, 3))).astype(np.float32)
    sig = rng.uniform(0.1, 1, (A, P)).astype(np.float32)
    inb = rng.uniform(0, 1, P) > 0.2
    nt, st = T(n).requires_grad_(True), T(sig).requires_grad_(True)
    t_r, t_i = lensless_ray_tracer(nt, st, T(pts), T(tx), T(rx), radar, inbounds=T(inb))
    sprm = dict(prm, f0=ts.FC_START)
    tref = orc.st_forward(n, sig, pts, tx, rx, inb, K, **sprm)
    assert rel_l2((t_r.detach().double() + 1j * t_i.detach().double()).cpu().numpy(),
                  tref) <= TOL
    g = rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))
    (t_r * T(g.real, torch.float32) + t_i * T(g.imag, torch.float32)).sum().backward()
    gn, gs = orc.st_backward(g.astype(np.complex64), n, sig, pts, tx, rx, inb, **sprm)
    assert rel_l2(nt.grad.cpu().numpy(), gn) <= 10 * TOL
    assert rel_l2(st.grad.cpu().numpy(), gs) <= 10 * TOL
    # depth resampler
    R, S, Tn = 29, 32, 64
    z = np.sort(rng.uniform(0.2, 0.5, (R, S)), axis=1).astype(np.float32)
    tz = np.sort(rng.uniform(0.15, 0.55, (R, Tn)), axis=1).astype(np.float32)
    nr = rng.standard_normal((R, S, 3)).astype(np.float32)
    sg = rng.uniform(0, 1, (1, R, S)).astype(np.float32)
    on, os_, ib = bilinear_ray_sampler(T(nr), T(sg), T(z), T(tz))
    rn, rs, rinb = orc.bilinear_forward(nr, sg, z, tz)
    assert np.array_equal(ib.cpu().numpy(), rinb)
    assert np.max(np.abs(on.double().cpu().numpy() - rn)) <= 1e-6
    assert np.max(np.abs(os_.double().cpu().numpy() - rs)) <= 1e-6
    # fused per-ray renderer through its helper
    b = ts.rays(40, 64, 64, seed=9, with_offset=True, uneven=(8, 64))
    bb = b.to(cuda)
    v_r, v_i = lensless_volume_tracer(bb.sdf, bb.grad, bb.dists, bb.z, bb.o, bb.d, bb.inv_s,
                                      bb.power, b.radar, reflectance=bb.refl,
                                      trans_offset=bb.offset, n_valid=bb.n_valid)
    r = b.radar
    vref = L.lensless_forward(b.sdf.double(), b.grad.double(), b.dists.double(),
                              L.points_f32(b.o, b.z, b.d), b.o.double(), b.d.double(),
                              b.inv_s.double(), b.power.double(), r["numADCSample"],
                              r["adcSampleRate"], r["As_sci"], r["fc_start"], ts.SPEED_OF_LIGHT,
                              ts.PATH_BIAS, refl=b.refl.double(), offset=b.offset.double(),
                              n_valid=b.n_valid.long())
    got = (v_r.double() + 1j * v_i.double()).cpu()
    assert float((got - vref).norm() / vref.norm()) <= TOL
