"""Matched-filter parity on the GPU: the gfx950 kernels (through the C ABI and
the drop-in ops) against the CPU oracle and against each other."""
import numpy as np
import pytest
import torch

from oracle import rf as orc
from rf_workloads import mf_scenes

pytestmark = pytest.mark.gpu

# Tolerances (volume rel-L2 = ||x - x64|| / ||x64||).
TOL_TUNED_VS_F64 = 1.0e-5     # tuned FP32+FP64-reduction kernel vs FP64 oracle
TOL_PLAIN32_VS_REF32 = 2.0e-4  # plain FP32 GPU form vs FP32 reference formula on CPU
TOL_F64 = 1.0e-10              # FP64 GPU form vs FP64 oracle


def rel_l2(x, ref):
    return float(np.linalg.norm(np.asarray(x) - ref) / np.linalg.norm(ref))


def run_mf(scene_or_arrays, dev, form="f32"):
    from geraf.ops import _native

    if isinstance(scene_or_arrays, mf_scenes.MFScene):
        sc = scene_or_arrays
        echo, tx, rx, vox, prm = sc.echo, sc.tx, sc.rx, sc.voxels, sc.params()
    else:
        echo, tx, rx, vox, prm = scene_or_arrays
    dt = torch.float64 if form == "f64" else torch.float32
    re = torch.as_tensor(np.ascontiguousarray(echo.real), device=dev, dtype=dt)
    im = torch.as_tensor(np.ascontiguousarray(echo.imag), device=dev, dtype=dt)
    t = torch.as_tensor(tx, device=dev, dtype=dt).contiguous()
    r = torch.as_tensor(rx, device=dev, dtype=dt).contiguous()
    v = torch.as_tensor(np.ascontiguousarray(np.asarray(vox).T), device=dev, dtype=dt)
    A, K = re.shape
    N = v.shape[1]
    o_r = torch.empty(N, device=dev, dtype=dt)
    o_i = torch.empty(N, device=dev, dtype=dt)
    params = _native.make_params(prm["f0"], prm["slope"], prm["fs"], prm["c"], prm["bias"])
    ptrs = list(map(_native.ptr, (re, im, t, r, o_r, o_i, v)))
    if form == "mma":
        size = int(_native.load().geraf_mf_forward_mma_workspace_bytes(A, K))
        ws = torch.empty(size, dtype=torch.uint8, device=dev)
        _native.call("geraf_mf_forward_f32_mma", *ptrs, A, K, N, params, _native.ptr(ws), size,
                     _native.stream_of(re))
    else:
        name = {"f32": "geraf_mf_forward_f32", "plain32": "geraf_mf_forward_plain_f32",
                "f64": "geraf_mf_forward_f64"}[form]
        _native.call(name, *ptrs, A, K, N, params, _native.stream_of(re))
    torch.cuda.synchronize()
    return o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy()


@pytest.fixture(scope="module")
def cfg1():
    sc = mf_scenes.config1(0)
    ref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels, **sc.params())
    return sc, ref


def test_cfg1_tuned_vs_fp64_oracle(cuda, cfg1):
    sc, ref = cfg1
    out = run_mf(sc, cuda)
    err = rel_l2(out, ref)
    print(f"cfg1 tuned rel-L2 vs FP64 oracle: {err:.3e}")
    assert err <= TOL_TUNED_VS_F64
    idx = mf_scenes.nearest_voxels(sc, sc.scatterers)
    peak = np.max(np.abs(np.abs(out[idx]) / np.abs(ref[idx]) - 1.0))
    phase = np.max(np.abs(np.angle(out[idx] / ref[idx])))
    print(f"cfg1 peak-magnitude err {peak:.3e}, phase err {phase:.3e} rad")
    assert peak < 1e-4 and phase < 1e-4


def test_cfg1_plain_forms(cuda, cfg1):
    sc, ref = cfg1
    ref32 = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels, form="ref32",
                           **sc.params())
    out32 = run_mf(sc, cuda, "plain32")
    assert rel_l2(out32, ref32) <= TOL_PLAIN32_VS_REF32
    out64 = run_mf(sc, cuda, "f64")
    assert rel_l2(out64, ref) <= TOL_F64


@pytest.mark.parametrize("A,K,N", [(1, 1, 1), (3, 37, 100), (7, 64, 65), (2, 32, 4097),
                                   (33, 96, 130), (5, 200, 1)])
def test_shapes_vs_oracle(cuda, A, K, N):
    rng = np.random.default_rng(A * 1000 + K * 10 + N)
    echo = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    tx = rng.uniform(-0.05, 0.05, (A, 3)).astype(np.float32)
    tx[:, 2] = 0.0
    rx = (tx + np.float32([0.002, 0, 0])).astype(np.float32)
    vox = rng.uniform(-0.1, 0.1, (N, 3)).astype(np.float32)
    vox[:, 2] += 0.3
    prm = dict(f0=mf_scenes.MF_F0, slope=float(np.float32(701500.0)), fs=1.25e6,
               c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)
    ref = orc.mf_forward(echo.real, echo.imag, tx, rx, vox, **prm)
    out = run_mf((echo, tx, rx, vox, prm), cuda)
    assert rel_l2(out, ref) <= 3e-6 * np.sqrt(K) + 1e-6


def _mma_case(A, K, N, mono, seed, row_scale=None):
    rng = np.random.default_rng(seed)
    echo = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    if row_scale is not None:
        echo = (echo * row_scale[:, None]).astype(np.complex64)
    tx = rng.uniform(-0.05, 0.05, (A, 3)).astype(np.float32)
    tx[:, 2] = 0.0
    rx = tx.copy() if mono else (tx + np.float32([0.002, 0.001, 0])).astype(np.float32)
    vox = rng.uniform(-0.1, 0.1, (N, 3)).astype(np.float32)
    vox[:, 2] += 0.3
    # a 77-81 GHz sweep over the K samples, as the workloads use
    fs = 1.25e6 * K / 64
    prm = dict(f0=mf_scenes.MF_F0, slope=float(np.float32(701500.0)), fs=fs,
               c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)
    return echo, tx, rx, vox, prm


@pytest.mark.parametrize("A,K,N,mono", [(1, 257, 1, False), (3, 300, 100, True),
                                        (5, 512, 1030, False), (2, 513, 65, True),
                                        (2, 1000, 700, False), (1, 1024, 257, True),
                                        (1, 2048, 300, False), (3, 2047, 129, True),
                                        (4, 1, 64, False), (6, 100, 63, True)])
def test_mma_forward_shapes_vs_oracle(cuda, A, K, N, mono):
    """The matrix-core forward over one to four 512-frequency tiles, partial
    tiles (K = 257, 513, 1000, 2047), K below one row of 32 (1, 100), ragged
    voxel counts (partial groups of 16 and waves of 64, N = 1), single
    antennas, monostatic and bistatic antennas, against the FP64 oracle."""
    echo, tx, rx, vox, prm = _mma_case(A, K, N, mono, seed=K * 7 + A)
    ref = orc.mf_forward(echo.real, echo.imag, tx, rx, vox, **prm)
    out = run_mf((echo, tx, rx, vox, prm), cuda, "mma")
    err = rel_l2(out, ref)
    print(f"mma A={A} K={K} N={N} mono={mono}: rel-L2 vs FP64 {err:.2e}")
    assert err <= TOL_TUNED_VS_F64


def test_mma_forward_per_antenna_scaling(cuda):
    """Echo rows 10^-30 .. 10^30 apart (the FP16 operands are scaled per
    antenna by a power of two), a row of zeros, and a row whose entries span
    twelve orders of magnitude: each antenna alone, and all together against
    the FP64 oracle."""
    A, K, N = 6, 512, 300
    scale = np.array([1e-30, 1e30, 0.0, 1.0, 3e-4, 7e5], dtype=np.float64)
    echo, tx, rx, vox, prm = _mma_case(A, K, N, False, seed=11, row_scale=scale)
    rng = np.random.default_rng(12)
    echo[3] = (echo[3] * 10.0 ** rng.uniform(-12, 0, K)).astype(np.complex64)
    for a in range(A):
        one = (echo[a:a + 1], tx[a:a + 1], rx[a:a + 1], vox, prm)
        ref = orc.mf_forward(echo[a:a + 1].real, echo[a:a + 1].imag, tx[a:a + 1], rx[a:a + 1],
                             vox, **prm)
        out = run_mf(one, cuda, "mma")
        if scale[a] == 0.0:
            assert np.all(out == 0)
            continue
        err = rel_l2(out, ref)
        print(f"antenna scale {scale[a]:.0e}: rel-L2 {err:.2e}")
        assert np.all(np.isfinite(out)) and err <= TOL_TUNED_VS_F64
    ref = orc.mf_forward(echo.real, echo.imag, tx, rx, vox, **prm)
    assert rel_l2(run_mf((echo, tx, rx, vox, prm), cuda, "mma"), ref) <= TOL_TUNED_VS_F64


def test_mma_forward_weak_scatterers_per_voxel(cuda):
    """A strong return and weak scatterers 60 and 80 dB below it (16x16
    aperture, K=512, FP64-synthesised echo): the weak scatterers' voxels,
    each on its own, against the FP64 oracle.  Their signal reaches every
    frequency bin at 1e-3 / 1e-4 of the strong one, so the FP16 hi/lo
    operands (2^-22 relative per entry of the row, scaled to the row
    maximum) leave them an absolute error of that order of the strong
    voxel's value: the bound is 2^-18 of the strong voxel's magnitude, and
    relative 1e-3 (-60 dB) / 1e-2 (-80 dB) on the weak voxels themselves."""
    K = 512
    fs = 1.25e6 * K / 64
    slope = float(np.float32(701500.0))
    ap = mf_scenes.planar_grid(16, 3.9e-3).astype(np.float32)
    c = np.array([0.0, 0.0, 0.30])
    rng = np.random.default_rng(21)
    offs = rng.uniform(-0.06, 0.06, (8, 3)) * np.array([1, 1, 0.5])
    scat = np.concatenate([c[None], c + offs])
    amp = np.array([1.0] + [1e-3] * 4 + [1e-4] * 4, dtype=np.complex128)
    amp[1:] *= np.exp(2j * np.pi * rng.uniform(0, 1, 8))
    echo = mf_scenes.synth_echo(ap, ap, scat, amp, K, fs, slope).numpy().astype(np.complex64)
    vox = scat.astype(np.float32)
    prm = dict(f0=mf_scenes.MF_F0, slope=slope, fs=fs, c=mf_scenes.SPEED_OF_LIGHT,
               bias=mf_scenes.PATH_BIAS)
    ref = orc.mf_forward(echo.real, echo.imag, ap, ap, vox, **prm)
    out = run_mf((echo, ap, ap, vox, prm), cuda, "mma")
    horner = run_mf((echo, ap, ap, vox, prm), cuda, "f32")
    strong = abs(ref[0])
    for name, sel, rel_tol in (("-60 dB", slice(1, 5), 1e-3), ("-80 dB", slice(5, 9), 1e-2)):
        err = np.abs(out[sel] - ref[sel])
        rel = err / np.abs(ref[sel])
        herr = np.abs(horner[sel] - ref[sel]) / np.abs(ref[sel])
        print(f"{name}: |ref| / strong {np.abs(ref[sel]).max() / strong:.1e}; matrix-core "
              f"max rel {rel.max():.2e} (abs / strong {err.max() / strong:.2e}); Horner max rel "
              f"{herr.max():.2e}", flush=True)
        assert np.all(err <= 2.0 ** -18 * strong) and np.all(rel <= rel_tol)
    assert abs(out[0] - ref[0]) <= 1e-5 * strong


def test_mma_forward_cfg1_and_drop_in_dispatch(cuda, cfg1):
    """The 16x16 / K=128 / 32^3 volume through the matrix-core entry point;
    and the drop-in op at K = 512 (it dispatches to the matrix-core form)
    against the Horner kernel."""
    from geraf.ops.RFMatchedFilter import matched_filter_cuda

    sc, ref = cfg1
    err = rel_l2(run_mf(sc, cuda, "mma"), ref)
    print(f"cfg1 through the matrix-core form: rel-L2 {err:.2e}")
    assert err <= TOL_TUNED_VS_F64
    echo, tx, rx, vox, prm = _mma_case(8, 512, 2000, False, seed=5)
    horner = run_mf((echo, tx, rx, vox, prm), cuda, "f32")
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    o_r, o_i = matched_filter_cuda(t(echo.real), t(echo.imag), t(tx), t(rx), t(vox), 8, 512,
                                   prm["fs"], prm["slope"], 2000)
    op = o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy()
    ref = orc.mf_forward(echo.real, echo.imag, tx, rx, vox, **prm)
    assert rel_l2(op, ref) <= TOL_TUNED_VS_F64
    assert rel_l2(op, horner) <= 2e-6 and not np.array_equal(op, horner)
    # K = 192 (above the op's switch to the matrix-core form at K = 161)
    echo, tx, rx, vox, prm = _mma_case(5, 192, 900, True, seed=6)
    o_r, o_i = matched_filter_cuda(t(echo.real), t(echo.imag), t(tx), t(rx), t(vox), 5, 192,
                                   prm["fs"], prm["slope"], 900)
    op = o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy()
    assert rel_l2(op, run_mf((echo, tx, rx, vox, prm), cuda, "mma")) == 0.0
    assert rel_l2(op, orc.mf_forward(echo.real, echo.imag, tx, rx, vox, **prm)) <= TOL_TUNED_VS_F64


def test_empty_volume(cuda):
    from geraf.ops.RFMatchedFilter import matched_filter_cuda

    re = torch.randn(4, 8, device=cuda)
    im = torch.randn(4, 8, device=cuda)
    pos = torch.zeros(4, 3, device=cuda)
    vox = torch.zeros(0, 3, device=cuda)
    o_r, o_i = matched_filter_cuda(re, im, pos, pos, vox, 4, 8, 1.25e6, 701500.0, 0)
    assert o_r.shape == (0,) and o_i.shape == (0,)


def _adjoint_case(dev, A=6, K=64, N=300, seed=3):
    rng = np.random.default_rng(seed)
    tx = rng.uniform(-0.05, 0.05, (A, 3)).astype(np.float32)
    rx = (tx + np.float32([0.001, 0.0005, 0])).astype(np.float32)
    vox = rng.uniform(-0.1, 0.1, (N, 3)).astype(np.float32)
    vox[:, 2] += 0.3
    g = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    return tx, rx, vox, g


def test_backward_vs_oracle_and_adjoint_identity(cuda):
    from geraf.ops.RFMatchedFilter import matched_filter_cuda

    A, K, N = 6, 64, 300
    tx, rx, vox, g = _adjoint_case(cuda, A, K, N)
    fs, As = 1.25e6, 701500.0
    prm = dict(f0=mf_scenes.MF_F0, slope=As, fs=fs, c=mf_scenes.SPEED_OF_LIGHT,
               bias=mf_scenes.PATH_BIAS)
    rng = np.random.default_rng(9)
    s = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    re = torch.tensor(s.real, device=cuda, requires_grad=True)
    im = torch.tensor(s.imag, device=cuda, requires_grad=True)
    t = torch.tensor(tx, device=cuda)
    r = torch.tensor(rx, device=cuda)
    v = torch.tensor(vox, device=cuda)
    o_r, o_i = matched_filter_cuda(re, im, t, r, v, A, K, fs, As, N)
    (o_r * torch.tensor(g.real, device=cuda) + o_i * torch.tensor(g.imag, device=cuda)).sum() \
        .backward()
    grad = re.grad.double().cpu().numpy() + 1j * im.grad.double().cpu().numpy()
    # d/ds of Re<out, g> is the adjoint applied to g (real-linear pairing).
    ref = orc.mf_backward(g, tx, rx, vox, K, norm=A, **prm)
    assert rel_l2(grad, ref) <= 1e-5
    # <MF s, g>_R == <s, MF^T g>_R in the oracle itself, and for the GPU pair.
    fwd = orc.mf_forward(s.real, s.imag, tx, rx, vox, **prm)
    lhs = np.sum(fwd.real * g.real + fwd.imag * g.imag)
    rhs = np.sum(s.real * ref.real + s.imag * ref.imag)
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    out = o_r.detach().double().cpu().numpy() + 1j * o_i.detach().double().cpu().numpy()
    lhs_gpu = np.sum(out.real * g.real + out.imag * g.imag)
    rhs_gpu = np.sum(s.real * grad.real + s.imag * grad.imag)
    assert abs(lhs_gpu - rhs_gpu) <= 1e-5 * abs(lhs_gpu)


@pytest.mark.parametrize("A,K,N", [(5, 200, 1500), (3, 300, 1100), (4, 512, 1030),
                                   (2, 1000, 700), (2, 1030, 600), (1, 129, 1),
                                   (2, 257, 513), (1, 1024, 511)])
def test_backward_matrix_tiles(cuda, A, K, N):
    """The matrix-core adjoint (128 < K <= 1024: one, two and four 256-sample
    tiles, partial tiles, the first K of each tile count, several voxel tiles
    with a partial last one, a single antenna and voxel) and the rotation
    adjoint beyond (K = 1030), against the FP64 oracle."""
    tx, rx, vox, g = _adjoint_case(cuda, A, K, N, seed=K)
    fs, As = 1.25e6 * K / 64, 701500.0
    prm = dict(f0=mf_scenes.MF_F0, slope=As, fs=fs, c=mf_scenes.SPEED_OF_LIGHT,
               bias=mf_scenes.PATH_BIAS)
    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    g_r, g_i = T(g.real), T(g.imag)
    o_r = torch.empty((A, K), device=cuda)
    o_i = torch.empty_like(o_r)
    from geraf.ops import _native

    params = _native.make_params(prm["f0"], prm["slope"], prm["fs"], prm["c"], prm["bias"])
    _native.call("geraf_mf_backward_f32", *map(_native.ptr, (g_r, g_i, T(tx), T(rx), o_r, o_i,
                                                                 T(vox.T))),
                 A, K, N, A, params, _native.stream_of(o_r))
    torch.cuda.synchronize()
    got = o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy()
    ref = orc.mf_backward(g, tx, rx, vox, K, norm=A, **prm)
    err = rel_l2(got, ref)
    print(f"adjoint A{A} K{K} N{N}: rel-L2 {err:.2e}")
    assert err <= 1e-5


def test_faster_matched_filter_equivalence(cuda):
    from geraf.ops.RFMatchedFilter import faster_matched_filter_cuda, matched_filter_cuda

    A, K, N, sel = 8, 64, 257, 3
    tx, rx, vox, g = _adjoint_case(cuda, A, K, N, seed=5)
    rng = np.random.default_rng(11)
    s = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    fs, As = 1.25e6, 701500.0
    re = torch.tensor(s.real, device=cuda, requires_grad=True)
    im = torch.tensor(s.imag, device=cuda, requires_grad=True)
    t, r, v = (torch.tensor(x, device=cuda) for x in (tx, rx, vox))
    o_r, o_i = matched_filter_cuda(re, im, t, r, v, A, K, fs, As, N)
    gr = torch.tensor(g.real, device=cuda)
    gi = torch.tensor(g.imag, device=cuda)
    (o_r * gr + o_i * gi).sum().backward()
    re2 = torch.tensor(s.real[:sel], device=cuda, requires_grad=True)
    im2 = torch.tensor(s.imag[:sel], device=cuda, requires_grad=True)
    f_r, f_i = faster_matched_filter_cuda(re2, im2, torch.tensor(s.real[sel:], device=cuda),
                                          torch.tensor(s.imag[sel:], device=cuda), t[:sel],
                                          r[:sel], t[sel:], r[sel:], v, A, sel, K, fs, As, N)
    (f_r * gr + f_i * gi).sum().backward()
    torch.testing.assert_close(f_r, o_r, rtol=0, atol=0)
    torch.testing.assert_close(f_i, o_i, rtol=0, atol=0)
    torch.testing.assert_close(re2.grad, re.grad[:sel], rtol=1e-5, atol=1e-7)
    torch.testing.assert_close(im2.grad, im.grad[:sel], rtol=1e-5, atol=1e-7)


@pytest.mark.parametrize("K", [64, 512])
def test_float64_and_half_dispatch(cuda, K):
    """float64 runs the FP64 form; float16 is computed in FP32 (K = 512: the
    matrix-core form) and returned as float16."""
    from geraf.ops.RFMatchedFilter import matched_filter_cuda

    A, N = 4, 100
    tx, rx, vox, _ = _adjoint_case(cuda, A, K, N, seed=7)
    rng = np.random.default_rng(1)
    s = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K)))
    fs, As = 1.25e6 * K / 64, 701500.0
    prm = dict(f0=mf_scenes.MF_F0, slope=As, fs=fs, c=mf_scenes.SPEED_OF_LIGHT,
               bias=mf_scenes.PATH_BIAS)
    ref = orc.mf_forward(s.real, s.imag, tx, rx, vox, **prm)
    args = [torch.tensor(x, device=cuda, dtype=torch.float64) for x in (s.real, s.imag, tx, rx, vox)]
    o_r, o_i = matched_filter_cuda(*args, A, K, fs, As, N)
    assert o_r.dtype == torch.float64
    assert rel_l2(o_r.cpu().numpy() + 1j * o_i.cpu().numpy(), ref) <= 1e-10
    h = [a.half() for a in args]
    o_r, o_i = matched_filter_cuda(*h, A, K, fs, As, N)
    assert o_r.dtype == torch.float16
    ref_h = orc.mf_forward(*(a.double().cpu().numpy() for a in h[:2]), *(a.float().cpu().numpy()
                           for a in h[2:]), **prm)
    assert rel_l2(o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy(), ref_h) < 2e-3


def test_cfg2_subvolume_vs_fp64(cuda):
    sc = mf_scenes.config2(0, device=str(cuda))
    rng = np.random.default_rng(0)
    sel = np.unique(np.concatenate([mf_scenes.nearest_voxels(sc, sc.scatterers[::25]),
                                    rng.integers(0, sc.N, 200)]))
    arrays = (sc.echo, sc.tx, sc.rx, sc.voxels[sel], sc.params())
    out = run_mf(arrays, cuda)
    out64 = run_mf(arrays, cuda, "f64")
    err = rel_l2(out, out64)
    print(f"cfg2 subset ({len(sel)} voxels) tuned rel-L2 vs GPU FP64: {err:.3e}")
    assert err <= TOL_TUNED_VS_F64
    few = sel[:8]
    ref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels[few], **sc.params())
    assert rel_l2(out64[:8], ref) <= 1e-9


# (A, R, N, spacing): 128 bins in one adjoint chunk; 5,200 bins in three
# chunks of 2,048 with the paths (~0.72-0.78 m) in bins ~2,400-2,600, i.e. in
# the second chunk (the reference skips d > spacing * R, so bins above R/2
# are never reached)
RANGE_BIN_CASES = [(12, 128, 3000, 0.008), (6, 5200, 2000, 1.5e-4)]


@pytest.mark.parametrize("case", RANGE_BIN_CASES)
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_range_bin_mf_vs_oracle(cuda, dtype, case):
    from geraf.ops.RFMatchedFilter import _C

    A, R, N, spacing = case
    rng = np.random.default_rng(21)
    tx = rng.uniform(-0.03, 0.03, (A, 3)).astype(np.float32)
    tx[:, 2] = 0
    rx = (tx + np.float32([0.001, 0, 0])).astype(np.float32)
    vox = (rng.uniform(-0.08, 0.08, (N, 3)) + [0, 0, 0.3]).astype(np.float32)
    s = (rng.standard_normal((A, R)) + 1j * rng.standard_normal((A, R))).astype(np.complex64)
    g = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    spacing = float(np.float32(spacing))
    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda, dtype=dtype)  # noqa: E731
    o_r = torch.empty(N, device=cuda, dtype=dtype)
    o_i = torch.empty_like(o_r)
    _C.MatchedFilter_fft_forward(T(s.real), T(s.imag), T(tx), T(rx), o_r, o_i, T(vox.T), A, R,
                                 spacing, 1.25e6, 701500.0, N)
    prm = dict(f0=mf_scenes.MF_F0, c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)
    ref = orc.mf_fft_forward(s.real, s.imag, tx, rx, vox, spacing, **prm)
    tol = 1e-6 if dtype == torch.float32 else 1e-12
    assert rel_l2(o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy(), ref) <= tol
    g_r = torch.empty((A, R), device=cuda, dtype=dtype)
    g_i = torch.empty_like(g_r)
    _C.MatchedFilter_fft_backward_realimag(T(g.real), T(g.imag), T(s.real), T(s.imag), T(tx),
                                           T(rx), g_r, g_i, T(vox.T), A, R, spacing, 1.25e6,
                                           701500.0, N)
    refg = orc.mf_fft_backward(g, tx, rx, vox, R, spacing, A, **prm)
    assert np.count_nonzero(refg) > 0
    assert rel_l2(g_r.double().cpu().numpy() + 1j * g_i.double().cpu().numpy(), refg) <= tol
    if dtype == torch.float32:
        # fixed-point bin sums: bit for bit the same on a second run
        g_r2, g_i2 = torch.empty_like(g_r), torch.empty_like(g_i)
        _C.MatchedFilter_fft_backward_realimag(T(g.real), T(g.imag), T(s.real), T(s.imag),
                                               T(tx), T(rx), g_r2, g_i2, T(vox.T), A, R,
                                               spacing, 1.25e6, 701500.0, N)
        assert torch.equal(g_r, g_r2) and torch.equal(g_i, g_i2)


def test_range_bin_adjoint_non_finite_gradient(cuda):
    """A NaN in the output gradient reaches the bins its voxel falls in (the
    FP32 form then sums in FP64, as the FP64 form does); the other bins keep
    their values."""
    from geraf.ops.RFMatchedFilter import _C

    rng = np.random.default_rng(22)
    A, R, N = 4, 128, 500
    tx = rng.uniform(-0.03, 0.03, (A, 3)).astype(np.float32)
    tx[:, 2] = 0
    rx = (tx + np.float32([0.001, 0, 0])).astype(np.float32)
    vox = (rng.uniform(-0.08, 0.08, (N, 3)) + [0, 0, 0.3]).astype(np.float32)
    g = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    g[7] = np.nan
    spacing = float(np.float32(0.008))
    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    g_r = torch.empty((A, R), device=cuda)
    g_i = torch.empty_like(g_r)
    z = torch.zeros((A, R), device=cuda)
    _C.MatchedFilter_fft_backward_realimag(T(g.real), T(g.imag), z, z, T(tx), T(rx), g_r, g_i,
                                           T(vox.T), A, R, spacing, 1.25e6, 701500.0, N)
    prm = dict(f0=mf_scenes.MF_F0, c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)
    refg = orc.mf_fft_backward(g, tx, rx, vox, R, spacing, A, **prm)
    out = g_r.double().cpu().numpy() + 1j * g_i.double().cpu().numpy()
    bad = np.isnan(refg)
    assert bad.any() and np.array_equal(np.isnan(out), bad)
    assert rel_l2(out[~bad], refg[~bad]) <= 1e-6


def test_adc_bank_selected_rows_matched_filter(cuda):
    """The device bank's selected-rows matched filter equals the full matched
    filter of the mixed bank, in value and in the selected rows' gradients."""
    from geraf.models.model_utils.adc_bank import ADCBank
    from rf_workloads import mf_scenes

    sc = mf_scenes.config1(0)
    A, K = sc.A, sc.K
    t_pos = torch.tensor(sc.tx, device=cuda)
    r_pos = torch.tensor(sc.rx, device=cuda)
    pts = torch.tensor(sc.voxels[::37], device=cuda)
    radar = dict(numADCSample=K, adcSampleRate=sc.fs, As_sci=sc.As_sci)
    bank = ADCBank(A, K, 4, cuda)
    rng = np.random.default_rng(3)
    for g in range(4):  # fill every group once
        n = len(bank.groups[g])
        bank.store(g, torch.tensor(rng.standard_normal((n, K)), dtype=torch.float32, device=cuda),
                   torch.tensor(rng.standard_normal((n, K)), dtype=torch.float32, device=cuda))
    bank.advance()
    n = len(bank.selected)
    sim_r = torch.tensor(rng.standard_normal((n, K)), dtype=torch.float32, device=cuda)
    sim_i = torch.tensor(rng.standard_normal((n, K)), dtype=torch.float32, device=cuda)
    outs = []
    for fast in (True, False):
        r = sim_r.clone().requires_grad_(True)
        i = sim_i.clone().requires_grad_(True)
        yr, yi = bank.matched_filter(r, i, t_pos, r_pos, pts, radar, selected_rows_only=fast)
        (yr.square().sum() + 0.5 * yi.sum()).backward()
        outs.append((yr.detach(), yi.detach(), r.grad, i.grad))
    for a, b in zip(*outs):
        rel = float((a - b).norm() / b.norm())
        assert rel < 1e-6, rel


def test_backproject_volume_matches_matched_filter(cuda, tmp_path):
    from geraf.models.model_utils import mf_volume as mv
    from geraf.ops.RFMatchedFilter import matched_filter_cuda
    from rf_workloads import mf_scenes

    sc = mf_scenes.config1(0)
    A, K = sc.A, sc.K
    re = torch.tensor(np.ascontiguousarray(sc.echo.real), device=cuda)
    im = torch.tensor(np.ascontiguousarray(sc.echo.imag), device=cuda)
    tx, rx = torch.tensor(sc.tx, device=cuda), torch.tensor(sc.rx, device=cuda)
    radar = dict(numADCSample=K, adcSampleRate=sc.fs, As_sci=sc.As_sci)
    bound, res = [-0.02, 0.02, -0.02, 0.02, 0.23, 0.27], [0.004, 0.004, 0.004]
    vol = mv.backproject(re, im, tx, rx, bound, res, radar, slab_voxels=300,
                         algo="direct")  # several slabs
    ax = [torch.tensor(a, dtype=torch.float32, device=cuda) for a in mv.grid_axes(bound, res)]
    g = torch.meshgrid(*ax, indexing="ij")
    vox = torch.stack([x.reshape(-1) for x in g], 1)
    r2, i2 = matched_filter_cuda(re, im, tx, rx, vox, A, K, sc.fs, sc.As_sci, vox.shape[0])
    # slabs of different sizes launch with different antenna groupings, so the
    # partial sums combine in another order: equal to FP32 rounding
    ref = torch.hypot(r2, i2)
    assert float((vol.reshape(-1) - ref).norm() / ref.norm()) < 1e-6
    # the producer writes the same cube
    d = tmp_path / "0_0_0"
    d.mkdir()
    for name, arr in (("adcr.npy", sc.echo.real), ("adci.npy", sc.echo.imag),
                      ("tpos.npy", sc.tx), ("rpos.npy", sc.rx)):
        np.save(d / name, np.asarray(arr, dtype=np.float32))
    cube = mv.produce_frame(str(d), bound, res, radar, device=str(cuda), algo="direct")
    assert np.array_equal(mv.read_sarimage(str(d / "sarimage.bin"), bound, res), cube)
    v = vol.cpu().numpy()
    assert np.linalg.norm(cube - v) / np.linalg.norm(v) < 1e-6


# ------------------------------------------------------------------ tiled (expansion) path

def _tiled(re, im, tx, rx, vox_nx3, K, fs, As, tol=1e-6):
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import tiled_forward
    return tiled_forward(re, im, tx, rx, vox_nx3, K, fs, As, tol)


def test_tiled_cfg1_vs_fp64_oracle(cuda):
    sc = mf_scenes.config1(0)
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    o_r, o_i, info = _tiled(t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx), t(sc.voxels),
                            sc.K, sc.fs, sc.As_sci)
    print(f"cfg1 tiled: {info}", flush=True)
    assert info["path"] == "tiled"
    ref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels, **sc.params())
    got = torch.complex(o_r.double(), o_i.double()).cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"cfg1 tiled rel-L2 vs FP64 oracle: {rel:.3e}", flush=True)
    assert rel <= 1e-5


@pytest.mark.parametrize("cap", [None, 1 << 20])
def test_tiled_cfg2_block_vs_gpu_fp64(cuda, cap, monkeypatch):
    """A compact 24^3 block of the 64x64 MIMO, K=512 grid (ragged: 13,824
    voxels = 27 tiles), against the plain FP64 GPU form; with a 1 MiB
    workspace the antennas run in many chunks."""
    from geraf.ops import _native
    from geraf.ops.RFMatchedFilter import tiled_matched_filter as tmf
    if cap is not None:
        monkeypatch.setattr(tmf, "WORKSPACE_CAP", cap)
    sc = mf_scenes.config2(0, device=str(cuda))
    g = sc.voxels.reshape(128, 128, 128, 3)[50:74, 40:64, 60:84].reshape(-1, 3)
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    re, im, tx, rx = t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx)
    o_r, o_i, info = _tiled(re, im, tx, rx, t(g), sc.K, sc.fs, sc.As_sci)
    print(f"cfg2 block tiled: {info}", flush=True)
    assert info["path"] == "tiled"
    d = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda, dtype=torch.float64)  # noqa
    n = g.shape[0]
    r64 = torch.empty(n, device=cuda, dtype=torch.float64)
    i64 = torch.empty_like(r64)
    re64, im64, tx64, rx64, g64 = d(sc.echo.real), d(sc.echo.imag), d(sc.tx), d(sc.rx), d(g.T)
    prm = _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias)
    _native.call("geraf_mf_forward_f64", *map(_native.ptr, (re64, im64, tx64, rx64, r64, i64, g64)),
                 sc.A, sc.K, n, prm, _native.stream_of(re64))
    ref = torch.complex(r64, i64)
    got = torch.complex(o_r.double(), o_i.double())
    rel = float((got - ref).norm() / ref.norm())
    print(f"cfg2 block tiled rel-L2 vs GPU FP64 (cap {cap}): {rel:.3e}", flush=True)
    assert rel <= 1e-5


def test_tiled_small_and_fallback_cases(cuda):
    from rf_workloads import mf_scenes as ms
    sc = ms.config1(0)
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    re, im, tx, rx = t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx)
    from geraf.ops.RFMatchedFilter import _C as mfC
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import TilePlan, tiled_forward
    for tile in (512, 128):
        for n in (1, 127, 129, 511, 513):
            v = t(sc.voxels[1000:1000 + n])
            plan = TilePlan(v, sc.K, mfC._params(sc.fs, sc.As_sci), 1e-6, tile=tile)
            o_r, o_i, info = tiled_forward(re, im, tx, rx, v, sc.K, sc.fs, sc.As_sci, 1e-6,
                                           plan=plan)
            ref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx,
                                 sc.voxels[1000:1000 + n], **sc.params())
            got = torch.complex(o_r.double(), o_i.double()).cpu().numpy()
            rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
            assert rel <= 1e-5, (tile, n, rel, info)
    # K = 1 runs the direct kernel
    o_r, o_i, info = _tiled(re[:, :1].contiguous(), im[:, :1].contiguous(), tx, rx,
                            t(sc.voxels[:600]), 1, sc.fs, sc.As_sci)
    assert info["path"] == "direct"
    # scattered points far apart: tiles too large for 32 terms -> direct kernel
    far = np.random.default_rng(0).uniform(-2, 2, (700, 3)).astype(np.float32)
    far[:, 2] += 3.0
    _, _, info = _tiled(re, im, tx, rx, t(far), sc.K, sc.fs, sc.As_sci)
    assert info["path"] == "direct"


def test_tiled_op_gradient_is_the_adjoint(cuda):
    from geraf.ops.RFMatchedFilter import matched_filter_cuda
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import matched_filter_tiled
    sc = mf_scenes.config1(0)
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    vox = t(sc.voxels[:4096])
    outs = []
    for fn in (matched_filter_tiled, matched_filter_cuda):
        re = t(sc.echo.real).requires_grad_(True)
        im = t(sc.echo.imag).requires_grad_(True)
        yr, yi = fn(re, im, t(sc.tx), t(sc.rx), vox, sc.A, sc.K, sc.fs, sc.As_sci, 4096)
        (yr * 0.3 - yi * 0.7).sum().backward()
        outs.append((yr.detach(), yi.detach(), re.grad, im.grad))
    for a, b in zip(*outs):
        assert float((a - b).norm() / b.norm()) < 1e-5


@pytest.mark.parametrize("tile", [512, 128])
@pytest.mark.parametrize("which", ["cfg1", "cfg2_block"])
def test_tiled_adjoint_matches_direct_and_is_the_transpose(cuda, which, tile):
    from geraf.ops import _native
    from geraf.ops.RFMatchedFilter import _C as mfC
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import (TilePlan, tiled_adjoint,
                                                                tiled_forward)
    if which == "cfg1":
        sc = mf_scenes.config1(0)
        vox = sc.voxels
    else:
        sc = mf_scenes.config2(0, device=str(cuda))
        vox = sc.voxels.reshape(128, 128, 128, 3)[50:74, 40:64, 60:84].reshape(-1, 3)
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    A, K, N = sc.A, sc.K, vox.shape[0]
    re, im, tx, rx, v = t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx), t(vox)
    rng = np.random.default_rng(5)
    g_r = t(rng.standard_normal(N).astype(np.float32))
    g_i = t(rng.standard_normal(N).astype(np.float32))
    prm = mfC._params(sc.fs, sc.As_sci)
    plan = TilePlan(v, K, prm, 1e-6, tile=tile)
    assert plan.usable and plan.tile == tile
    a_r, a_i = tiled_adjoint(g_r, g_i, tx, rx, plan, A, K, A, sc.fs, sc.As_sci, v)
    d_r, d_i = torch.empty((A, K), device=cuda), torch.empty((A, K), device=cuda)
    vt = v.t().contiguous()
    _native.call("geraf_mf_backward_f32", *map(_native.ptr, (g_r, g_i, tx, rx, d_r, d_i, vt)),
                 A, K, N, A, _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias),
                 _native.stream_of(g_r))
    got, ref = torch.complex(a_r.double(), a_i.double()), torch.complex(d_r.double(), d_i.double())
    rel = float((got - ref).norm() / ref.norm())
    # <MF s, g> = <s, MF^T g> for the tiled pair (MF^T divided by A like MF)
    f_r, f_i, _ = tiled_forward(re, im, tx, rx, v, K, sc.fs, sc.As_sci, 1e-6, plan=plan)
    lhs = float((f_r.double() * g_r.double() + f_i.double() * g_i.double()).sum())
    rhs = float((re.double() * a_r.double() + im.double() * a_i.double()).sum())
    print(f"{which} tile {tile}: tiled adjoint vs direct rel-L2 {rel:.3e} ({plan.terms} terms); "
          f"<MF s,g>={lhs:.6e} <s,MF^T g>={rhs:.6e}", flush=True)
    assert rel <= 1e-5
    assert abs(lhs - rhs) <= 1e-5 * (abs(lhs) + abs(rhs))
    # the forward of this plan against the direct kernel
    o_r, o_i = torch.empty(N, device=cuda), torch.empty(N, device=cuda)
    _native.call("geraf_mf_forward_f32", *map(_native.ptr, (re, im, tx, rx, o_r, o_i, vt)), A, K,
                 N, _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias),
                 _native.stream_of(re))
    fwd = float((torch.complex(f_r.double(), f_i.double()) -
                 torch.complex(o_r.double(), o_i.double())).norm() /
                torch.complex(o_r.double(), o_i.double()).norm())
    assert fwd <= 1e-5, fwd


@pytest.mark.parametrize("K,tile", [(2, 512), (37, 512), (100, 512), (300, 512), (1100, 512),
                                    (1100, 128)])
def test_tiled_pair_vs_fp64_oracle_over_k(cuda, K, tile):
    """Tiled forward and adjoint against the FP64 oracle for K that is not a
    multiple of the kernels' steps: partial 32-sample steps of the moments,
    1, 2 and 4 waves of 64 frequencies per block in the adjoint evaluation,
    a second 1024-sample echo chunk (K = 1100).  The sample rate grows with K
    so that the bandwidth, and so the term count, stays that of the
    config-1 chirp; 64 antennas, 2,048 voxels."""
    from geraf.ops.RFMatchedFilter import _C as mfC
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import (TilePlan, tiled_adjoint,
                                                                tiled_forward)
    sc = mf_scenes.config1(0)
    A, N = 64, 2048
    fs = sc.fs * max(1.0, (K - 1) / (sc.K - 1))
    prm = dict(sc.params(), fs=fs)
    rng = np.random.default_rng(K)
    echo = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    vox = np.ascontiguousarray(sc.voxels[5000:5000 + N])
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    re, im, tx, rx, v = t(echo.real), t(echo.imag), t(sc.tx[:A]), t(sc.rx[:A]), t(vox)
    plan = TilePlan(v, K, mfC._params(fs, sc.As_sci), 1e-6, tile=tile)
    assert plan.usable and plan.tiled_fraction == 1.0, (K, tile, plan.tiled_fraction)
    o_r, o_i, info = tiled_forward(re, im, tx, rx, v, K, fs, sc.As_sci, 1e-6, plan=plan)
    assert info["path"] == "tiled"
    ref = orc.mf_forward(echo.real, echo.imag, sc.tx[:A], sc.rx[:A], vox, **prm)
    got = torch.complex(o_r.double(), o_i.double()).cpu().numpy()
    frel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    g = (rng.standard_normal(N) + 1j * rng.standard_normal(N)).astype(np.complex64)
    a_r, a_i = tiled_adjoint(t(g.real), t(g.imag), tx, rx, plan, A, K, A, fs, sc.As_sci, v)
    aref = orc.mf_backward(g, sc.tx[:A], sc.rx[:A], vox, K, norm=A, **prm)
    agot = torch.complex(a_r.double(), a_i.double()).cpu().numpy()
    arel = np.linalg.norm(agot - aref) / np.linalg.norm(aref)
    print(f"K={K} tile {tile}: {plan.terms} terms; forward rel-L2 {frel:.2e}, adjoint {arel:.2e}",
          flush=True)
    assert frel <= 1e-5 and arel <= 1e-5, (frel, arel)


@pytest.mark.parametrize("tile", [512, 128])
def test_tiled_mixed_plan_forward_and_adjoint(cuda, tile):
    """A compact block plus scattered far points: the block's tiles run the
    expansion, the scattered tiles the direct kernel, in both directions."""
    from geraf.ops import _native
    from geraf.ops.RFMatchedFilter import _C as mfC
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import (TilePlan, tiled_adjoint,
                                                                tiled_forward)
    sc = mf_scenes.config1(0)
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    far = np.random.default_rng(2).uniform(-1.5, 1.5, (1500, 3)).astype(np.float32)
    far[:, 2] += 2.0
    vox = np.concatenate([sc.voxels[:8192], far])
    A, K, N = sc.A, sc.K, vox.shape[0]
    re, im, tx, rx, v = t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx), t(vox)
    plan = TilePlan(v, K, mfC._params(sc.fs, sc.As_sci), 1e-6, tile=tile)
    assert plan.usable and 0.5 < plan.tiled_fraction < 1.0, plan.tiled_fraction
    o_r, o_i, info = tiled_forward(re, im, tx, rx, v, K, sc.fs, sc.As_sci, 1e-6, plan=plan)
    ref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, vox, **sc.params())
    got = torch.complex(o_r.double(), o_i.double()).cpu().numpy()
    rel = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    rng = np.random.default_rng(9)
    g_r, g_i = t(rng.standard_normal(N).astype(np.float32)), t(rng.standard_normal(N).astype(np.float32))
    a_r, a_i = tiled_adjoint(g_r, g_i, tx, rx, plan, A, K, A, sc.fs, sc.As_sci, v)
    d_r, d_i = torch.empty((A, K), device=cuda), torch.empty((A, K), device=cuda)
    vt = v.t().contiguous()
    _native.call("geraf_mf_backward_f32", *map(_native.ptr, (g_r, g_i, tx, rx, d_r, d_i, vt)),
                 A, K, N, A, _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias),
                 _native.stream_of(g_r))
    arel = float((torch.complex(a_r, a_i) - torch.complex(d_r, d_i)).norm() /
                 torch.complex(d_r, d_i).norm())
    print(f"mixed plan: tiled fraction {plan.tiled_fraction:.3f}, buckets "
          f"{[b.terms for b in plan.buckets]}; forward rel-L2 {rel:.2e}, adjoint {arel:.2e}",
          flush=True)
    assert rel <= 1e-5 and arel <= 1e-5
