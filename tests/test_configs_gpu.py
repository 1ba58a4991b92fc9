"""Parity at the full workload sizes, on the paths the benchmark and the step
workloads run:

* the 64x64 MIMO, K=512, 128^3 matched filter (the full echo, A = 4,096):
  the matrix-core forward, the drop-in op and the adjoint against the GPU FP64
  forms on >= 4,096 voxels (every scatterer-nearest voxel among them), the
  FP64 forms pinned to the CPU oracle;
* the stage-1-style step (8,192 rays x 64 samples, valid counts 16..64,
  K=64): the network-produced tracer inputs through the fused tracer, forward
  and all five gradients against the FP64 oracle on 256 rays, and the step's
  matched-filter prediction against the CPU oracle's per-ray backprojection on
  1,024 target points;
* the stage-2-style tracer (32,768 rays x 256 samples, K=512, calibration
  offsets): forward and gradients on 256 rays against the FP64 oracle;
* the 256^3 target (128x128 aperture, K=1024: two matrix-core tiles): the
  direct kernel and the tiled form against the GPU FP64 form on 256 voxels of
  the central 64^3.
"""
import numpy as np
import pytest
import torch

from oracle import rf as orc
from rf_workloads import mf_scenes
from rf_workloads import tracer_scenes as ts
from test_mf_gpu import rel_l2, run_mf
from test_tracer_gpu import _lensless_oracle, rel

pytestmark = pytest.mark.gpu

TOL_MF = 1e-5     # FP32 matched filter vs FP64, volume rel-L2
TOL_FWD = 1e-5    # tracer forward vs FP64 oracle
TOL_GRAD = 1e-4   # tracer gradients vs FP64 oracle


@pytest.fixture(scope="module")
def mimo(cuda):
    return mf_scenes.config2(0, device=str(cuda))


def _scatterer_errors(out, ref, n_near):
    peak = float(np.max(np.abs(np.abs(out[:n_near]) / np.abs(ref[:n_near]) - 1.0)))
    phase = float(np.max(np.abs(np.angle(out[:n_near] / ref[:n_near]))))
    return peak, phase


def test_mimo64_k512_forward_full_echo(cuda, mimo):
    from geraf.ops.RFMatchedFilter import matched_filter_cuda

    sc = mimo
    near = np.unique(mf_scenes.nearest_voxels(sc, sc.scatterers))
    rng = np.random.default_rng(7)
    extra = np.setdiff1d(rng.integers(0, sc.N, 1024), near)
    sel = np.concatenate([near, extra])
    assert sel.size >= 4096 and sc.A == 4096 and sc.K == 512
    print(f"64x64 MIMO K=512: {sel.size} voxels ({near.size} scatterer-nearest)", flush=True)
    arrays = (sc.echo, sc.tx, sc.rx, sc.voxels[sel], sc.params())
    mma = run_mf(arrays, cuda, "mma")
    f64 = run_mf(arrays, cuda, "f64")
    err = rel_l2(mma, f64)
    peak, phase = _scatterer_errors(mma, f64, near.size)
    print(f"  matrix-core forward vs GPU FP64: rel-L2 {err:.2e}, peak {peak:.2e}, "
          f"phase {phase:.2e} rad", flush=True)
    assert err <= TOL_MF and peak <= 1e-4 and phase <= 1e-4
    # the drop-in op runs the same kernel for K = 512: bit-identical
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    o_r, o_i = matched_filter_cuda(t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx),
                                   t(sc.voxels[sel]), sc.A, sc.K, sc.fs, sc.As_sci, sel.size)
    op = o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy()
    assert np.array_equal(op, mma)
    # the GPU FP64 form against the CPU oracle on a few voxels
    few = np.concatenate([sel[:4], sel[-4:]])
    ref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels[few], **sc.params())
    pin = rel_l2(np.concatenate([f64[:4], f64[-4:]]), ref)
    print(f"  GPU FP64 vs CPU oracle on 8 voxels: {pin:.2e}", flush=True)
    assert pin <= 1e-9


def test_mimo64_k512_adjoint_full_echo(cuda, mimo):
    from geraf.ops import _native

    sc = mimo
    rng = np.random.default_rng(8)
    sel = np.unique(np.concatenate([mf_scenes.nearest_voxels(sc, sc.scatterers[::2]),
                                    rng.integers(0, sc.N, 2048)]))
    assert sel.size >= 4096
    g = (rng.standard_normal(sel.size) + 1j * rng.standard_normal(sel.size))
    prm = _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias)
    outs = []
    for dt, name in ((torch.float32, "geraf_mf_backward_f32"),
                     (torch.float64, "geraf_mf_backward_f64")):
        T = lambda x: torch.tensor(np.ascontiguousarray(x), device=cuda, dtype=dt)  # noqa
        ins = [T(g.real), T(g.imag), T(sc.tx), T(sc.rx)]
        o_r = torch.empty((sc.A, sc.K), device=cuda, dtype=dt)
        o_i = torch.empty_like(o_r)
        vox = T(sc.voxels[sel].T)
        _native.call(name, *map(_native.ptr, ins + [o_r, o_i, vox]), sc.A, sc.K, sel.size, sc.A,
                     prm, _native.stream_of(o_r))
        torch.cuda.synchronize()
        outs.append(o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy())
    err = rel_l2(outs[0], outs[1])
    print(f"64x64 MIMO K=512 adjoint on {sel.size} voxels vs GPU FP64: rel-L2 {err:.2e}",
          flush=True)
    assert err <= TOL_MF
    # the FP64 adjoint against the CPU oracle on a few antenna rows
    rows = np.array([0, 1234, 4095])
    ref = orc.mf_backward(g, sc.tx[rows], sc.rx[rows], sc.voxels[sel], sc.K, norm=sc.A,
                          **sc.params())
    assert rel_l2(outs[1][rows], ref) <= 1e-9


def _volume_target(dev, n=128, side=0.20, center=(0.0, 0.0, 0.30)):
    """A smooth non-negative target volume over the matched-filter grid (the
    tracer and the prediction do not depend on it)."""
    _, axes = mf_scenes.voxel_grid(n, side, center)
    g = [torch.tensor(a, dtype=torch.float32) for a in axes]
    X, Y, Z = torch.meshgrid(*g, indexing="ij")
    vol = torch.exp(-((X ** 2 + Y ** 2 + (Z - 0.3) ** 2) / 0.05 ** 2)).to(dev)
    return vol.to(torch.complex64), tuple(axes)


def _subset_batch(b, n, **over):
    """The first n rays of a device RayBatch as a CPU RayBatch (fields
    replaced by `over`)."""
    f = lambda t: None if t is None else t[:n].detach().cpu()  # noqa: E731
    fields = dict(sdf=f(b.sdf), grad=f(b.grad), dists=f(b.dists), z=f(b.z), o=f(b.o), d=f(b.d),
                  refl=f(b.refl), offset=f(b.offset), n_valid=f(b.n_valid),
                  inv_s=b.inv_s.detach().cpu().reshape(()), power=b.power.detach().cpu().reshape(()))
    fields.update(over)
    return ts.RayBatch(radar=b.radar, anneal=b.anneal, **fields)


def _tracer_vs_oracle_on_subset(b, dev, n, label):
    """Fused tracer forward + backward on the whole device batch b, with the
    output gradient nonzero only on the first n rays; forward, the five
    gradients and exact zeros elsewhere against the FP64 oracle on those n."""
    from geraf.ops.RFSignalTracer.lensless_tracer import lensless_signal_tracer

    leaves = {k: getattr(b, k).detach().clone().float().requires_grad_(True)
              for k in ("sdf", "grad", "refl", "inv_s", "power")}
    o_r, o_i = lensless_signal_tracer(leaves["sdf"], leaves["grad"], b.dists, b.z, b.o, b.d,
                                      leaves["inv_s"], leaves["power"],
                                      reflectance=leaves["refl"], trans_offset=b.offset,
                                      n_valid=b.n_valid, cos_anneal_ratio=b.anneal,
                                      numADCSample=b.radar["numADCSample"],
                                      adcSampleRate=b.radar["adcSampleRate"],
                                      fc_start=b.radar["fc_start"], As_sci=b.radar["As_sci"])
    R, K = o_r.shape
    rng = np.random.default_rng(R)
    g = np.zeros((R, K), np.complex128)
    g[:n] = rng.standard_normal((n, K)) + 1j * rng.standard_normal((n, K))
    gt_r = torch.tensor(g.real, dtype=torch.float32, device=dev)
    gt_i = torch.tensor(g.imag, dtype=torch.float32, device=dev)
    (o_r * gt_r + o_i * gt_i).sum().backward()
    out = (o_r.detach().double() + 1j * o_i.detach().double()).cpu().numpy()[:n]
    sub = _subset_batch(b, n)
    ref, rg = _lensless_oracle(sub, g[:n])
    err = rel(out, ref)
    print(f"{label}: forward on {n} rays rel-L2 {err:.2e}", flush=True)
    assert err <= TOL_FWD
    for k in ("sdf", "grad", "refl"):
        got = leaves[k].grad.double().cpu()
        e = rel(got[:n].numpy().ravel(), rg[k].numpy().ravel())
        print(f"  grad {k}: rel-L2 {e:.2e}", flush=True)
        assert e <= TOL_GRAD, k
        assert not torch.any(got[n:]), k  # zero output gradient there
    for k in ("inv_s", "power"):
        e = abs(float(leaves[k].grad) - rg[k].item()) / rg[k + "_abs"].item()
        print(f"  grad {k}: err/sum|c| {e:.2e}", flush=True)
        assert e <= TOL_GRAD, k
    return o_r.detach(), o_i.detach()


def test_stage1_workload_tracer_and_prediction(cuda):
    from rf_workloads import stage_step as st

    vol, axes = _volume_target(cuda)
    step = st.StageStep(st.STAGE1, vol, axes, cuda, form="fused")
    assert (step.cfg.rays, step.cfg.samples, step.cfg.K) == (8192, 64, 64)
    nv = step.b.n_valid
    assert int(nv.min()) >= 16 and int(nv.max()) <= 64 and int(nv.min()) < int(nv.max())
    with torch.no_grad():
        p = step.parts()
    # the tracer inputs the networks produced, with the calibration offset
    b = ts.RayBatch(p["sdf"], p["grad"].float(), step.b.dists, step.b.z, step.b.o, step.b.d,
                    p["refl"], p["offset"], nv, p["inv_s"].reshape(()), p["power"].reshape(()),
                    step.b.radar, 1.0)
    assert torch.any(p["offset"] != 0)
    sim_r, sim_i = _tracer_vs_oracle_on_subset(b, cuda, 256, "stage-1 8192x64 K=64 tracer")
    # the op rerun on the same inputs gives the step's signals bit for bit
    assert torch.equal(sim_r, p["sim_r"]) and torch.equal(sim_i, p["sim_i"])
    # the prediction: |MF| of the per-ray signals (each ray its own antenna,
    # 1/rays normalisation) at 1,024 targets, on the CPU oracle
    rng = np.random.default_rng(4)
    sel = np.sort(rng.choice(step.tgt.shape[0], 1024, replace=False))
    o = step.b.o.cpu().numpy()
    sig = (p["sim_r"].double() + 1j * p["sim_i"].double()).cpu().numpy()
    r = step.radar
    ref = orc.mf_forward(sig.real, sig.imag, o, o, step.tgt[sel].cpu().numpy(),
                         f0=mf_scenes.MF_F0, slope=r["As_sci"], fs=r["adcSampleRate"],
                         c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)
    got = (p["mf_r"].double() + 1j * p["mf_i"].double()).cpu().numpy()[sel]
    err = rel_l2(got, ref)
    pred = p["pred"][sel, 0].double().cpu().numpy()
    perr = float(np.max(np.abs(pred - np.abs(ref))) / np.max(np.abs(ref)))
    print(f"stage-1 prediction on 1024 targets vs CPU oracle: rel-L2 {err:.2e}, "
          f"|MF| max err {perr:.2e} (matched filter {step.mf_algo}, "
          f"tiled fraction {step.plan.tiled_fraction:.3f})", flush=True)
    assert err <= TOL_MF and perr <= TOL_MF
    # ground truth: the volume sample divided by trans_power and sub_adc
    assert r["sub_adc"] == 8
    torch.testing.assert_close(p["gt"][:, 0], step.gt / 8, rtol=0, atol=0)


def test_stage2_size_tracer(cuda):
    b = ts.rays(32768, 256, 512, seed=0, with_offset=True).to(cuda)
    _tracer_vs_oracle_on_subset(b, cuda, 256, "32768x256 K=512 tracer")


def test_stage2_target_volume_central_64(cuda):
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import tiled_forward

    sc = mf_scenes.config2(0, device=str(cuda), n_side=256, n_ap=128, K=1024, slab_z=(0, 1))
    assert sc.A == 16384 and sc.K == 1024
    lin = np.linspace(-0.10, 0.10, 256)  # the axes of the 256^3 grid over 20 cm at z = 30 cm
    ax = [(lin + c).astype(np.float32) for c in (0.0, 0.0, 0.30)]
    # the central 64^3 block, and 256 voxels in it (the scatterer-nearest first)
    lo = 96
    gx, gy, gz = np.meshgrid(ax[0][lo:lo + 64], ax[1][lo:lo + 64], ax[2][lo:lo + 64],
                             indexing="ij")
    block = np.stack([gx.ravel(), gy.ravel(), gz.ravel()], 1).astype(np.float32)
    idx = []
    for d in range(3):
        step = float(ax[d][1] - ax[d][0])
        idx.append(np.rint((sc.scatterers[:, d] - ax[d][0]) / step).astype(np.int64) - lo)
    inside = np.all([(i >= 0) & (i < 64) for i in idx], axis=0)
    near = np.unique((idx[0][inside] * 64 + idx[1][inside]) * 64 + idx[2][inside])
    rng = np.random.default_rng(64)
    sel = np.concatenate([near, np.setdiff1d(rng.integers(0, 64 ** 3, 512), near)])[:256]
    assert sel.size == 256
    print(f"256^3 target: 256 voxels of the central 64^3 ({min(near.size, 256)} "
          "scatterer-nearest)", flush=True)
    arrays = (sc.echo, sc.tx, sc.rx, block[sel], sc.params())
    f64 = run_mf(arrays, cuda, "f64")
    mma = run_mf(arrays, cuda, "mma")
    err = rel_l2(mma, f64)
    print(f"  direct (matrix-core, 2 tiles) vs GPU FP64: rel-L2 {err:.2e}", flush=True)
    assert err <= TOL_MF
    t = lambda x: torch.as_tensor(np.ascontiguousarray(x), device=cuda)  # noqa: E731
    o_r, o_i, info = tiled_forward(t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx),
                                   t(block), sc.K, sc.fs, sc.As_sci, 1e-6)
    tiled = (o_r.double() + 1j * o_i.double()).cpu().numpy()[sel]
    terr = rel_l2(tiled, f64)
    print(f"  tiled form on the 64^3 block ({info['path']}, {info['terms']} terms) vs GPU "
          f"FP64: rel-L2 {terr:.2e}", flush=True)
    assert info["path"] == "tiled" and terr <= TOL_MF
    few = sel[:2]
    ref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, block[few], **sc.params())
    assert rel_l2(f64[:2], ref) <= 1e-9
