"""One launch of each hot kernel on its benchmark workload, for rocprofv3
counter passes, plus a 1 GiB device copy whose FETCH_SIZE / WRITE_SIZE
readings calibrate those counters (known traffic: 1 GiB read, 1 GiB written).

    rocprofv3 --pmc <counters> -d <dir> -o pmc -- python3 geraf-sens_amd/rf_workloads/profile_driver.py

Kernels, in dispatch order: the copy; mf_forward_horner_kernel,
mf_mma_prep_kernel + mf_forward_mma_kernel (the matrix-core form) and
mf_backward_kernel on the 64x64 MIMO / K=512 / 128^3 workload;
lensless_forward_kernel and
lensless_backward_kernel on 4,096 rays x 128 samples, K=256, then (with
--stage2) once more on 32,768 rays x 256 samples, K=512.  --tracer-only skips
the copy and the matched filter.  --stage1-step N instead runs two warm-up
and N timed stage-1-style training steps (8,192 rays x 64 samples, K=64;
fused form, BF16 MLPs) for a kernel trace of the whole step; the target
volume is a smooth synthetic blob on the 32^3 grid, which fixes the work but
not the values.  --stage2-step N does the same for the stage-2-style step
(32,768 rays x 256 samples, K=512, consistency term; one warm-up step).
--tiled runs the per-tile expansion matched filter and its
adjoint on the 64x64 MIMO, K=512, 128^3 workload (twice: the second call is
the one to read).  --range-bin runs the range-bin matched filter and its adjoint (FP32 forms) on
the 64x64 MIMO aperture and the 128^3 voxels with 512 bins, three times.
--ap-ops runs the [A,P] signal tracer forward and backward
at its training shape (2,048 antenna pairs, 51,456 points, K=64) and the
depth resampler (804 rays x 32 -> 64 and 32,768 x 256 -> 512 samples), each
three times.  --sdf-network runs the hand-written SDF network (the RF
stages' 8 x 256 configuration, BF16) on 8,388,608 points: one warm-up and
two passes of sdf, features and normals with the full backward.
"""
import argparse
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
sys.path[:0] = [PKG]

import numpy as np  # noqa: E402
import torch  # noqa: E402

from geraf.ops import _native  # noqa: E402
from geraf.ops.RFSignalTracer import _C as stC  # noqa: E402
from rf_workloads import mf_scenes, tracer_scenes as ts  # noqa: E402


def run_tracer(dev, bb) -> None:
    R, S, Kt = bb.R, bb.S, bb.radar["numADCSample"]
    tp = stC.tracer_params(bb.radar["fc_start"], bb.radar["As_sci"], bb.radar["adcSampleRate"])
    P = _native.ptr
    inv_s = bb.inv_s.reshape(1).contiguous()
    power = bb.power.reshape(1).contiguous()
    ins = [P(bb.sdf), P(bb.grad), P(bb.refl), P(bb.dists), P(bb.z), P(bb.o), P(bb.d),
           P(bb.offset), P(bb.n_valid), P(inv_s), P(power)]
    lr = torch.empty((R, Kt), device=dev)
    li = torch.empty_like(lr)
    st = _native.stream_of(lr)
    carry = torch.empty((R, (S + 63) // 64), device=dev)
    _native.call("geraf_lensless_forward_f32", *ins, P(lr), P(li), P(None), P(carry), R, S, Kt,
                 bb.anneal, tp, st)
    g_s = torch.empty((R, S), device=dev)
    g_g = torch.empty((R, S, 3), device=dev)
    g_r = torch.empty((R, S), device=dev)
    g_i = torch.empty(R, device=dev)
    g_p = torch.empty(R, device=dev)
    _native.call("geraf_lensless_backward_f32", P(lr), P(li), *ins, P(g_s), P(g_g), P(g_r),
                 P(g_i), P(g_p), P(carry), R, S, Kt, bb.anneal, tp, st)
    torch.cuda.synchronize()


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--stage2", action="store_true")
    ap.add_argument("--tracer-only", action="store_true")
    ap.add_argument("--stage1-step", type=int, default=0)
    ap.add_argument("--stage2-step", type=int, default=0)
    ap.add_argument("--form", default="fused", choices=("fused", "fp32_autograd"))
    ap.add_argument("--tiled", action="store_true")
    ap.add_argument("--ap-ops", action="store_true")
    ap.add_argument("--sdf-network", action="store_true")
    ap.add_argument("--range-bin", action="store_true")
    args = ap.parse_args()
    dev = torch.device("cuda:0")
    if args.sdf_network:
        sdf_network(dev)
        return
    if args.ap_ops:
        ap_ops(dev)
        return
    if args.tiled:
        tiled(dev)
        return
    if args.range_bin:
        range_bin(dev)
        return
    if args.stage1_step:
        stage_steps(dev, args.stage1_step, args.form)
        return
    if args.stage2_step:
        stage_steps(dev, args.stage2_step, "fused", stage=2)
        return
    if args.tracer_only:
        trace_sizes(dev, args.stage2)
        return
    print("profile_driver: 1 GiB copy", flush=True)
    a = torch.empty(1 << 28, device=dev)
    b = torch.empty_like(a)
    a.fill_(1.0)
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b

    print("profile_driver: matched filter 64x64 MIMO, K=512, 128^3", flush=True)
    sc = mf_scenes.config2(0, device=str(dev))
    A, K, N = sc.A, sc.K, sc.N
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=dev)  # noqa: E731
    re, im, tx, rx, vox = t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx), t(sc.voxels.T)
    o_r = torch.empty(N, device=dev)
    o_i = torch.empty(N, device=dev)
    prm = _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias)
    _native.call("geraf_mf_forward_f32", *map(_native.ptr, (re, im, tx, rx, o_r, o_i, vox)),
                 A, K, N, prm, _native.stream_of(re))
    torch.cuda.synchronize()
    print("profile_driver: matched filter, matrix-core form, same workload", flush=True)
    ws_bytes = int(_native.load().geraf_mf_forward_mma_workspace_bytes(A, K))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    _native.call("geraf_mf_forward_f32_mma", *map(_native.ptr, (re, im, tx, rx, o_r, o_i, vox)),
                 A, K, N, prm, _native.ptr(ws), ws_bytes, _native.stream_of(re))
    torch.cuda.synchronize()
    del ws
    print("profile_driver: matched-filter adjoint, same workload", flush=True)
    a_r = torch.empty((A, K), device=dev)
    a_i = torch.empty_like(a_r)
    _native.call("geraf_mf_backward_f32", *map(_native.ptr, (o_r, o_i, tx, rx, a_r, a_i, vox)),
                 A, K, N, A, prm, _native.stream_of(re))
    torch.cuda.synchronize()

    trace_sizes(dev, args.stage2)


def sdf_network(dev, M: int = 8388608) -> None:
    """The hand-written SDF network's pass (fused_sdf: forward, normals and
    the backward with the double-backward terms) on M random points."""
    from geraf.models.networks.fused_sdf import fused_sdf
    from geraf.models.networks.sdf_network import SDFNetwork

    torch.manual_seed(0)
    net = SDFNetwork(3, 257, 256, 8, skip_in=(4,), multires=6,
                     compute_dtype=torch.bfloat16).to(dev)
    x = torch.rand(M, 3, device=dev) * 1.6 - 0.8
    g_sdf = torch.randn(M, device=dev)
    g_feat = (torch.randn(M, 256, device=dev) * 0.05).to(torch.bfloat16)
    g_n = torch.randn(M, 3, device=dev)
    for i in range(3):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record()
        sdf, feat, n = fused_sdf(net, x)
        torch.autograd.backward([sdf, feat, n], [g_sdf, g_feat, g_n])
        ev[1].record()
        torch.cuda.synchronize()
        print(f"profile_driver: SDF network pass {i} on {M} points: "
              f"{ev[0].elapsed_time(ev[1]):.1f} ms", flush=True)
        del sdf, feat, n
        net.zero_grad(set_to_none=True)


def tiled(dev) -> None:
    """The per-tile expansion matched filter and its adjoint on the 64x64
    MIMO, K=512, 128^3 workload (one call each, after one warm-up call)."""
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import tiled_adjoint, tiled_forward

    sc = mf_scenes.config2(0, device=str(dev))
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=dev)  # noqa: E731
    re, im, tx, rx, vox = t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx), t(sc.voxels)
    for rep in range(2):
        print(f"profile_driver: tiled matched filter, call {rep}", flush=True)
        o_r, o_i, info = tiled_forward(re, im, tx, rx, vox, sc.K, sc.fs, sc.As_sci, 1e-6)
        tiled_adjoint(o_r, o_i, tx, rx, info["plan"], sc.A, sc.K, sc.A, sc.fs, sc.As_sci, vox)
        torch.cuda.synchronize()
    print("profile_driver: done", flush=True)


def range_bin(dev) -> None:
    """The range-bin matched filter (FP32 forms) and its adjoint on the 64x64
    MIMO aperture and 128^3 voxels, 512 bins of 2.5 mm (bench.py's range_bin
    leg): one warm-up call each, then two."""
    sc = mf_scenes.config2(0, device=str(dev))
    t = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float32, device=dev)  # noqa
    A, N, R, spacing = sc.A, sc.N, 512, float(np.float32(0.0025))
    rng = np.random.default_rng(5)
    sr, si = t(rng.standard_normal((A, R))), t(rng.standard_normal((A, R)))
    gr, gi = t(rng.standard_normal(N)), t(rng.standard_normal(N))
    tx, rx, vox = t(sc.tx), t(sc.rx), t(sc.voxels.T)
    o_r, o_i = torch.empty(N, device=dev), torch.empty(N, device=dev)
    b_r, b_i = torch.empty((A, R), device=dev), torch.empty((A, R), device=dev)
    prm = _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias)
    P = _native.ptr
    st = _native.stream_of(o_r)
    for rep in range(3):
        print(f"profile_driver: range-bin matched filter, call {rep}", flush=True)
        _native.call("geraf_mf_fft_forward_f32", P(sr), P(si), P(tx), P(rx), P(o_r), P(o_i),
                     P(vox), A, R, N, spacing, prm, st)
        _native.call("geraf_mf_fft_backward_f32", P(gr), P(gi), P(tx), P(rx), P(b_r), P(b_i),
                     P(vox), A, R, N, A, spacing, prm, st)
        torch.cuda.synchronize()
    print("profile_driver: done", flush=True)


def stage_steps(dev, n: int, form: str, stage: int = 1) -> None:
    import time

    from rf_workloads import stage_step as st

    sc = mf_scenes.config1(0)
    g = [torch.tensor(a) for a in sc.axes]
    X, Y, Z = torch.meshgrid(*g, indexing="ij")
    vol = torch.exp(-((X ** 2 + Y ** 2 + (Z - 0.25) ** 2) / 0.03 ** 2)).to(dev)
    cfg = st.STAGE1 if stage == 1 else st.STAGE2
    step = st.StageStep(cfg, vol.to(torch.complex64), tuple(sc.axes), dev, form=form)
    for _ in range(2 if stage == 1 else 1):
        step.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        step.step()
    torch.cuda.synchronize()
    print(f"profile_driver: stage-{stage} step [{form}] "
          f"{(time.perf_counter() - t0) / n * 1e3:.1f} ms", flush=True)


def ap_ops(dev) -> None:
    c = ts.ap_training_case(dev)
    A, P, K, r = c["A"], c["P"], c["K"], c["radar"]
    print(f"profile_driver: [A,P] tracer A={A} P={P} K={K}", flush=True)
    prm = stC.tracer_params(r["fc_start"], r["As_sci"], r["adcSampleRate"])
    Pp = _native.ptr
    o_r = torch.empty((A, K), device=dev)
    o_i = torch.empty_like(o_r)
    gn = torch.empty((P, 3), device=dev)
    gs = torch.empty((A, P), device=dev)
    st = _native.stream_of(o_r)
    geo = [Pp(c[k]) for k in ("normals", "sigmas", "points", "tx", "rx", "inbounds")]
    for _ in range(3):
        _native.call("geraf_st_forward_f32", *geo, Pp(o_r), Pp(o_i), A, P, K, prm, st)
        _native.call("geraf_st_backward_f32", Pp(o_r), Pp(o_i), *geo, Pp(gn), Pp(gs), A, P, K,
                     prm, st)
    torch.cuda.synchronize()
    for R, S, T in ((804, 32, 64), (32768, 256, 512)):
        print(f"profile_driver: depth resampler {R} rays {S} -> {T}", flush=True)
        z = torch.sort(torch.rand(R, S, device=dev) * 0.3 + 0.2, dim=1).values
        tz = torch.sort(torch.rand(R, T, device=dev) * 0.32 + 0.19, dim=1).values
        nn_ = torch.randn(R, S, 3, device=dev)
        sg = torch.rand(1, R, S, device=dev)
        on = torch.empty((R, T, 3), device=dev)
        os_ = torch.empty((1, R, T), device=dev)
        ib = torch.empty((R, T), dtype=torch.bool, device=dev)
        g_n = torch.empty((R, S, 3), device=dev)
        g_s = torch.empty((1, R, S), device=dev)
        for _ in range(3):
            _native.call("geraf_bilinear_forward_f32", Pp(nn_), Pp(sg), Pp(z), Pp(tz), Pp(on),
                         Pp(os_), Pp(ib), 1, R, S, T, st)
            _native.call("geraf_bilinear_backward_f32", Pp(on), Pp(os_), Pp(z), Pp(tz), Pp(g_n),
                         Pp(g_s), 1, R, S, T, st)
        torch.cuda.synchronize()
    print("profile_driver: done", flush=True)


def trace_sizes(dev, stage2: bool) -> None:
    print("profile_driver: lensless tracer 4096 x 128, K=256", flush=True)
    run_tracer(dev, ts.config3(0).to(dev))
    if stage2:
        print("profile_driver: lensless tracer 32768 x 256, K=512", flush=True)
        run_tracer(dev, ts.rays(32768, 256, 512, seed=0).to(dev))
    print("profile_driver: done", flush=True)


if __name__ == "__main__":
    main()
