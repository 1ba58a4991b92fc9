#!/usr/bin/env python3
"""Benchmark of the GeRaF RF hot path on one MI355X.

A step is one full matched-filter backprojection of the 64x64 MIMO, K=512,
128^3-voxel workload (4.398e12 voxel*antenna*sample terms) by the path the
drop-in op takes for it, the matrix-core forward (FP32-accurate products
from an FP16 hi/lo split), inputs resident in HBM.  Prints ONE JSON line with:

* value        - terms/s over the timed steps (HIP events + synchronize);
* roofline     - algorithmic flops (8 per term: one complex multiply-add) per
                 launch / mean launch duration, against the 2.5 PFLOP/s FP16
                 matrix peak of the unit that runs the products, with the same
                 figure against the 157.3 TFLOP/s FP32 peak beside it;
* horner_form  - the vector-ALU (two-level Horner) kernel on the same volume;
* cpu_baseline - the reference formula restated in torch-CPU FP32 (the
                 reference has no CPU path), timed on a bounded voxel sample;
* errors       - volume rel-L2, peak-magnitude and phase error of the tuned
                 kernel against the plain FP64 GPU form on a voxel subset
                 (the FP64 form itself pinned to the CPU FP64 oracle);
* plain forms  - the plain FP32 (per-term sincosf) and plain FP64 GPU forms
                 timed on a voxel slab, and the tuned speed-up over them.
Further legs in the same line: the matched-filter adjoint, the tiled
expansion form, the fused lensless tracer (4,096 x 128, K=256 with its CPU
baseline, and 32,768 x 256, K=512), the stage-1-style step against its
all-FP32 autograd form, the 256^3 target volume (direct slab, whole volume by
tiles, FP64 check on the central 64^3) and the stage-2-style step.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "geraf-sens_amd"))
sys.path.insert(0, ROOT)

VALU_FP32_PEAK_TFLOPS = 157.3
MFMA_FP16_PEAK_TFLOPS = 2500.0
HBM_PEAK_GBS = 8000.0
FLOPS_PER_TERM = 8
MMA_FLOPS_PER_TERM = 24  # 4 real products per complex term, each as 3 FP16 products


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes (profiles/traffic.json, calibrated against a 1 GiB copy),
    or None if absent."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)[kernel]
        return t["fetch_bytes"] + t["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def bench_tiled(dev, sc, re, im, tx, rx, sel, ref, direct_ms):
    """The per-tile expansion form of the matched filter on the full 64x64
    MIMO, K=512, 128^3 volume, timed end to end (Morton ordering, tile
    bounds, the expansion kernels and the scatter back), with its errors
    against the FP64 GPU form on the bench's voxel subset."""
    import numpy as np
    import torch

    from geraf.ops.RFMatchedFilter.tiled_matched_filter import tiled_forward

    vox = torch.tensor(np.ascontiguousarray(sc.voxels), device=dev)
    run = lambda: tiled_forward(re, im, tx, rx, vox, sc.K, sc.fs, sc.As_sci, 1e-6)  # noqa: E731
    o_r, o_i, info = run()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 3
    e0.record()
    for _ in range(reps):
        o_r, o_i, info = run()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    terms = float(sc.A) * sc.K * sc.N
    out = torch.complex(o_r.double(), o_i.double()).cpu().numpy()[sel]
    rel = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    nsc = len(np.unique(mf_scenes_nearest(sc)))
    peak = float(np.max(np.abs(np.abs(out[:nsc]) / np.abs(ref[:nsc]) - 1.0)))
    phase = float(np.max(np.abs(np.angle(out[:nsc] / ref[:nsc]))))
    # the adjoint by the same expansion, against the direct adjoint kernel
    from geraf.ops import _native
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import tiled_adjoint
    plan = info["plan"]
    a_r, a_i = tiled_adjoint(o_r, o_i, tx, rx, plan, sc.A, sc.K, sc.A, sc.fs, sc.As_sci, vox)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        a_r, a_i = tiled_adjoint(o_r, o_i, tx, rx, plan, sc.A, sc.K, sc.A, sc.fs, sc.As_sci, vox)
    e1.record()
    torch.cuda.synchronize()
    adj_ms = e0.elapsed_time(e1) / reps
    d_r, d_i = torch.empty_like(a_r), torch.empty_like(a_i)
    vt = torch.tensor(np.ascontiguousarray(sc.voxels.T), device=dev)
    _native.call("geraf_mf_backward_f32", *map(_native.ptr, (o_r, o_i, tx, rx, d_r, d_i, vt)),
                 sc.A, sc.K, sc.N, sc.A,
                 _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias),
                 _native.stream_of(o_r))
    torch.cuda.synchronize()
    adj_rel = float((torch.complex(a_r, a_i) - torch.complex(d_r, d_i)).norm() /
                    torch.complex(d_r, d_i).norm())
    log(f"tiled adjoint: {adj_ms:.1f} ms, {terms / (adj_ms / 1e3):.3e} terms/s; rel-L2 vs the "
        f"direct adjoint {adj_rel:.2e}")
    log(f"tiled expansion: {ms:.1f} ms ({info['terms']} terms, Z={info['Z']:.3f}), "
        f"{terms / (ms / 1e3):.3e} terms/s, {direct_ms / ms:.1f}x the direct kernel; rel-L2 vs "
        f"FP64 {rel:.3e}, peak {peak:.2e}, phase {phase:.2e} rad")
    return {"metric": "MF voxel*antenna*freq terms/s (per-tile expansion form)",
            "value": terms / (ms / 1e3), "unit": "terms/s", "ms_per_call": ms,
            "speedup_vs_direct": direct_ms / ms, "terms": info["terms"], "Z": info["Z"],
            "adjoint": {"ms_per_call": adj_ms, "terms_per_s": terms / (adj_ms / 1e3),
                        "rel_l2_vs_direct_adjoint": adj_rel},
            "truncation_tol": 1e-6, "errors_vs_fp64": {"volume_rel_l2": rel,
                                                      "peak_mag_err_max": peak,
                                                      "phase_err_max_rad": phase},
            "workload": "matched filter, 64x64 MIMO, K=512, 128^3 voxels (Morton 8^3 tiles)"}


def plain_forms(dev, sc, re, im, tx, rx, vox, prm, stream, n_vox: int = 262144,
                n_ant: int = 64, reps: int = 3):
    """terms/s of the plain FP32 (per-term sincosf, the reference formula) and
    plain FP64 GPU forms on n_vox voxels x n_ant antennas of the workload: a
    launch that fills the part (n_vox / 256 blocks); one warm-up launch, then
    the median of `reps` launches, each between its own events."""
    import numpy as np
    import torch

    from geraf.ops import _native

    K = sc.K
    n_vox = min(n_vox, sc.N)
    n_ant = min(n_ant, sc.A)
    out = {"voxels": n_vox, "antennas": n_ant, "K": K, "launches": reps}
    for name, dt, key in (("geraf_mf_forward_plain_f32", torch.float32, "fp32_per_term_sincos"),
                          ("geraf_mf_forward_f64", torch.float64, "fp64")):
        ins = [x[:n_ant].to(dt).contiguous() for x in (re, im, tx, rx)]
        vs = vox[:, :n_vox].to(dt).contiguous()
        o_r = torch.empty(n_vox, device=dev, dtype=dt)
        o_i = torch.empty_like(o_r)
        ptrs = list(map(_native.ptr, ins + [o_r, o_i, vs]))
        call = lambda: _native.call(name, *ptrs, n_ant, K, n_vox, prm, stream)  # noqa: E731
        call()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            call()
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        out[key] = n_vox * float(n_ant) * K / (float(np.median(ms)) / 1e3)
        out[key + "_ms"] = float(np.median(ms))
    return out


def mf_scenes_nearest(sc):
    from rf_workloads import mf_scenes
    return mf_scenes.nearest_voxels(sc, sc.scatterers)


# FP64 operations per (voxel, antenna) pair of the range-bin forms that the
# roofline counts: two distances (3 subtractions, 1 multiply, 2 FMAs (4
# flops), 1 square root each = 9) and the carrier-phase reduction (2
# multiplies, 1 rounding, 1 subtraction = 4).
RANGE_BIN_FP64_FLOPS_PER_PAIR = 22
VALU_FP64_PEAK_TFLOPS = 78.6


def bench_range_bin(dev, sc, tx, rx, vox, prm, reps: int = 3):
    """Range-bin (FFT) matched filter and its adjoint on the 64x64 MIMO
    aperture and the 128^3 voxels, with range-compressed echoes of 512 bins
    (spacing 2.5 mm: the round trips of 0.55-1.07 m fall in bins ~110-215).
    FP32 forms timed per launch (median of `reps`); errors against the FP64
    forms on a voxel subset (forward) and on 64 antennas (adjoint)."""
    import numpy as np
    import torch

    from geraf.ops import _native

    A, N = sc.A, sc.N
    R, spacing = 512, float(np.float32(0.0025))
    rng = np.random.default_rng(5)
    s = rng.standard_normal((A, R)) + 1j * rng.standard_normal((A, R))
    g = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    f32 = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float32, device=dev)  # noqa
    f64 = lambda x: torch.tensor(np.ascontiguousarray(x), dtype=torch.float64, device=dev)  # noqa
    sr, si, gr, gi = f32(s.real), f32(s.imag), f32(g.real), f32(g.imag)
    o_r = torch.empty(N, device=dev)
    o_i = torch.empty_like(o_r)
    b_r = torch.empty((A, R), device=dev)
    b_i = torch.empty_like(b_r)
    P = _native.ptr
    st = _native.stream_of(o_r)
    fwd = lambda: _native.call("geraf_mf_fft_forward_f32", P(sr), P(si), P(tx), P(rx),  # noqa
                               P(o_r), P(o_i), P(vox), A, R, N, spacing, prm, st)
    bwd = lambda: _native.call("geraf_mf_fft_backward_f32", P(gr), P(gi), P(tx), P(rx),  # noqa
                               P(b_r), P(b_i), P(vox), A, R, N, A, spacing, prm, st)

    def per_launch(fn):
        fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(reps)]
        for e0, e1 in evs:
            e0.record()
            fn()
            e1.record()
        torch.cuda.synchronize()
        return float(np.median([e0.elapsed_time(e1) for e0, e1 in evs]))

    f_ms, b_ms = per_launch(fwd), per_launch(bwd)
    first = (b_r.clone(), b_i.clone())
    bwd()
    torch.cuda.synchronize()
    deterministic = bool(torch.equal(first[0], b_r) and torch.equal(first[1], b_i))
    pairs = float(A) * N
    res = {"workload": f"range-bin matched filter, 64x64 MIMO ({A} antennas), {N} voxels "
                       f"(128^3), {R} range bins of {spacing} m",
           "pairs": pairs, "range_bins": R,
           "kernels": {"forward": "mf_fft_forward_kernel<float>",
                       "backward": "mf_fft_backward_kernel<float, 4> (4 antennas per block at 512 bins)"},
           "adjoint_bitwise_repeatable": deterministic}
    for name, ms in (("forward", f_ms), ("backward", b_ms)):
        tf = RANGE_BIN_FP64_FLOPS_PER_PAIR * pairs / (ms / 1e3) / 1e12
        res[name] = {"ms": ms, "pairs_per_s": pairs / (ms / 1e3),
                     "roofline": {"bound": "valu (fp64)", "achieved_tflops": tf,
                                  "peak": VALU_FP64_PEAK_TFLOPS, "frac": tf / VALU_FP64_PEAK_TFLOPS,
                                  "flops_per_pair": RANGE_BIN_FP64_FLOPS_PER_PAIR}}
    # FP64 forms: the forward on 4,096 voxels, the adjoint on the first 64 antennas
    sel = np.sort(rng.choice(N, 4096, replace=False))
    vsel = vox[:, torch.tensor(sel, device=dev)].double().contiguous()
    r64 = torch.empty(len(sel), device=dev, dtype=torch.float64)
    i64 = torch.empty_like(r64)
    tx64, rx64 = tx.double().contiguous(), rx.double().contiguous()
    s64r, s64i = f64(s.real), f64(s.imag)
    _native.call("geraf_mf_fft_forward_f64", P(s64r), P(s64i), P(tx64), P(rx64), P(r64), P(i64),
                 P(vsel), A, R, len(sel), spacing, prm, _native.stream_of(r64))
    torch.cuda.synchronize()
    out = torch.complex(o_r.double(), o_i.double())[torch.tensor(sel, device=dev)]
    ref = torch.complex(r64, i64)
    res["forward"]["rel_l2_vs_fp64"] = float((out - ref).norm() / ref.norm())
    A2 = 64
    gb_r = torch.empty((A2, R), device=dev, dtype=torch.float64)
    gb_i = torch.empty_like(gb_r)
    vox64 = vox.double().contiguous()
    g64r, g64i = f64(g.real), f64(g.imag)  # kept alive across the call
    _native.call("geraf_mf_fft_backward_f64", P(g64r), P(g64i), P(tx64), P(rx64),
                 P(gb_r), P(gb_i), P(vox64), A2, R, N, A, spacing, prm, _native.stream_of(gb_r))
    torch.cuda.synchronize()
    outb = torch.complex(b_r[:A2].double(), b_i[:A2].double())
    refb = torch.complex(gb_r, gb_i)
    res["backward"]["rel_l2_vs_fp64"] = float((outb - refb).norm() / refb.norm())
    log(f"range-bin MF ({R} bins): forward {f_ms:.2f} ms, adjoint {b_ms:.2f} ms "
        f"({pairs / f_ms * 1e3:.3e} / {pairs / b_ms * 1e3:.3e} pairs/s); rel-L2 vs FP64 "
        f"{res['forward']['rel_l2_vs_fp64']:.2e} / {res['backward']['rel_l2_vs_fp64']:.2e}; "
        f"adjoint repeatable: {deterministic}")
    return res


def bench_adjoint(dev, sc, re, im, tx, rx, vox, g_r, g_i, prm, sel):
    """Matched-filter adjoint (grad w.r.t. the echo) on the full 64x64 MIMO,
    K=512, 128^3 workload, with the forward output as the incoming gradient;
    accuracy against the FP64 GPU adjoint with the gradient restricted to the
    voxel subset `sel` (zero elsewhere)."""
    import numpy as np
    import torch

    from geraf.ops import _native

    A, K, N = sc.A, sc.K, sc.N
    P = _native.ptr
    a_r = torch.empty((A, K), device=dev)
    a_i = torch.empty_like(a_r)
    st = _native.stream_of(a_r)

    def call(gr, gi, vx, n, outr, outi, name="geraf_mf_backward_f32"):
        _native.call(name, P(gr), P(gi), P(tx), P(rx), P(outr), P(outi), P(vx), A, K, n, A, prm,
                     st)

    call(g_r, g_i, vox, N, a_r, a_i)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    reps = 2
    e0.record()
    for _ in range(reps):
        call(g_r, g_i, vox, N, a_r, a_i)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    rate = float(A) * K * N / (ms / 1e3)
    # accuracy on the subset
    idx = torch.tensor(sel, device=dev)
    vs = vox[:, idx].contiguous()
    gr_s, gi_s = g_r[idx].contiguous(), g_i[idx].contiguous()
    n = int(idx.numel())
    call(gr_s, gi_s, vs, n, a_r, a_i)
    d = lambda t: t.double().contiguous()  # noqa: E731
    b_r = torch.empty((A, K), device=dev, dtype=torch.float64)
    b_i = torch.empty_like(b_r)
    # keep every FP64 operand referenced until the call has run
    tx64, rx64, gr64, gi64, vs64 = d(tx), d(rx), d(gr_s), d(gi_s), d(vs)
    _native.call("geraf_mf_backward_f64", P(gr64), P(gi64), P(tx64), P(rx64), P(b_r), P(b_i),
                 P(vs64), A, K, n, A, prm, st)
    torch.cuda.synchronize()
    got = torch.complex(a_r.double(), a_i.double())
    ref = torch.complex(b_r, b_i)
    rel = float((got - ref).norm() / ref.norm())
    log(f"adjoint: {ms:.1f} ms, {rate:.3e} terms/s; rel-L2 vs FP64 adjoint on {n} voxels "
        f"{rel:.2e}")
    return {"metric": "MF adjoint voxel*antenna*freq terms/s", "value": rate, "unit": "terms/s",
            "ms_per_launch": ms,
            "kernel": "mf_backward_mma16_kernel<2>" if 256 < K <= 512 else
                      "mf_backward_mma16_kernel" if 128 < K <= 1024 else "mf_backward_kernel",
            "rel_l2_vs_fp64": rel, "check_voxels": n,
            "workload": "matched-filter adjoint, 64x64 MIMO, K=512, 128^3 voxels"}


def bench_stage_step(dev, target_volume, axes, steps: int = 5, warmup: int = 3):
    """Stage-1-style training step (8,192 rays x 64 samples, valid counts
    16..64, K=64), fused form vs the all-FP32 autograd form.  Three warm-up
    steps: the third step of a fresh process carries a one-time ~70 ms (the
    first two do not reach it), the steps after it are steady."""
    import torch

    from rf_workloads import stage_step as st

    out = {"workload": "stage-1-style step: 8192 rays x 64 samples (valid 16..64), K=64, "
                       "SDF MLP 8x256 + signal MLP 4x256, transmittance calibration, "
                       "Charbonnier loss of |MF| prediction vs the 128^3 volume / sub_adc"}
    for form in ("fused", "fp32_autograd"):
        step = st.StageStep(st.STAGE1, target_volume, axes, dev, form=form)
        for _ in range(warmup):
            step.step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            loss = step.step()
        torch.cuda.synchronize()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        samples = int(step.valid.sum())
        out[form] = {"ms_per_step": ms, "valid_samples_per_step": samples,
                     "samples_per_s": samples / ms * 1e3, "last_loss": loss,
                     "matched_filter": step.mf_algo,
                     "mf_tiled_fraction": step.plan.tiled_fraction if step.plan else 0.0,
                     "mf_bucket_terms": [b.terms for b in step.plan.buckets] if step.plan
                     else None}
        log(f"stage step [{form}]: {ms:.2f} ms/step, {samples} valid samples, loss {loss:.4f}")
        del step
        torch.cuda.empty_cache()
    out["speedup_vs_fp32_autograd"] = out["fp32_autograd"]["ms_per_step"] / \
        out["fused"]["ms_per_step"]
    return out


def bench_large_target(dev):
    """The 256^3 target volume of the stage-2 workload (128x128 aperture,
    K=1024): one 256x256x4 slab (1/64 of the volume) timed; terms/s is
    independent of the slab size."""
    import numpy as np
    import torch

    from geraf.ops import _native
    from rf_workloads import mf_scenes

    sc = mf_scenes.config2(0, device=str(dev), n_side=256, n_ap=128, K=1024, slab_z=(0, 4))
    A, K, N = sc.A, sc.K, sc.N
    t = lambda x: torch.tensor(np.ascontiguousarray(x), device=dev)  # noqa: E731
    args = [t(sc.echo.real), t(sc.echo.imag), t(sc.tx), t(sc.rx)]
    o_r = torch.empty(N, device=dev)
    o_i = torch.empty_like(o_r)
    vox = t(sc.voxels.T)
    prm = _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias)
    ws_bytes = int(_native.load().geraf_mf_forward_mma_workspace_bytes(A, K))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    call = lambda: _native.call("geraf_mf_forward_f32_mma",  # noqa: E731
                                *map(_native.ptr, args + [o_r, o_i, vox]), A, K, N, prm,
                                _native.ptr(ws), ws_bytes, _native.stream_of(o_r))
    call()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    call()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    rate = A * K * N / (ms / 1e3)
    log(f"256^3 target slab ({N} voxels, A={A}, K={K}): {ms:.1f} ms, {rate:.3e} terms/s; "
        f"full volume {A * K * 256**3 / rate:.1f} s")
    res = {"workload": "MF target 256^3 (128x128 aperture, K=1024), 256x256x4 slab",
           "terms_per_s": rate, "slab_ms": ms, "full_volume_s": A * K * 256 ** 3 / rate}

    # the whole 256^3 volume by the per-tile expansion form, checked against
    # the direct kernel on the slab
    from geraf.ops.RFMatchedFilter.tiled_matched_filter import tiled_forward

    lin = np.linspace(-0.10, 0.10, 256)  # the grid of mf_scenes.voxel_grid(256, 0.20, centre)
    ax = [torch.tensor((lin + c).astype(np.float32), device=dev) for c in (0.0, 0.0, 0.30)]
    gx, gy, gz = torch.meshgrid(*ax, indexing="ij")
    full = torch.stack([gx.reshape(-1), gy.reshape(-1), gz.reshape(-1)], dim=1)
    del gx, gy, gz
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f_r, f_i, info = tiled_forward(args[0], args[1], args[2], args[3], full, K, sc.fs, sc.As_sci,
                                   1e-6)
    torch.cuda.synchronize()
    full_s = time.perf_counter() - t0
    slab_idx = torch.arange(256 * 256, device=dev) * 256  # iz = 0 plane of the full grid
    sl = torch.stack([slab_idx + z for z in range(4)], dim=1).reshape(-1)
    got = torch.complex(f_r[sl].double(), f_i[sl].double())
    ref = torch.complex(o_r.double(), o_i.double())
    rel = float((got - ref).norm() / ref.norm())
    terms_full = float(A) * K * 256 ** 3
    log(f"256^3 target, whole volume by tile expansion: {full_s:.2f} s ({info['terms']} terms, "
        f"Z={info['Z']:.3f}), {terms_full / full_s:.3e} terms/s; rel-L2 vs the direct kernel on "
        f"the slab {rel:.2e}")
    res["tiled_full_volume"] = {"seconds": full_s, "terms_per_s": terms_full / full_s,
                                "terms": info["terms"], "Z": info["Z"],
                                "rel_l2_vs_direct_on_slab": rel,
                                "speedup_vs_direct": (A * K * 256 ** 3 / rate) / full_s}

    # errors against the FP64 GPU form on 256 voxels of the central 64^3
    # subvolume (a full 64^3 subvolume is 4.4e15 FP64 terms at this aperture: a
    # sample)
    rng = np.random.default_rng(64)
    ijk = rng.integers(96, 160, (256, 3))
    idx = torch.tensor((ijk[:, 0] * 256 + ijk[:, 1]) * 256 + ijk[:, 2], device=dev)
    def d64(x):
        return torch.tensor(np.ascontiguousarray(x), device=dev, dtype=torch.float64)
    ops = [d64(sc.echo.real), d64(sc.echo.imag), d64(sc.tx), d64(sc.rx)]
    sub = full[idx].double().t().contiguous()
    s_r = torch.empty(idx.numel(), device=dev, dtype=torch.float64)
    s_i = torch.empty_like(s_r)
    t0 = time.perf_counter()
    _native.call("geraf_mf_forward_f64", *map(_native.ptr, ops + [s_r, s_i, sub]), A, K,
                 idx.numel(), prm, _native.stream_of(s_r))
    torch.cuda.synchronize()
    f64_s = time.perf_counter() - t0
    ref64 = torch.complex(s_r, s_i).cpu().numpy()
    got64 = torch.complex(f_r[idx].double(), f_i[idx].double()).cpu().numpy()
    err64 = float(np.linalg.norm(got64 - ref64) / np.linalg.norm(ref64))
    mag = float(np.max(np.abs(np.abs(got64) - np.abs(ref64))) / np.max(np.abs(ref64)))
    ph = float(np.max(np.abs(np.angle(got64 * np.conj(ref64)))))
    log(f"256^3 target vs FP64 on {idx.numel()} voxels of the central 64^3 ({f64_s:.1f} s of "
        f"FP64): rel-L2 {err64:.2e}, peak-magnitude {mag:.2e}, max phase {ph:.2e} rad")
    res["tiled_full_volume"]["vs_fp64_central_64c_sample"] = {
        "voxels": int(idx.numel()), "rel_l2": err64, "peak_magnitude_err": mag,
        "max_phase_err_rad": ph}
    vol = torch.complex(f_r, f_i).reshape(256, 256, 256)
    return res, vol, tuple(a.cpu() for a in ax)


def bench_stage2_step(dev, target_volume, axes, steps: int = 2, warmup: int = 1):
    """Stage-2-style step: 32,768 rays x 256 samples (8.4M samples), K=512,
    fused tracer, BF16 SDF + signal MLPs, the vision-SDF consistency term on
    262,144 points, matched filter of all rays onto all samples (tiled), L1
    against |MF| of the 256^3 target."""
    import torch

    from rf_workloads import stage_step as st

    torch.cuda.reset_peak_memory_stats()
    step = st.StageStep(st.STAGE2, target_volume, axes, dev, form="fused")
    for _ in range(warmup):
        step.step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        loss = step.step()
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    samples = int(step.valid.sum())
    peak = torch.cuda.max_memory_allocated() / 2 ** 30
    out = {"workload": "stage-2-style step: 32768 rays x 256 samples, K=512, hand-written BF16 "
                       "SDF MLP 8x256 (normals and double backward) + signal MLP 4x256, "
                       "transmittance calibration, consistency term on 262144 points, "
                       "Charbonnier loss of |MF| prediction vs the 256^3 target / sub_adc",
           "ms_per_step": ms, "samples_per_step": samples, "samples_per_s": samples / ms * 1e3,
           "last_loss": loss, "peak_memory_gib": peak,
           "mf_tiled_fraction": step.plan.tiled_fraction,
           "mf_bucket_terms": [b.terms for b in step.plan.buckets]}
    log(f"stage-2 step: {ms:.0f} ms/step, {samples} samples, peak {peak:.1f} GiB, "
        f"MF tiled fraction {step.plan.tiled_fraction:.3f}")
    del step
    torch.cuda.empty_cache()
    return out


def bench_sdf_network(dev, M: int = 8388608, reps: int = 3):
    """The SDF network of the RF stages (8 x 256 softplus layers, skip at 4,
    multires 6, BF16) at the stage-2 size: one pass of sdf, features and
    normals with the full backward (the double-backward terms included),
    hand-written (fused_sdf) vs torch's BF16 network under autograd; and each
    layer-product / weight-gradient form of the hand-written network on its
    own, timed per launch, with its algorithmic HBM bytes (the BF16 / FP16
    tensors it must read and write, [M,256] each) against the 8 TB/s peak and
    a device copy of the same size."""
    import torch

    from geraf.models.networks import fused_sdf as fs
    from geraf.models.networks.sdf_network import SDFNetwork

    BF, HF = torch.bfloat16, torch.float16
    torch.manual_seed(0)
    net = SDFNetwork(3, 257, 256, 8, skip_in=(4,), multires=6, compute_dtype=BF).to(dev)
    x = torch.rand(M, 3, device=dev) * 1.6 - 0.8
    g_sdf = torch.randn(M, device=dev)
    g_feat = (torch.randn(M, 256, device=dev) * 0.05).to(BF)
    g_n = torch.randn(M, 3, device=dev)

    def fused_pass():
        sdf, feat, n = fs.fused_sdf(net, x)
        torch.autograd.backward([sdf, feat, n], [g_sdf, g_feat, g_n])

    def torch_pass():
        out, n = net.sdf_and_gradient(x.clone(), float_out=False)
        torch.autograd.backward([out[:, 0], out[:, 1:], n],
                                [g_sdf.to(out.dtype), g_feat.to(out.dtype), g_n.to(n.dtype)])

    def timed(fn, n=reps):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(n):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        return sorted(ts)[len(ts) // 2]

    out = {"workload": f"SDF network pass (sdf, features, normals + backward) on {M} points",
           "points": M}
    out["fused_ms"] = timed(fused_pass)
    net.zero_grad(set_to_none=True)
    torch.cuda.empty_cache()
    out["torch_bf16_ms"] = timed(torch_pass)
    out["speedup_vs_torch_bf16"] = out["torch_bf16_ms"] / out["fused_ms"]
    net.zero_grad(set_to_none=True)
    torch.cuda.empty_cache()
    log(f"SDF network pass on {M} points: hand-written {out['fused_ms']:.1f} ms, torch BF16 "
        f"{out['torch_bf16_ms']:.1f} ms")

    # the reflectance network on the same points (features from the SDF
    # network's output rows, as the stage steps feed it), forward + backward
    from geraf.models.networks.fused_reflectance import fused_reflectance
    from geraf.models.networks.sdf_network import ReflectivePowerNetwork
    rnet = ReflectivePowerNetwork(256, compute_dtype=BF).to(dev)
    with torch.no_grad():
        feat_rows = fs.fused_sdf(net, x, with_normals=False)[1]
    g_r = torch.randn(M, device=dev)

    def refl_pass(fused):
        f = feat_rows.detach().requires_grad_(True)
        r = fused_reflectance(rnet, x, f) if fused else rnet.reflectance(x, f)
        torch.autograd.backward([r], [g_r.to(r.dtype)])

    out["reflectance_network"] = {
        "workload": f"reflectance network forward + backward on {M} points",
        "fused_ms": timed(lambda: refl_pass(True)),
        "torch_bf16_ms": timed(lambda: refl_pass(False))}
    rn = out["reflectance_network"]
    rn["speedup_vs_torch_bf16"] = rn["torch_bf16_ms"] / rn["fused_ms"]
    log(f"reflectance network on {M} points: hand-written {rn['fused_ms']:.1f} ms, torch BF16 "
        f"{rn['torch_bf16_ms']:.1f} ms")
    del rnet, feat_rows
    torch.cuda.empty_cache()

    # the kernel forms the pass runs, per launch, on [M,256] operands
    params = [t for l in range(9)
              for t in (getattr(net, f"lin{l}").weight, getattr(net, f"lin{l}").bias)]
    ws = [w.detach() for w in params[0::2]]
    hidden, _ = fs._hidden_pack(ws)
    rev, _ = fs._reverse_pack(ws)
    table = fs._bias_table([b.detach().float() for b in params[1::2]])
    w_sdf = fs._sdf_row(params)
    xa = (torch.randn(M, 256, device=dev) * 0.5).to(BF)
    qa = (torch.randn(M, 256, device=dev) * 0.3).to(HF)
    ea = torch.randn(M, 256, device=dev).to(BF)
    ya = torch.empty(M, 256, dtype=BF, device=dev)
    y2 = torch.empty(M, 256, dtype=BF, device=dev)
    hh = torch.empty(8, M, 256, dtype=BF, device=dev)
    hq = torch.empty(8, M, 256, dtype=HF, device=dev)
    sdf = torch.empty(M, device=dev)
    W = fs.pack(torch.randn(256, 256, device=dev) / 16, False)
    dw = torch.zeros(256, 256, device=dev)
    db = torch.zeros(256, device=dev)
    tensor_gb = M * 256 * 2 / 1e9

    def primal():
        fs._sweep(xa, rev, hq, hh[:7], hh[:7])

    forms = [
        ("device copy (torch)", 2, lambda: ya.copy_(xa)),
        ("chained forward, training (h_l and q_l of 8 layers written)", 16,
         lambda: fs._chain(x, hidden, table, h=hh, q=hq)),
        ("chained normal-pass sweep, 7 layers (q_l read, u_l s_l written)", 15,
         lambda: fs._sweep(xa[:1].expand(M, 256), rev, hq, None, hh[:7], x2=hq[7])),
        ("chained primal-adjoint sweep, 7 layers (q_l, t_l read, zbar_l written)", 22, primal),
        ("product + second-order term (vbar s, u s, q in; vbar s, t out)", 5,
         lambda: fs.gemm(xa, W, 256, 256, ya, y2=y2, e0=qa, e1=ea, epilogue=7)),
        ("product + hidden-layer adjoint", 4,
         lambda: fs.gemm(xa, W, 256, 256, ya, e0=qa, e1=ea, epilogue=5)),
        ("weight gradient", 2, lambda: fs.wgrad(xa, ea, 256, 256, dw, db=db)),
    ]
    kernels = []
    for name, tensors, fn in forms:
        ms = timed(fn, 5)
        gbs = tensors * tensor_gb / ms * 1e3
        kernels.append({"form": name, "ms": ms, "algorithmic_gb": tensors * tensor_gb,
                        "achieved_gbs": gbs, "hbm_frac": gbs / (HBM_PEAK_GBS)})
        log(f"  {name}: {ms:.3f} ms, {gbs / 1e3:.2f} TB/s")
    copy_gbs = kernels[0]["achieved_gbs"]
    for k in kernels[1:]:
        k["vs_device_copy"] = k["achieved_gbs"] / copy_gbs
    # the value-only chain moves 16 B per point: bound by its matrix and
    # vector work (0.918 MFLOP of products per point)
    ms = timed(lambda: fs._chain(x, hidden, table, w_sdf=w_sdf, sdf=sdf), 5)
    flops = 2.0 * (256 * 39 + 6 * 256 * 256 + 217 * 256 + 256) * M
    out["value_chain"] = {"form": "chained forward, sdf only", "ms": ms,
                          "tflops": flops / ms / 1e9, "mfma_frac": flops / ms / 1e9 / 2500.0}
    log(f"  chained forward, sdf only: {ms:.3f} ms, {flops / ms / 1e9:.0f} TFLOP/s of products")
    out["kernels"] = kernels
    del xa, qa, ea, ya, y2, hh, hq
    torch.cuda.empty_cache()
    return out


def bench_ap_tracer(dev, reps: int = 10):
    """The [A,P] signal-tracer op (SignalTracer_forward / _backward_normal_sigma)
    at the training shape of ap_training_case, each direction timed per launch
    (median of `reps`), with its FP32 and HBM roofline fractions; accuracy of
    one antenna row and 2,048 points against the FP64 oracle."""
    import numpy as np
    import torch

    from geraf.ops import _native
    from geraf.ops.RFSignalTracer import _C as stC

    from rf_workloads.tracer_scenes import ap_training_case

    c = ap_training_case(dev)
    A, P, K, r = c["A"], c["P"], c["K"], c["radar"]
    prm = stC.tracer_params(r["fc_start"], r["As_sci"], r["adcSampleRate"])
    o_r = torch.empty((A, K), device=dev)
    o_i = torch.empty_like(o_r)
    rng = np.random.default_rng(1)
    g_r = torch.tensor(rng.standard_normal((A, K)), dtype=torch.float32, device=dev)
    g_i = torch.tensor(rng.standard_normal((A, K)), dtype=torch.float32, device=dev)
    gn = torch.empty((P, 3), device=dev)
    gsg = torch.empty((A, P), device=dev)
    Pp = _native.ptr
    st = _native.stream_of(o_r)
    geo = [Pp(c["normals"]), Pp(c["sigmas"]), Pp(c["points"]), Pp(c["tx"]), Pp(c["rx"]),
           Pp(c["inbounds"])]
    fwd = lambda: _native.call("geraf_st_forward_f32", *geo, Pp(o_r), Pp(o_i), A, P, K,  # noqa
                               prm, st)
    bwd = lambda: _native.call("geraf_st_backward_f32", Pp(g_r), Pp(g_i), *geo, Pp(gn),  # noqa
                               Pp(gsg), A, P, K, prm, st)

    def per_launch(fn):
        fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(reps)]
        for e0, e1 in evs:
            e0.record()
            fn()
            e1.record()
        torch.cuda.synchronize()
        return float(np.median([e0.elapsed_time(e1) for e0, e1 in evs]))

    f_ms, b_ms = per_launch(fwd), per_launch(bwd)
    terms = float(A) * P * K
    # algorithmic work: 8 flops per (antenna, point, sample) each way; bytes:
    # forward sigmas + points/normals/inbounds + output, backward gradient row
    # + sigmas + points/normals/inbounds + sigma and normal gradients
    fb = A * P * 4 + P * 25 + A * K * 8
    bb = A * K * 8 + A * P * 8 + P * 25 + P * 12
    res = {"workload": f"[A,P] signal tracer, A={A} antenna pairs (one bank of a 64x64 MIMO "
                       f"aperture), P={P} points (804 rays x 64 target samples), K={K}",
           "terms": terms}
    for name, ms, by in (("forward", f_ms, fb), ("backward", b_ms, bb)):
        tf = 8.0 * terms / (ms / 1e3) / 1e12
        gbs = by / (ms / 1e3) / 1e9
        res[name] = {"ms": ms, "terms_per_s": terms / (ms / 1e3),
                     "roofline": {"bound": "valu", "achieved_tflops": tf,
                                  "fp32_frac": tf / VALU_FP32_PEAK_TFLOPS, "hbm_gbs": gbs,
                                  "hbm_frac": gbs / HBM_PEAK_GBS, "bytes": by}}
    res["kernels"] = {"forward": "st_forward_kernel", "backward": "st_backward_kernel"}
    # accuracy on one antenna row / a point subset against the FP64 oracle
    try:
        from oracle import rf as orc
        from rf_workloads import tracer_scenes as ts
        kw = dict(f0=r["fc_start"], slope=r["As_sci"], fs=r["adcSampleRate"],
                  c=ts.SPEED_OF_LIGHT, bias=ts.PATH_BIAS)
        cpu = {k: v.cpu().numpy() for k, v in c.items() if torch.is_tensor(v)}
        sel = slice(0, 2048)
        fwd()
        torch.cuda.synchronize()
        ref = orc.st_forward(cpu["normals"], cpu["sigmas"][:1], cpu["points"], cpu["tx"][:1],
                             cpu["rx"][:1], cpu["inbounds"], K, **kw)
        got = (o_r[:1].double() + 1j * o_i[:1].double()).cpu().numpy()
        res["fwd_rel_l2_vs_fp64_row0"] = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        bwd()
        torch.cuda.synchronize()
        g = (g_r.double() + 1j * g_i.double()).cpu().numpy()
        gn_ref, gs_ref = orc.st_backward(g, cpu["normals"][sel], cpu["sigmas"][:, sel],
                                         cpu["points"][sel], cpu["tx"], cpu["rx"],
                                         cpu["inbounds"][sel], **kw)
        # the op divides by its own P; the subset call divides by 2,048
        scale = 2048.0 / P
        res["bwd_normals_rel_l2_vs_fp64"] = float(
            np.linalg.norm(gn[sel].double().cpu().numpy() - gn_ref * scale) /
            np.linalg.norm(gn_ref * scale))
        res["bwd_sigmas_rel_l2_vs_fp64"] = float(
            np.linalg.norm(gsg[:, sel].double().cpu().numpy() - gs_ref * scale) /
            np.linalg.norm(gs_ref * scale))
    except Exception as e:  # the oracle is a checker; report, do not fail the bench
        log(f"[A,P] tracer oracle check skipped: {e}")
    log(f"[A,P] tracer A={A} P={P} K={K}: forward {f_ms:.3f} ms "
        f"({res['forward']['roofline']['fp32_frac']:.3f} of FP32 peak), backward {b_ms:.3f} ms "
        f"({res['backward']['roofline']['fp32_frac']:.3f}); errors "
        f"{ {k: v for k, v in res.items() if 'rel_l2' in k} }")
    return res


def bench_bilinear(dev, reps: int = 20):
    """The depth resampler (BilinearRaysample_forward / _backward) at the
    training shape (804 rays, 32 source and 64 target samples, one sigma row)
    and at 32,768 rays x 256 -> 512 samples, per-launch medians, with the HBM
    fraction of its algorithmic bytes."""
    import numpy as np
    import torch

    from geraf.ops import _native

    out = {}
    for R, S, T in ((804, 32, 64), (32768, 256, 512)):
        rng = np.random.default_rng(R)
        z = np.sort(rng.uniform(0.2, 0.5, (R, S)), axis=1).astype(np.float32)
        tz = np.sort(rng.uniform(0.19, 0.51, (R, T)), axis=1).astype(np.float32)
        nr = rng.standard_normal((R, S, 3)).astype(np.float32)
        sg = rng.uniform(0, 1, (1, R, S)).astype(np.float32)
        t = lambda x: torch.tensor(x, device=dev)  # noqa: E731
        zz, tt, nn, ss = t(z), t(tz), t(nr), t(sg)
        on = torch.empty((R, T, 3), device=dev)
        os_ = torch.empty((1, R, T), device=dev)
        ib = torch.empty((R, T), dtype=torch.bool, device=dev)
        gn = torch.empty((R, S, 3), device=dev)
        gs = torch.empty((1, R, S), device=dev)
        P = _native.ptr
        st = _native.stream_of(zz)
        fwd = lambda: _native.call("geraf_bilinear_forward_f32", P(nn), P(ss), P(zz), P(tt),  # noqa
                                   P(on), P(os_), P(ib), 1, R, S, T, st)
        bwd = lambda: _native.call("geraf_bilinear_backward_f32", P(on), P(os_), P(zz),  # noqa
                                   P(tt), P(gn), P(gs), 1, R, S, T, st)
        res = {}
        # algorithmic bytes: forward reads the sources (normals, sigma, depths)
        # and target depths, writes normals, sigma, inbounds; the backward
        # reads the target gradients and both depth rows, writes the source
        # gradients
        for name, fn, by in (("forward", fwd, R * S * 20 + R * T * 4 + R * T * 17),
                             ("backward", bwd, R * T * 16 + R * (S + T) * 4 + R * S * 16)):
            fn()
            torch.cuda.synchronize()
            evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                   for _ in range(reps)]
            for e0, e1 in evs:
                e0.record()
                fn()
                e1.record()
            torch.cuda.synchronize()
            ms = float(np.median([e0.elapsed_time(e1) for e0, e1 in evs]))
            gbs = by / (ms / 1e3) / 1e9
            res[name] = {"ms": ms, "bytes": by, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS}
        out[f"{R}x{S}to{T}"] = res
        log(f"bilinear {R} rays {S}->{T}: forward {res['forward']['ms'] * 1e3:.1f} us "
            f"({res['forward']['hbm_frac']:.3f} of HBM), backward "
            f"{res['backward']['ms'] * 1e3:.1f} us ({res['backward']['hbm_frac']:.3f})")
    out["workload"] = ("depth resampler: 804 rays x 32 -> 64 samples (training shape) and "
                       "32768 rays x 256 -> 512 samples, one sigma row")
    return out


def emitting_fraction(b) -> float:
    """Share of valid samples whose monostatic specular term is not culled
    (inc.n <= 0 and spec >= 1e-6; unit normals, inc = ray direction)."""
    import torch

    n = torch.nn.functional.normalize(b.grad, dim=-1)
    idn = (n * b.d[:, None, :]).sum(-1)
    spec = 2.0 * idn * idn - 1.0
    return float(((idn <= 0) & (spec >= 1e-6)).float().mean())


def samples_per_call(R: int, S: int) -> int:
    return R * S


def bench_tracer(dev, steps: int, warmup: int, cpu_seconds: float, no_cpu: bool,
                 rays: int = 4096, samples: int = 128, K: int = 256, oracle_check: bool = True):
    """Fused lensless tracer, forward + backward, on `rays` x `samples`, K
    frequencies (default: the 4,096 x 128, K=256 workload).  Returns a dict
    for the JSON line."""
    import numpy as np
    import torch

    from geraf.ops import _native
    from geraf.ops.RFSignalTracer import _C as stC
    from rf_workloads import tracer_scenes as ts

    b0 = ts.rays(rays, samples, K, seed=0)
    b = b0.to(dev)
    live = emitting_fraction(b)
    R, S, K = b.R, b.S, b.radar["numADCSample"]
    prm = stC.tracer_params(b.radar["fc_start"], b.radar["As_sci"], b.radar["adcSampleRate"])
    o_r = torch.empty((R, K), device=dev)
    o_i = torch.empty_like(o_r)
    rng = np.random.default_rng(0)
    g_r = torch.tensor(rng.standard_normal((R, K)), dtype=torch.float32, device=dev)
    g_i = torch.tensor(rng.standard_normal((R, K)), dtype=torch.float32, device=dev)
    gs = torch.empty((R, S), device=dev)
    gg = torch.empty((R, S, 3), device=dev)
    grf = torch.empty((R, S), device=dev)
    gis = torch.empty(R, device=dev)
    gpw = torch.empty(R, device=dev)
    inv_s = b.inv_s.reshape(1).contiguous()
    power = b.power.reshape(1).contiguous()
    ntile = (S + 63) // 64
    carry = torch.empty((R, ntile), device=dev)  # the forward's tile carries
    P = _native.ptr
    stream = _native.stream_of(o_r)
    ins = [P(b.sdf), P(b.grad), P(b.refl), P(b.dists), P(b.z), P(b.o), P(b.d), P(None), P(None),
           P(inv_s), P(power)]

    def fwd():
        _native.call("geraf_lensless_forward_f32", *ins, P(o_r), P(o_i), P(None), P(carry), R,
                     S, K, b.anneal, prm, stream)

    def bwd(tc=carry):
        _native.call("geraf_lensless_backward_f32", P(g_r), P(g_i), *ins, P(gs), P(gg), P(grf),
                     P(gis), P(gpw), P(tc), R, S, K, b.anneal, prm, stream)

    for _ in range(warmup):
        fwd()
        bwd()
    torch.cuda.synchronize()

    def burst(fn, n):
        """Mean time of n back-to-back launches between one pair of events on
        the launch stream (launch gaps included, event overhead amortised)."""
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    def per_launch(fn, n):
        """Median duration of n launches, each between its own pair of events on
        the launch stream (what a kernel trace reports per dispatch)."""
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(n)]
        for e0, e1 in evs:
            e0.record()
            fn()
            e1.record()
        torch.cuda.synchronize()
        return float(np.median([e0.elapsed_time(e1) for e0, e1 in evs]))

    n = max(steps, 4) if samples_per_call(R, S) > (1 << 22) else max(steps, 50)
    f_ms = per_launch(fwd, n)
    b_ms = per_launch(bwd, n)
    b2_ms = per_launch(lambda: bwd(None), n)  # without the carries: two walks
    pair_ms = f_ms + b_ms
    burst_pair_ms = burst(lambda: (fwd(), bwd()), n)
    samples = R * S
    rate = samples / (pair_ms / 1e3)
    # algorithmic traffic and work: the forward reads 28 B per sample (sdf,
    # 3-vector gradient, reflectance, dists, z), 24 B of origin and direction
    # per ray and writes the K complex outputs; the backward reads the same
    # inputs and the K complex gradients and writes 20 B per sample (sdf,
    # gradient and reflectance gradients) plus two per-ray partials; the
    # forward writes and the backward reads one carry per 64-sample tile
    bytes_fwd = samples * 28 + R * (24 + 8 * K + 4 * ntile) + 8
    bytes_bwd = samples * (28 + 20) + R * (24 + 8 * K + 8 + 4 * ntile) + 8
    flops = 16.0 * K * samples * live  # only emitting samples carry signal work
    gbs = (bytes_fwd + bytes_bwd) / (pair_ms / 1e3) / 1e9
    tfs = flops / (pair_ms / 1e3) / 1e12
    log(f"tracer {R}x{S} K={K}: fwd {f_ms:.4f} ms, bwd {b_ms:.4f} ms (per launch; "
        f"{b2_ms:.4f} ms without the forward's carries), pair "
        f"{pair_ms:.4f} ms (back to back {burst_pair_ms:.4f} ms), {rate:.3e} samples/s, "
        f"{gbs:.0f} GB/s, {tfs:.1f} TFLOP/s (emitting fraction {live:.3f})")

    # accuracy against the FP64 oracle on a ray subset: the output gradient is
    # nonzero only on the first n rays, so all five gradients (inv_s and
    # power included) are the subset's
    errs = {}
    try:
        if not oracle_check:
            raise RuntimeError("not requested for this size")
        from oracle import lensless as L
        n = 256
        sub = {k: getattr(b0, k)[:n].double().clone().requires_grad_(True)
               for k in ("sdf", "grad", "refl")}
        inv = b0.inv_s.double().reshape(1, 1).expand(n, S).clone().requires_grad_(True)
        pw = b0.power.double().reshape(1, 1).expand(n, S).clone().requires_grad_(True)
        r = b0.radar
        ref = L.lensless_forward(sub["sdf"], sub["grad"], b0.dists[:n].double(),
                                 L.points_f32(b0.o[:n], b0.z[:n], b0.d[:n]),
                                 b0.o[:n].double(), b0.d[:n].double(), inv, pw, K,
                                 r["adcSampleRate"], r["As_sci"], r["fc_start"],
                                 ts.SPEED_OF_LIGHT, ts.PATH_BIAS, refl=sub["refl"],
                                 anneal=b0.anneal)
        gt = torch.complex(g_r[:n].double().cpu(), g_i[:n].double().cpu())
        (ref.real * gt.real + ref.imag * gt.imag).sum().backward()
        m_r, m_i = torch.zeros_like(g_r), torch.zeros_like(g_i)
        m_r[:n], m_i[:n] = g_r[:n], g_i[:n]
        fwd()
        _native.call("geraf_lensless_backward_f32", P(m_r), P(m_i), *ins, P(gs), P(gg),
                     P(grf), P(gis), P(gpw), P(carry), R, S, K, b.anneal, prm, stream)
        torch.cuda.synchronize()
        out = torch.complex(o_r[:n].double(), o_i[:n].double()).cpu()
        errs["rays_checked"] = n
        errs["fwd_rel_l2"] = float((out - ref.detach()).norm() / ref.detach().norm())
        for name, mine, theirs in (("sdf", gs[:n], sub["sdf"].grad),
                                   ("gradients", gg[:n], sub["grad"].grad),
                                   ("reflectance", grf[:n], sub["refl"].grad)):
            m = mine.double().cpu()
            errs[f"grad_{name}_rel_l2"] = float((m - theirs).norm() / theirs.norm())
        # the two scalars: error relative to the sum of |per-sample terms|
        for name, mine, leaf in (("inv_s", gis, inv), ("power", gpw, pw)):
            errs[f"grad_{name}_err_over_sum_abs"] = float(
                abs(float(mine.double().sum()) - float(leaf.grad.sum())) /
                float(leaf.grad.abs().sum()))
        log(f"tracer errors vs FP64 oracle ({n} rays): {errs}")
    except Exception as e:  # a checker failure is reported, not fatal to the timing
        log(f"tracer oracle check skipped: {e}")

    cpu = None
    if not no_cpu:
        from oracle import cpu_baseline
        crate, threads, desc = cpu_baseline.time_tracer_sample(b0, cpu_seconds,
                                                               threads=len(os.sched_getaffinity(0)))
        cpu = {"value": crate, "unit": "samples/s", "cores": threads, "kind": "port",
               "sample": "torch-CPU FP32 restatement of the reference composition: " + desc}
        log(f"tracer cpu baseline {crate:.3e} samples/s")
    return {
        "metric": "tracer fwd+bwd samples/s",
        "value": rate, "unit": "samples/s",
        "workload": f"fused lensless tracer, {R} rays x {S} samples, K={K}",
        "fwd_ms": f_ms, "bwd_ms": b_ms, "bwd_two_walk_ms": b2_ms, "fwd_bwd_ms": pair_ms,
        "timing": "median per-launch kernel durations (events around each launch); "
                  "back-to-back pair for comparison",
        "back_to_back_pair_ms": burst_pair_ms, "emitting_fraction": live,
        "roofline": {"hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS, "fp32_tflops": tfs,
                     "fp32_frac": tfs / VALU_FP32_PEAK_TFLOPS,
                     "fp32_units": "MFMA (signal for every K, h_s for K > 128), VALU otherwise; the "
                                   "FP32 peak is 157.3 TFLOP/s on either",
                     "bytes_per_step": bytes_fwd + bytes_bwd, "flops_per_step": flops},
        "errors_vs_fp64": errs,
        "cpu_baseline": cpu,
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tracer", action="store_true")
    ap.add_argument("--no-stage", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); "
              "the RF hot path is benchmarked on a single MI355X.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from geraf.ops import _native
    from rf_workloads import mf_scenes

    dev = torch.device("cuda:0")
    log("building workload (config: 64x64 MIMO, K=512, 128^3)")
    sc = mf_scenes.config2(0, device=str(dev))
    A, K, N = sc.A, sc.K, sc.N
    re = torch.tensor(np.ascontiguousarray(sc.echo.real), device=dev)
    im = torch.tensor(np.ascontiguousarray(sc.echo.imag), device=dev)
    tx = torch.tensor(sc.tx, device=dev)
    rx = torch.tensor(sc.rx, device=dev)
    vox = torch.tensor(np.ascontiguousarray(sc.voxels.T), device=dev)
    o_r = torch.empty(N, device=dev)
    o_i = torch.empty(N, device=dev)
    prm = _native.make_params(sc.f0, sc.As_sci, sc.fs, sc.c, sc.bias)
    stream = _native.stream_of(re)
    ptrs = [_native.ptr(t) for t in (re, im, tx, rx, o_r, o_i, vox)]
    ws_bytes = int(_native.load().geraf_mf_forward_mma_workspace_bytes(A, K))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)

    def step():
        # the drop-in op's path for K = 512: the matrix-core forward
        _native.call("geraf_mf_forward_f32_mma", *ptrs, A, K, N, prm, _native.ptr(ws), ws_bytes,
                     stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"timing {args.steps} steps")
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        starts[i].record()
        step()
        ends[i].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    launch_ms = [s.elapsed_time(e) for s, e in zip(starts, ends)]
    ms_step = 1e3 * wall / args.steps
    terms = float(A) * K * N
    value = terms * args.steps / wall
    mean_launch_s = float(np.mean(launch_ms)) / 1e3
    achieved = FLOPS_PER_TERM * terms / mean_launch_s / 1e12
    log(f"tuned: {ms_step:.2f} ms/step, {value:.3e} terms/s, launch mean "
        f"{1e3 * mean_launch_s:.2f} ms (min {min(launch_ms):.2f}, max {max(launch_ms):.2f})")

    # the vector-ALU form (two-level Horner), for comparison on the same volume
    h_r, h_i = torch.empty_like(o_r), torch.empty_like(o_i)
    hptrs = [_native.ptr(t) for t in (re, im, tx, rx, h_r, h_i, vox)]
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    _native.call("geraf_mf_forward_f32", *hptrs, A, K, N, prm, stream)
    e1.record()
    torch.cuda.synchronize()
    horner_ms = e0.elapsed_time(e1)
    hdiff = float(torch.linalg.vector_norm(torch.complex(h_r - o_r, h_i - o_i)) /
                  torch.linalg.vector_norm(torch.complex(h_r, h_i)))
    log(f"Horner form: {horner_ms:.2f} ms, {terms / horner_ms * 1e3:.3e} terms/s; "
        f"rel-L2 between the two forms {hdiff:.2e}")

    # ---- accuracy against the FP64 forms on a voxel subset
    rng = np.random.default_rng(0)
    sel = np.unique(np.concatenate([mf_scenes.nearest_voxels(sc, sc.scatterers),
                                    rng.integers(0, N, 4096)]))
    out = torch.complex(o_r.double(), o_i.double()).cpu().numpy()[sel]
    d = lambda a: torch.tensor(np.ascontiguousarray(a), device=dev, dtype=torch.float64)  # noqa
    re64, im64, tx64, rx64 = d(sc.echo.real), d(sc.echo.imag), d(sc.tx), d(sc.rx)
    vsel64 = d(sc.voxels[sel].T)
    ns = len(sel)
    g_r = torch.empty(ns, device=dev, dtype=torch.float64)
    g_i = torch.empty_like(g_r)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    _native.call("geraf_mf_forward_f64", *map(_native.ptr, (re64, im64, tx64, rx64, g_r, g_i,
                                                          vsel64)), A, K, ns, prm,
                 _native.stream_of(re64))
    torch.cuda.synchronize()
    f64_s = time.perf_counter() - t1
    ref = torch.complex(g_r, g_i).cpu().numpy()
    rel = float(np.linalg.norm(out - ref) / np.linalg.norm(ref))
    nsc = len(np.unique(mf_scenes.nearest_voxels(sc, sc.scatterers)))
    peak = float(np.max(np.abs(np.abs(out[:nsc]) / np.abs(ref[:nsc]) - 1.0)))
    phase = float(np.max(np.abs(np.angle(out[:nsc] / ref[:nsc]))))
    log(f"rel-L2 vs FP64 on {ns} voxels: {rel:.3e}; peak {peak:.2e}; phase {phase:.2e} rad "
        f"(FP64 reference: {f64_s:.2f} s)")

    # the plain FP32 (reference formula, per-term sincosf) and FP64 forms on a
    # launch that fills the part: 262,144 voxels (1,024 blocks of 256) with the
    # first 64 antennas; terms/s does not depend on the sizes.  One warm-up
    # launch, then the median of three, each between its own events.
    plain = plain_forms(dev, sc, re, im, tx, rx, vox, prm, stream)
    p32_rate, f64_rate = plain["fp32_per_term_sincos"], plain["fp64"]
    log(f"plain FP32 {p32_rate:.3e} terms/s, plain FP64 {f64_rate:.3e} terms/s "
        f"({plain['voxels']} voxels x {plain['antennas']} antennas x K={K})")

    # the FP64 GPU form pinned to the CPU FP64 oracle on a few voxels
    pin_err = None
    try:
        from oracle import rf as orc
        few = sel[:6]
        oref = orc.mf_forward(sc.echo.real, sc.echo.imag, sc.tx, sc.rx, sc.voxels[few],
                              **sc.params())
        pin_err = float(np.linalg.norm(ref[:6] - oref) / np.linalg.norm(oref))
    except Exception as e:  # the oracle is a checker; report, do not fail the bench
        log(f"oracle pin skipped: {e}")

    adjoint = bench_adjoint(dev, sc, re, im, tx, rx, vox, o_r, o_i, prm, sel)
    range_bin = bench_range_bin(dev, sc, tx, rx, vox, prm)
    tiled = bench_tiled(dev, sc, re, im, tx, rx, sel, ref, 1e3 * mean_launch_s)

    stage = large = stage2 = None
    if not args.no_stage:
        axes = tuple(torch.tensor(a) for a in sc.axes)
        vol = torch.complex(o_r, o_i).reshape(sc.grid_shape)
        stage = bench_stage_step(dev, vol, axes)
        large, vol2, axes2 = bench_large_target(dev)
        stage2 = bench_stage2_step(dev, vol2, axes2)
        del vol2
        stage2["sdf_network"] = bench_sdf_network(dev)

    tracer = None
    if not args.no_tracer:
        tracer = bench_tracer(dev, max(args.steps, 5), max(args.warmup, 1), args.cpu_seconds / 2,
                              args.no_cpu)
        tracer["stage2_size"] = bench_tracer(dev, max(args.steps, 3), 1, 0.0, True, rays=32768,
                                             samples=256, K=512)
        # the same rays at the reference's training K (its stage configs' 64
        # samples per chirp), where the pair's bytes outweigh its flops
        tracer["training_k"] = bench_tracer(dev, max(args.steps, 3), 1, 0.0, True, rays=32768,
                                            samples=256, K=64)
        tracer["ap_op_training_shape"] = bench_ap_tracer(dev)
        tracer["bilinear_resampler"] = bench_bilinear(dev)

    # the CPU legs last: the launch-bound stage steps are timed before any
    # multi-threaded CPU work
    cpu = None
    if not args.no_cpu:
        from oracle import cpu_baseline
        ncores = len(os.sched_getaffinity(0))
        rate, threads, desc = cpu_baseline.time_mf_sample(sc, args.cpu_seconds, threads=ncores)
        cpu = {"value": rate, "unit": "terms/s", "cores": threads, "kind": "port",
               "sample": f"torch-CPU FP32 restatement of the reference formula "
                         f"(per-term cos/sin) on {desc}"}
        log(f"cpu baseline {rate:.3e} terms/s on {threads} threads")

    line = {
        "metric": "MF voxel*antenna*freq terms/s",
        "value": value,
        "unit": "terms/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-accurate products (fp16 hi/lo split on MFMA, fp32 accumulation; "
                 "fp64 distances and phase reduction)",
        "data": "synthetic (seeded box+torus scene, FP64-synthesised echo, 10 dB SNR)",
        "config": {"workload": "matched filter, 64x64 MIMO virtual array, K=512 (77.37-81.37 "
                               "GHz), 128^3 voxels over a 20 cm cube at 30 cm, 5000 scatterers",
                   "A": A, "K": K, "N": N, "terms": terms},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": MFMA_FP16_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / MFMA_FP16_PEAK_TFLOPS,
                     "traffic": profiled_traffic("mf_forward_mma_kernel<2>"),
                     "kernel": "mf_forward_mma_kernel<2> (with mf_mma_prep_kernel)",
                     "mean_launch_ms": 1e3 * mean_launch_s,
                     "flops_per_launch": FLOPS_PER_TERM * terms,
                     "fp32_peak_frac": achieved / VALU_FP32_PEAK_TFLOPS,
                     "mfma_flops_issued_per_term": MMA_FLOPS_PER_TERM,
                     "mfma_issued_frac": achieved * MMA_FLOPS_PER_TERM / FLOPS_PER_TERM /
                     MFMA_FP16_PEAK_TFLOPS,
                     "note": "achieved counts 8 flops per term (one complex multiply-add); "
                             "the products run on the FP16 matrix cores as a hi/lo split "
                             "(4 real products x 3 = 24 issued flops per term), so the FP32 "
                             "peak no longer bounds the kernel (fp32_peak_frac); its limiter "
                             "is the vector work that forms and splits the phasor operands"},
        "horner_form": {"kernel": "mf_forward_horner_kernel<4,2>", "ms": horner_ms,
                        "terms_per_s": terms / horner_ms * 1e3, "rel_l2_vs_mma": hdiff,
                        "fp32_peak_frac": FLOPS_PER_TERM * terms / horner_ms * 1e3 / 1e12 /
                        VALU_FP32_PEAK_TFLOPS,
                        "traffic": profiled_traffic("mf_forward_horner_kernel<4,2>")},
        "cpu_baseline": cpu,
        "errors_vs_fp64": {"volume_rel_l2": rel, "peak_mag_err_max": peak,
                           "phase_err_max_rad": phase, "voxels": ns,
                           "fp64_gpu_vs_cpu_oracle_rel_l2": pin_err},
        "plain_forms_terms_per_s": dict(plain, tuned_speedup_vs_fp64=value / f64_rate,
                                        tuned_speedup_vs_fp32=value / p32_rate),
        "adjoint": adjoint,
        "range_bin": range_bin,
        "tiled_expansion": tiled,
        "tracer": tracer,
        "stage1_step": stage,
        "stage2_target_volume": large,
        "stage2_step": stage2,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
