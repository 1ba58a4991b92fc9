"""The matrix-core matched-filter forward without a workspace
(geraf_mf_forward_f32_mma_noalloc, static device scratch) against the same
operator with a caller-owned workspace (geraf_mf_forward_f32_mma), through the
C ABI: two calls in a row, the second with a smaller input, each bit-identical
to the workspace form; sizes beyond the scratch are rejected on the host."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import rf as orc
from rf_workloads import mf_scenes

pytestmark = pytest.mark.gpu


def _case(A, K, N, seed):
    rng = np.random.default_rng(seed)
    echo = (rng.standard_normal((A, K)) + 1j * rng.standard_normal((A, K))).astype(np.complex64)
    tx = rng.uniform(-0.06, 0.06, (A, 3)).astype(np.float32)
    tx[:, 2] = 0.0
    rx = (tx + np.float32([0.002, 0.0, 0.0])).astype(np.float32)
    vox = rng.uniform(-0.1, 0.1, (N, 3)).astype(np.float32)
    vox[:, 2] += 0.3
    prm = dict(f0=mf_scenes.MF_F0, slope=float(np.float32(701500.0)), fs=1.25e6 * K / 64,
               c=mf_scenes.SPEED_OF_LIGHT, bias=mf_scenes.PATH_BIAS)
    return echo, tx, rx, vox, prm


def _both(dev, echo, tx, rx, vox, prm):
    from geraf.ops import _native

    T = lambda x: torch.tensor(np.ascontiguousarray(x), device=dev)  # noqa: E731
    re, im, t, r, v = T(echo.real), T(echo.imag), T(tx), T(rx), T(np.asarray(vox).T)
    A, K = re.shape
    N = v.shape[1]
    params = _native.make_params(prm["f0"], prm["slope"], prm["fs"], prm["c"], prm["bias"])
    outs = []
    for name in ("geraf_mf_forward_f32_mma", "geraf_mf_forward_f32_mma_noalloc"):
        o_r = torch.full((N,), float("nan"), device=dev)
        o_i = torch.full((N,), float("nan"), device=dev)
        ptrs = list(map(_native.ptr, (re, im, t, r, o_r, o_i, v)))
        if name.endswith("noalloc"):
            _native.call(name, *ptrs, A, K, N, params, _native.stream_of(re))
        else:
            size = int(_native.load().geraf_mf_forward_mma_workspace_bytes(A, K))
            ws = torch.empty(size, dtype=torch.uint8, device=dev)
            _native.call(name, *ptrs, A, K, N, params, _native.ptr(ws), size,
                         _native.stream_of(re))
        torch.cuda.synchronize()
        outs.append(o_r.double().cpu().numpy() + 1j * o_i.double().cpu().numpy())
    return outs


def test_noalloc_matches_workspace_form_twice(cuda):
    # a large case (4,096 antennas, K = 512: 32 MiB of operands in the
    # 48 MiB scratch), then a smaller one that reads only part of what the
    # first left
    for A, K, N, seed in ((4096, 512, 3000, 1), (33, 300, 700, 2)):
        print(f"no-alloc MMA forward A={A} K={K} N={N}", flush=True)
        echo, tx, rx, vox, prm = _case(A, K, N, seed)
        ws_out, na_out = _both(cuda, echo, tx, rx, vox, prm)
        assert np.all(np.isfinite(na_out))
        assert np.array_equal(na_out, ws_out), float(np.max(np.abs(na_out - ws_out)))
        few = np.arange(0, N, max(1, N // 6))[:6]
        ref = orc.mf_forward(echo.real, echo.imag, tx, rx, vox[few], **prm)
        err = float(np.linalg.norm(na_out[few] - ref) / np.linalg.norm(ref))
        print(f"  bit-identical to the workspace form; rel-L2 vs FP64 oracle on {len(few)} "
              f"voxels {err:.2e}", flush=True)
        assert err <= 1e-5


def test_noalloc_rejects_sizes_beyond_the_scratch():
    from geraf.ops import _native

    lib = _native.load()
    prm = _native.make_params(mf_scenes.MF_F0, 701500.0, 1.25e6)
    fn = lib.geraf_mf_forward_f32_mma_noalloc
    # 6,142 antennas at K = 512 need more than 48 MiB: refused before any GPU use
    st = fn(*([None] * 7), 6142, 512, 10, ctypes.byref(prm), None)
    assert st == 1 and b"static scratch" in lib.geraf_rf_last_error()
    st = fn(*([None] * 7), 2458, 1025, 10, ctypes.byref(prm), None)
    assert st == 1 and b"static scratch" in lib.geraf_rf_last_error()
    st = fn(*([None] * 7), 2, 2049, 10, ctypes.byref(prm), None)
    assert st == 1 and b"2048" in lib.geraf_rf_last_error()
    # the limits themselves are accepted (empty volume: nothing is launched)
    assert fn(*([None] * 7), 6141, 512, 0, ctypes.byref(prm), None) == 0
    assert fn(*([None] * 7), 2457, 1025, 0, ctypes.byref(prm), None) == 0
