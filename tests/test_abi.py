"""C-ABI and host-logic checks that need no GPU: the library loads, exports
every function include/geraf_rf.h declares, reports errors through the status
code, and the drop-in ops validate their arguments like the reference."""
import ctypes
import os
import re

import pytest
import torch

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "geraf_rf.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(geraf_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from geraf.ops import _native

    lib = _native.load()
    names = header_functions()
    assert len(names) >= 10
    for name in names:
        assert hasattr(lib, name), name
    # the Python binding table covers the same set
    assert sorted(_native.exported_symbols()) == names


def test_version_and_error_reporting_without_gpu():
    from geraf.ops import _native

    lib = _native.load()
    assert b"gfx950" in lib.geraf_rf_version()
    # a NULL params struct is rejected on the host before any GPU call
    st = lib.geraf_mf_forward_f32(None, None, None, None, None, None, None, 1, 1, 1, None, None)
    assert st == 1
    assert b"params" in lib.geraf_rf_last_error()
    prm = _native.make_params(773.704, 701500.0, 1.25e6)
    st = lib.geraf_mf_forward_f32(None, None, None, None, None, None, None, -1, 1, 1,
                                  ctypes.byref(prm), None)
    assert st == 1 and b"non-negative" in lib.geraf_rf_last_error()
    bad = _native.make_params(773.704, 701500.0, 0.0)
    st = lib.geraf_st_forward_f32(*([None] * 8), 1, 1, 1, ctypes.byref(bad), None)
    assert st == 1 and b"adcSampleRate" in lib.geraf_rf_last_error()
    st = lib.geraf_bilinear_forward_f32(*([None] * 7), 1, 1, 1, 1, None)
    assert st == 1 and b"Nz >= 2" in lib.geraf_rf_last_error()
    # empty work is a no-op that touches no pointer
    st = lib.geraf_mf_forward_f32(None, None, None, None, None, None, None, 4, 8, 0,
                                  ctypes.byref(prm), None)
    assert st == 0
    # the matrix-core forward: its K limit, its workspace contract and empty work
    assert lib.geraf_mf_forward_mma_workspace_bytes(-1, 8) == -1
    need = lib.geraf_mf_forward_mma_workspace_bytes(3, 513)
    # 3 antennas x 3 tiles of 256 frequencies x 4 operands x 64 lanes x 16 B,
    # and the per-antenna scales
    assert need >= 3 * 3 * 4 * 64 * 16 + 3 * 4
    mma = lib.geraf_mf_forward_f32_mma
    st = mma(*([None] * 7), 2, 2049, 5, ctypes.byref(prm), None, 0, None)
    assert st == 1 and b"2048" in lib.geraf_rf_last_error()
    st = mma(*([ctypes.c_void_p(256)] * 7), 3, 513, 5, ctypes.byref(prm), None, need, None)
    assert st == 1 and b"NULL" in lib.geraf_rf_last_error()
    st = mma(*([ctypes.c_void_p(256)] * 7), 3, 513, 5, ctypes.byref(prm),
             ctypes.c_void_p(4096), need - 1, None)
    assert st == 1 and b"workspace too small" in lib.geraf_rf_last_error()
    st = mma(*([None] * 7), 3, 513, 0, ctypes.byref(prm), None, 0, None)
    assert st == 0


def test_call_raises_runtime_error_with_message():
    from geraf.ops import _native

    with pytest.raises(RuntimeError, match="non-negative"):
        _native.call("geraf_mf_backward_f32", *([None] * 7), -1, 1, 1, 1,
                     _native.make_params(773.704, 1.0, 1.0), None)


def test_drop_in_ops_reject_host_tensors_like_the_reference():
    from geraf.ops.RFMatchedFilter import matched_filter_cuda
    from geraf.ops.RFMatchedFilter import _C as mfC
    from geraf.ops.RFSignalTracer import signal_tracer_cuda
    from geraf.ops.RFSignalTracer import _C as stC

    x = torch.zeros(4, 8)
    p = torch.zeros(4, 3)
    with pytest.raises(AssertionError, match="CUDA"):
        matched_filter_cuda(x, x, p, p, torch.zeros(5, 3), 4, 8, 1.0, 1.0, 5)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        mfC.MatchedFilter_forward(x, x, p, p, torch.zeros(5), torch.zeros(5), torch.zeros(3, 5),
                                  4, 8, 1.0, 1.0, 5)
    with pytest.raises(ValueError, match="CUDA"):
        signal_tracer_cuda(p, x, p, p, p, torch.ones(4, dtype=torch.bool), 8, 1.0, 1.0, 1.0, 4, 4)
    with pytest.raises(RuntimeError, match="must be a CUDA tensor"):
        stC.BilinearRaysample_forward(torch.zeros(2, 3, 3), torch.zeros(1, 2, 3),
                                      torch.zeros(2, 3), torch.zeros(2, 4),
                                      torch.zeros(2, 4, 3), torch.zeros(1, 2, 4),
                                      torch.zeros(2, 4, dtype=torch.bool), 1, 2, 3, 4)


def test_tracer_rejects_non_differentiable_mode():
    from geraf.ops.RFSignalTracer.signal_tracer_cuda import _SignalTracerFunction

    with pytest.raises(ValueError, match="differentiable"):
        _SignalTracerFunction.forward(None, *([None] * 6), 8, 1.0, 1.0, 1.0, 1, 1, False, True)


def test_scalar_parameters_are_rounded_like_the_reference_binding():
    from geraf.ops.RFMatchedFilter import _C as mfC
    import numpy as np

    prm = mfC._params(1234567.891, 3149606.299212598)
    assert prm.fs == float(np.float32(1234567.891))
    assert prm.slope == float(np.float32(3149606.299212598))
    assert prm.f0 == float(np.float32(773.704)) and prm.c == float(np.float32(2.9979))
    assert prm.path_bias == float(np.float32(0.15))


def test_model_utils_reject_other_backends():
    from geraf.models.model_utils.matched_filter import base_matched_filter
    from geraf.models.model_utils.ray_sampler import bilinear_ray_sampler
    from geraf.models.model_utils.ray_tracer import lensless_ray_tracer

    t = torch.zeros(2, 3)
    with pytest.raises(ValueError, match="Unsupported"):
        base_matched_filter(t, t, t, t, t, {}, backend="torch")
    with pytest.raises(ValueError, match="Unsupported"):
        lensless_ray_tracer(t, t, t, t, t, {}, backend="torch")
    with pytest.raises(ValueError, match="Unsupported"):
        bilinear_ray_sampler(torch.zeros(1, 2, 3), torch.zeros(1, 1, 2), torch.zeros(1, 2),
                             torch.zeros(1, 2), backend="torch")
