"""ctypes binding of libgeraf_rf.so, the gfx950 C-ABI library behind the ops.

The library is built in-tree (``make -C geraf-sens_amd/csrc``) into
``geraf/ops/_lib/libgeraf_rf.so``.  There is no fallback: if the library is
missing or fails to load, every op raises ``RuntimeError`` naming the path.

Each wrapper takes torch tensors, passes their device pointers, the reference
radar constants as a ``geraf_fmcw_params`` struct and the current HIP stream
of the tensors' device, and turns a non-zero status into ``RuntimeError``
carrying ``geraf_rf_last_error()``.
"""
from __future__ import annotations

import ctypes
import os
import threading
from typing import Optional

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libgeraf_rf.so")

# Reference constants (MatchedFilter_cuda_kernel.cu:14-16, SignalTracer_cuda_kernel.cu:14-15).
SPEED_OF_LIGHT_SCALED = 2.9979
MF_START_FREQUENCY_SCALED = 773.704
PATH_LENGTH_BIAS = 0.15


class FmcwParams(ctypes.Structure):
    _fields_ = [
        ("f0", ctypes.c_double),
        ("slope", ctypes.c_double),
        ("fs", ctypes.c_double),
        ("c", ctypes.c_double),
        ("path_bias", ctypes.c_double),
    ]


def make_params(f0: float, slope: float, fs: float, c: float = SPEED_OF_LIGHT_SCALED,
                path_bias: float = PATH_LENGTH_BIAS) -> FmcwParams:
    return FmcwParams(float(f0), float(slope), float(fs), float(c), float(path_bias))


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_PARAMS = ctypes.POINTER(FmcwParams)

# name -> argument types (all return int status).  Kept in the order of
# include/geraf_rf.h.
_SIGNATURES = {
    "geraf_mf_forward_f32": [_P] * 7 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_mf_forward_f32_mma": [_P] * 7 + [_I64] * 3 + [_PARAMS, _P, _I64, _P],
    "geraf_mf_forward_f32_mma_noalloc": [_P] * 7 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_mf_forward_plain_f32": [_P] * 7 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_mf_forward_f64": [_P] * 7 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_mf_backward_f32": [_P] * 7 + [_I64] * 4 + [_PARAMS, _P],
    "geraf_mf_backward_f64": [_P] * 7 + [_I64] * 4 + [_PARAMS, _P],
    "geraf_mf_fft_forward_f32": [_P] * 7 + [_I64] * 3 + [ctypes.c_double, _PARAMS, _P],
    "geraf_mf_fft_forward_f64": [_P] * 7 + [_I64] * 3 + [ctypes.c_double, _PARAMS, _P],
    "geraf_mf_fft_backward_f32": [_P] * 7 + [_I64] * 4 + [ctypes.c_double, _PARAMS, _P],
    "geraf_mf_fft_backward_f64": [_P] * 7 + [_I64] * 4 + [ctypes.c_double, _PARAMS, _P],
    "geraf_st_forward_f32": [_P] * 8 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_st_forward_f64": [_P] * 8 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_st_backward_f32": [_P] * 10 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_st_backward_f64": [_P] * 10 + [_I64] * 3 + [_PARAMS, _P],
    "geraf_bilinear_forward_f32": [_P] * 7 + [_I64] * 4 + [_P],
    "geraf_bilinear_forward_f64": [_P] * 7 + [_I64] * 4 + [_P],
    "geraf_bilinear_backward_f32": [_P] * 6 + [_I64] * 4 + [_P],
    "geraf_bilinear_backward_f64": [_P] * 6 + [_I64] * 4 + [_P],
    "geraf_lensless_forward_f32": [_P] * 15 + [_I64] * 3 + [ctypes.c_float, _PARAMS, _P],
    "geraf_lensless_backward_f32": [_P] * 19 + [_I64] * 3 + [ctypes.c_float, _PARAMS, _P],
    "geraf_softplus_bf16": [_P] * 2 + [_I64, ctypes.c_float, ctypes.c_float, _P],
    "geraf_softplus_grad_bf16": [_P] * 3 + [_I64, ctypes.c_float, ctypes.c_float, _P],
    "geraf_softplus_grad2_bf16": [_P] * 5 + [_I64, ctypes.c_float, ctypes.c_float, _P],
    "geraf_intersect_sphere_f32": [_P] * 4 + [_I64, ctypes.c_float, _P],
    "geraf_calibration_offset_f32": [_P] * 5 + [_I64] * 2 + [_P],
    "geraf_mlp_pack_weights": [_P, _I64, ctypes.c_int32, ctypes.c_float, _I64, _I64, _P, _P],
    "geraf_mlp_gemm_bf16": [_P, _I64, _P, _I64, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P, _I64,
                            _I64, _I64, _I64, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                            ctypes.c_float, _P],
    "geraf_mlp_wgrad_bf16": [_P, _I64, _P, _I64, ctypes.c_int32, _P, _I64, _P, _I64,
                             ctypes.c_int32, _P, _I64, _P, ctypes.c_int32, _I64, _I64, _I64,
                             ctypes.c_float, ctypes.c_float, _P, _I64, _P],
    "geraf_mlp_positional_encoding": [_P, _I64, ctypes.c_int32, ctypes.c_float, _P, _I64,
                                      ctypes.c_int32, ctypes.c_int32, _P],
    "geraf_mlp_pe_gradient": [_P, _I64, ctypes.c_int32, ctypes.c_float, _P, _I64, _P, _I64, _P,
                              _P],
    "geraf_mlp_pe_gradient_adjoint": [_P, _I64, ctypes.c_int32, ctypes.c_float, _P, _P, _I64,
                                      ctypes.c_int32, ctypes.c_int32, _P],
    "geraf_mlp_chain_adjoint_bf16": [_P, _I64, _P, _I64, _P, ctypes.c_int32, _P,
                                     ctypes.c_int32, ctypes.c_int32, _P, _I64, _I64, _P, _I64,
                                     _I64, _P, _I64, _I64, _P, _I64, _I64, _P],
    "geraf_mlp_sdf_chain_bf16": [_P, _I64, ctypes.c_int32, ctypes.c_float, _P, _P, _P,
                                 _P, _P, _I64, _I64, _P, _I64, _I64,
                                 ctypes.c_float, ctypes.c_float, _P],
    "geraf_mf_forward_tiled_f32": [_P] * 8 + [_I64] * 3 + [ctypes.c_int32] * 2 +
                                  [_P, _I64, _PARAMS, _P],
    "geraf_mf_backward_tiled_f32": [_P] * 8 + [_I64] * 4 + [ctypes.c_int32] * 2 +
                                   [_P, _I64, _PARAMS, _P],
}

_lock = threading.Lock()
_lib: Optional[ctypes.CDLL] = None


def load() -> ctypes.CDLL:
    """Load (once) and return the C-ABI library; raise RuntimeError if absent."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"geraf RF native library not found at {LIB_PATH}; build it with "
                "`make -C geraf-sens_amd/csrc` (hipcc --offload-arch=gfx950)."
            )
        # torch first: it brings libamdhip64.so.7, which the library then shares.
        lib = ctypes.CDLL(LIB_PATH)
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = ctypes.c_int
        lib.geraf_rf_last_error.restype = ctypes.c_char_p
        lib.geraf_rf_last_error.argtypes = []
        lib.geraf_rf_version.restype = ctypes.c_char_p
        lib.geraf_rf_version.argtypes = []
        lib.geraf_mf_tiled_workspace_bytes.restype = ctypes.c_int64
        lib.geraf_mf_tiled_workspace_bytes.argtypes = [_I64, _I64, _I64, ctypes.c_int32,
                                                       ctypes.c_int32]
        lib.geraf_mf_forward_mma_workspace_bytes.restype = ctypes.c_int64
        lib.geraf_mf_forward_mma_workspace_bytes.argtypes = [_I64, _I64]
        for name, args in (("geraf_mlp_pack_bytes", [_I64, _I64]),
                           ("geraf_mlp_wgrad_workspace_bytes", [_I64, _I64, _I64]),
                           ("geraf_mlp_sdf_chain_pack_bytes", [ctypes.c_int32])):
            getattr(lib, name).restype = ctypes.c_int64
            getattr(lib, name).argtypes = args
        _lib = lib
    return _lib


def exported_symbols():
    return list(_SIGNATURES) + ["geraf_rf_last_error", "geraf_rf_version",
                                "geraf_mf_tiled_workspace_bytes",
                                "geraf_mf_forward_mma_workspace_bytes", "geraf_mlp_pack_bytes",
                                "geraf_mlp_wgrad_workspace_bytes",
                                "geraf_mlp_sdf_chain_pack_bytes"]


def version() -> str:
    return load().geraf_rf_version().decode()


def stream_of(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def ptr(t: Optional[torch.Tensor]) -> ctypes.c_void_p:
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def call(name: str, *args) -> None:
    lib = load()
    status = getattr(lib, name)(*args)
    if status != 0:
        msg = lib.geraf_rf_last_error().decode(errors="replace")
        raise RuntimeError(f"{name} failed (status {status}): {msg}")
