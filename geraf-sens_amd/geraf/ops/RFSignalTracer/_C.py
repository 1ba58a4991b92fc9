"""Drop-in for the reference's pybind module ``geraf.ops.RFSignalTracer._C``.

Same names, argument order and in-place output contract as
RFSignalTracer/csrc/vision.cpp:70-97 and the checks of
SignalTracer_cuda.cu:10-178 (failures raise ``RuntimeError``).  float32 runs
the tuned gfx950 kernels, float64 the FP64 kernels.  Scalars are rounded to
FP32 on entry like the reference's ``float`` parameters; the speed of light
and the 0.15 m path bias are the reference's FP32 constants
(SignalTracer_cuda_kernel.cu:14-15).

Deviation: ``SignalTracer_backward_normal_sigma`` writes zeros for points with
``inbounds == False`` (the reference leaves them uninitialised,
SignalTracer_cuda_kernel.cu:471-473), and the bilinear backward is a
deterministic gather (the reference scatters with float atomics).
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _native

_F32 = lambda x: float(np.float32(x))  # noqa: E731

SPEED_OF_LIGHT = _F32(_native.SPEED_OF_LIGHT_SCALED)
PATH_BIAS = _F32(_native.PATH_LENGTH_BIAS)


def _check(cond: bool, msg: str) -> None:
    if not cond:
        raise RuntimeError(msg)


def _cuda(t: torch.Tensor, name: str) -> None:
    _check(t.is_cuda, f"`{name}` must be a CUDA tensor.")
    _check(t.is_contiguous(), f"`{name}` must be contiguous.")


def _float_dtype(tensors, what: str) -> torch.dtype:
    dtype = tensors[0].dtype
    _check(dtype in (torch.float32, torch.float64), f"{what} supports float32/float64, got {dtype}")
    for t in tensors[1:]:
        _check(t.dtype == dtype, f"expected scalar type {dtype} but found {t.dtype}")
    return dtype


def tracer_params(fc_start: float, As_sci: float, adc_sample_rate: float) -> _native.FmcwParams:
    return _native.make_params(_F32(fc_start), _F32(As_sci), _F32(adc_sample_rate),
                               SPEED_OF_LIGHT, PATH_BIAS)


def _sfx(dtype: torch.dtype) -> str:
    return "f32" if dtype == torch.float32 else "f64"


def SignalTracer_forward(normals, sigmas, points, tx_pos, rx_pos, inbounds, output_real,
                         output_imag, num_adc_samples: int, adc_sample_rate: float,
                         start_frequency: float, chirp_slope_scaled: float, total_ant_size: int,
                         total_point_size: int) -> None:
    """In-place [A,P] tracer forward; reference: SignalTracer_cuda.cu:101-142."""
    names = ("normals", "sigmas", "points", "tx_pos", "rx_pos", "inbounds", "output_real",
             "output_imag")
    ts = (normals, sigmas, points, tx_pos, rx_pos, inbounds, output_real, output_imag)
    for t, n in zip(ts, names):
        _cuda(t, n)
    A, P, K = int(total_ant_size), int(total_point_size), int(num_adc_samples)
    _check(inbounds.dtype == torch.bool, "`inbounds` must be bool.")
    _check(normals.dim() == 2 and normals.size(1) == 3, "`normals` must have shape [P, 3].")
    _check(points.dim() == 2 and points.size(1) == 3, "`points` must have shape [P, 3].")
    _check(sigmas.dim() == 2, "`sigmas` must have shape [A, P].")
    _check(tx_pos.dim() == 2 and tx_pos.size(1) == 3, "`tx_pos` must have shape [A, 3].")
    _check(rx_pos.dim() == 2 and rx_pos.size(1) == 3, "`rx_pos` must have shape [A, 3].")
    _check(inbounds.dim() == 1, "`inbounds` must have shape [P].")
    _check(output_real.dim() == 2, "`output_real` must have shape [A, S].")
    _check(output_imag.dim() == 2, "`output_imag` must have shape [A, S].")
    _check(normals.size(0) == P, "`normals` point dimension mismatch.")
    _check(points.size(0) == P, "`points` point dimension mismatch.")
    _check(sigmas.size(0) == A and sigmas.size(1) == P, "`sigmas` shape mismatch.")
    _check(tx_pos.size(0) == A, "`tx_pos` antenna dimension mismatch.")
    _check(rx_pos.size(0) == A, "`rx_pos` antenna dimension mismatch.")
    _check(inbounds.size(0) == P, "`inbounds` point dimension mismatch.")
    _check(tuple(output_real.shape) == (A, K), "`output_real` shape mismatch.")
    _check(tuple(output_imag.shape) == (A, K), "`output_imag` shape mismatch.")
    dtype = _float_dtype((normals, sigmas, points, tx_pos, rx_pos, output_real, output_imag),
                         "signal tracer")
    prm = tracer_params(start_frequency, chirp_slope_scaled, adc_sample_rate)
    with torch.cuda.device(normals.device):
        _native.call(f"geraf_st_forward_{_sfx(dtype)}",
                     *map(_native.ptr, (normals, sigmas, points, tx_pos, rx_pos, inbounds,
                                        output_real, output_imag)),
                     A, P, K, prm, _native.stream_of(normals))


def SignalTracer_backward_normal_sigma(grad_output_real, grad_output_imag, normals, sigmas,
                                       points, tx_pos, rx_pos, inbounds, grad_normals,
                                       grad_sigmas, num_adc_samples: int,
                                       adc_sample_rate: float, start_frequency: float,
                                       chirp_slope_scaled: float, total_ant_size: int,
                                       total_point_size: int) -> None:
    """In-place [A,P] tracer backward; reference: SignalTracer_cuda.cu:144-178."""
    names = ("grad_output_real", "grad_output_imag", "normals", "sigmas", "points", "tx_pos",
             "rx_pos", "inbounds", "grad_normals", "grad_sigmas")
    ts = (grad_output_real, grad_output_imag, normals, sigmas, points, tx_pos, rx_pos, inbounds,
          grad_normals, grad_sigmas)
    for t, n in zip(ts, names):
        _cuda(t, n)
    A, P, K = int(total_ant_size), int(total_point_size), int(num_adc_samples)
    _check(inbounds.dtype == torch.bool, "`inbounds` must be bool.")
    _check(normals.dim() == 2 and normals.size(0) == P and normals.size(1) == 3,
           "`normals` must have shape [P, 3].")
    _check(points.dim() == 2 and points.size(0) == P and points.size(1) == 3,
           "`points` must have shape [P, 3].")
    _check(sigmas.dim() == 2 and tuple(sigmas.shape) == (A, P), "`sigmas` shape mismatch.")
    _check(tuple(tx_pos.shape) == (A, 3), "`tx_pos` must have shape [A, 3].")
    _check(tuple(rx_pos.shape) == (A, 3), "`rx_pos` must have shape [A, 3].")
    _check(inbounds.dim() == 1 and inbounds.size(0) == P, "`inbounds` must have shape [P].")
    _check(tuple(grad_output_real.shape) == (A, K), "`grad_output_real` shape mismatch.")
    _check(tuple(grad_output_imag.shape) == (A, K), "`grad_output_imag` shape mismatch.")
    _check(grad_normals.dim() == 2 and grad_normals.size(1) == 3,
           "`grad_normals` must have shape [P, 3].")
    _check(grad_sigmas.dim() == 2, "`grad_sigmas` must have shape [A, P].")
    _check(grad_normals.size(0) == P, "`grad_normals` point dimension mismatch.")
    _check(tuple(grad_sigmas.shape) == (A, P), "`grad_sigmas` shape mismatch.")
    dtype = _float_dtype((grad_output_real, grad_output_imag, normals, sigmas, points, tx_pos,
                          rx_pos, grad_normals, grad_sigmas), "signal tracer")
    prm = tracer_params(start_frequency, chirp_slope_scaled, adc_sample_rate)
    with torch.cuda.device(normals.device):
        _native.call(f"geraf_st_backward_{_sfx(dtype)}", *map(_native.ptr, ts), A, P, K, prm,
                     _native.stream_of(normals))


def BilinearRaysample_forward(normals, sigmas, z_vals, target_z_vals, output_normals,
                              output_sigmas, output_inbounds, num_antennas: int, num_rays: int,
                              num_samples: int, num_target_samples: int) -> None:
    """In-place resampler forward; reference: SignalTracer_cuda.cu:10-58."""
    names = ("normals", "sigmas", "z_vals", "target_z_vals", "output_normals", "output_sigmas",
             "output_inbounds")
    ts = (normals, sigmas, z_vals, target_z_vals, output_normals, output_sigmas, output_inbounds)
    for t, n in zip(ts, names):
        _cuda(t, n)
    A, R, S, T = int(num_antennas), int(num_rays), int(num_samples), int(num_target_samples)
    _check(normals.dim() == 3, "`normals` must have shape [R, S, 3].")
    _check(sigmas.dim() == 3, "`sigmas` must have shape [A, R, S].")
    _check(z_vals.dim() == 2, "`z_vals` must have shape [R, S].")
    _check(target_z_vals.dim() == 2, "`target_z_vals` must have shape [R, T].")
    _check(output_normals.dim() == 3, "`output_normals` must have shape [R, T, 3].")
    _check(output_sigmas.dim() == 3, "`output_sigmas` must have shape [A, R, T].")
    _check(output_inbounds.dim() == 2, "`output_inbounds` must have shape [R, T].")
    _check(output_inbounds.dtype == torch.bool, "`output_inbounds` must be bool.")
    _check(tuple(normals.shape) == (R, S, 3), "`normals` shape mismatch.")
    _check(tuple(sigmas.shape) == (A, R, S), "`sigmas` shape mismatch.")
    _check(tuple(z_vals.shape) == (R, S), "`z_vals` shape mismatch.")
    _check(tuple(target_z_vals.shape) == (R, T), "`target_z_vals` shape mismatch.")
    _check(tuple(output_normals.shape) == (R, T, 3), "`output_normals` shape mismatch.")
    _check(tuple(output_sigmas.shape) == (A, R, T), "`output_sigmas` shape mismatch.")
    _check(tuple(output_inbounds.shape) == (R, T), "`output_inbounds` shape mismatch.")
    dtype = _float_dtype((normals, sigmas, z_vals, target_z_vals, output_normals, output_sigmas),
                         "bilinear ray sampler")
    with torch.cuda.device(normals.device):
        _native.call(f"geraf_bilinear_forward_{_sfx(dtype)}", *map(_native.ptr, ts), A, R, S, T,
                     _native.stream_of(normals))


def BilinearRaysample_backward(grad_output_normals, grad_output_sigmas, z_vals, target_z_vals,
                               grad_normals, grad_sigmas, num_antennas: int, num_rays: int,
                               num_samples: int, num_target_samples: int) -> None:
    """In-place resampler backward (every element written); reference:
    SignalTracer_cuda.cu:60-99."""
    names = ("grad_output_normals", "grad_output_sigmas", "z_vals", "target_z_vals",
             "grad_normals", "grad_sigmas")
    ts = (grad_output_normals, grad_output_sigmas, z_vals, target_z_vals, grad_normals,
          grad_sigmas)
    for t, n in zip(ts, names):
        _cuda(t, n)
    A, R, S, T = int(num_antennas), int(num_rays), int(num_samples), int(num_target_samples)
    _check(grad_output_normals.dim() == 3, "`grad_output_normals` must have shape [R, T, 3].")
    _check(grad_output_sigmas.dim() == 3, "`grad_output_sigmas` must have shape [A, R, T].")
    _check(grad_normals.dim() == 3, "`grad_normals` must have shape [R, S, 3].")
    _check(grad_sigmas.dim() == 3, "`grad_sigmas` must have shape [A, R, S].")
    _check(tuple(grad_output_normals.shape) == (R, T, 3), "`grad_output_normals` shape mismatch.")
    _check(tuple(grad_output_sigmas.shape) == (A, R, T), "`grad_output_sigmas` shape mismatch.")
    _check(tuple(z_vals.shape) == (R, S), "`z_vals` shape mismatch.")
    _check(tuple(target_z_vals.shape) == (R, T), "`target_z_vals` shape mismatch.")
    _check(tuple(grad_normals.shape) == (R, S, 3), "`grad_normals` shape mismatch.")
    _check(tuple(grad_sigmas.shape) == (A, R, S), "`grad_sigmas` shape mismatch.")
    dtype = _float_dtype(ts, "bilinear ray sampler")
    with torch.cuda.device(z_vals.device):
        _native.call(f"geraf_bilinear_backward_{_sfx(dtype)}", *map(_native.ptr, ts), A, R, S, T,
                     _native.stream_of(z_vals))


def get_compiler_version() -> str:
    return _native.version()


def get_cuda_version() -> str:
    return f"HIP {torch.version.hip}"


def has_cuda() -> bool:
    return True
