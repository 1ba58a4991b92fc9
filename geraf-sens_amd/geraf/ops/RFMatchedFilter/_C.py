"""Drop-in for the reference's pybind module ``geraf.ops.RFMatchedFilter._C``.

Same function names, argument order and in-place output contract as
RFMatchedFilter/csrc/vision.cpp:70-97 and the checks of
MatchedFilter_cuda.cu:10-107 (failures raise ``RuntimeError`` like
``TORCH_CHECK``).  Each call forwards device pointers to libgeraf_rf.so on the
current HIP stream of the input's device; nothing runs on the host.

Scalar arguments are rounded to FP32 on entry, as the reference's C++
signatures (``float adc_sample_rate, float chirp_slope_scaled``) do, and the
reference's hard-coded FP32 constants (start frequency 773.704, c = 2.9979,
path bias 0.15; MatchedFilter_cuda_kernel.cu:14-16) are used at their FP32
values.  float32 tensors run the tuned kernel; float64 tensors run the FP64
kernel; float16 tensors are widened to float32 around the tuned kernel.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _native

_F32 = lambda x: float(np.float32(x))  # noqa: E731  (pybind float conversion)

START_FREQUENCY = _F32(_native.MF_START_FREQUENCY_SCALED)
SPEED_OF_LIGHT = _F32(_native.SPEED_OF_LIGHT_SCALED)
PATH_BIAS = _F32(_native.PATH_LENGTH_BIAS)


def _check(cond: bool, msg: str) -> None:
    if not cond:
        raise RuntimeError(msg)


def _check_common(real, imag, tx_pos, rx_pos, voxel) -> None:
    for t, n in ((real, "real"), (imag, "imag"), (tx_pos, "tx_pos"), (rx_pos, "rx_pos"),
                 (voxel, "voxel")):
        _check(t.is_cuda, f"`{n}` must be a CUDA tensor.")
    for t, n in ((real, "real"), (imag, "imag"), (tx_pos, "tx_pos"), (rx_pos, "rx_pos"),
                 (voxel, "voxel")):
        _check(t.is_contiguous(), f"`{n}` must be contiguous.")
    _check(real.dim() == 2, "`real` must have shape [A, S].")
    _check(imag.dim() == 2, "`imag` must have shape [A, S].")
    _check(tx_pos.dim() == 2, "`tx_pos` must have shape [A, 3].")
    _check(rx_pos.dim() == 2, "`rx_pos` must have shape [A, 3].")
    _check(voxel.dim() == 2, "`voxel` must have shape [3, N].")
    _check(real.shape == imag.shape, "`real` and `imag` must match.")
    _check(tx_pos.size(1) == 3, "`tx_pos` must store XYZ triplets.")
    _check(rx_pos.size(1) == 3, "`rx_pos` must store XYZ triplets.")


def _check_dtypes(*tensors) -> torch.dtype:
    dtype = tensors[0].dtype
    _check(dtype in (torch.float32, torch.float64, torch.float16),
           f"matched filter supports float32/float64/float16, got {dtype}")
    for t in tensors[1:]:
        _check(t.dtype == dtype, f"expected scalar type {dtype} but found {t.dtype}")
    return dtype


def _params(adc_sample_rate: float, chirp_slope_scaled: float) -> _native.FmcwParams:
    return _native.make_params(START_FREQUENCY, _F32(chirp_slope_scaled), _F32(adc_sample_rate),
                               SPEED_OF_LIGHT, PATH_BIAS)


def MatchedFilter_forward(real, imag, tx_pos, rx_pos, output_real, output_imag, voxel,
                          total_ant_size: int, num_adc_samples: int, adc_sample_rate: float,
                          chirp_slope_scaled: float, num_voxels: int) -> None:
    """In-place forward; reference: MatchedFilter_cuda.cu:39-69,345-371."""
    _check_common(real, imag, tx_pos, rx_pos, voxel)
    for t, n in ((output_real, "output_real"), (output_imag, "output_imag")):
        _check(t.is_cuda, f"`{n}` must be a CUDA tensor.")
        _check(t.is_contiguous(), f"`{n}` must be contiguous.")
        _check(t.dim() == 1, f"`{n}` must have shape [N].")
    A, K, N = int(total_ant_size), int(num_adc_samples), int(num_voxels)
    _check(real.size(0) == A, "`real` antenna dimension mismatch.")
    _check(real.size(1) == K, "`real` sample dimension mismatch.")
    _check(imag.size(0) == A, "`imag` antenna dimension mismatch.")
    _check(imag.size(1) == K, "`imag` sample dimension mismatch.")
    _check(tx_pos.size(0) == A, "`tx_pos` antenna dimension mismatch.")
    _check(rx_pos.size(0) == A, "`rx_pos` antenna dimension mismatch.")
    _check(voxel.size(0) == 3, "`voxel` must be laid out as [3, N].")
    _check(voxel.size(1) == N, "`voxel` count mismatch.")
    _check(output_real.size(0) == N, "`output_real` voxel dimension mismatch.")
    _check(output_imag.size(0) == N, "`output_imag` voxel dimension mismatch.")
    dtype = _check_dtypes(real, imag, tx_pos, rx_pos, output_real, output_imag, voxel)
    params = _params(adc_sample_rate, chirp_slope_scaled)
    with torch.cuda.device(real.device):
        if dtype == torch.float16:
            f = [t.float() for t in (real, imag, tx_pos, rx_pos, voxel)]
            o_r = torch.empty(N, dtype=torch.float32, device=real.device)
            o_i = torch.empty_like(o_r)
            _forward_f32(*f[:4], o_r, o_i, f[4], A, K, N, params)
            output_real.copy_(o_r)
            output_imag.copy_(o_i)
            return
        if dtype == torch.float32:
            _forward_f32(real, imag, tx_pos, rx_pos, output_real, output_imag, voxel, A, K, N,
                         params)
            return
        _native.call("geraf_mf_forward_f64", *map(_native.ptr, (real, imag, tx_pos, rx_pos,
                                                                output_real, output_imag, voxel)),
                     A, K, N, params, _native.stream_of(real))


# K range of the matrix-core forward (geraf_mf_forward_f32_mma).  Its time is
# flat up to K = 512 (one 512-frequency tile), the Horner kernel's grows with
# K: at 4,096 antennas x 262,144 voxels Horner takes 9.2, 15.2 and 29.2 ms at
# K = 64, 128 and 256, the matrix-core form 19.0-19.8 ms, so they cross near
# K = 168 (and at 64x64 x 128^3 the matrix-core form is 2.2x faster at 512).
# The instruction handles up to 2048 samples per antenna.
MMA_MIN_K = 161
MMA_MAX_K = 2048


def _forward_f32(real, imag, tx_pos, rx_pos, out_r, out_i, voxel, A, K, N, params) -> None:
    ptrs = map(_native.ptr, (real, imag, tx_pos, rx_pos, out_r, out_i, voxel))
    st = _native.stream_of(real)
    if MMA_MIN_K <= K <= MMA_MAX_K and A > 0 and N > 0:
        size = int(_native.load().geraf_mf_forward_mma_workspace_bytes(A, K))
        ws = torch.empty(size, dtype=torch.uint8, device=real.device)
        _native.call("geraf_mf_forward_f32_mma", *ptrs, A, K, N, params, _native.ptr(ws), size,
                     st)
        return
    _native.call("geraf_mf_forward_f32", *ptrs, A, K, N, params, st)


def MatchedFilter_backward_realimag(grad_output_real, grad_output_imag, real, imag, tx_pos,
                                    rx_pos, grad_real, grad_imag, voxel, total_ant_size: int,
                                    num_adc_samples: int, adc_sample_rate: float,
                                    chirp_slope_scaled: float, num_voxels: int) -> None:
    """In-place adjoint into the echoes; reference: MatchedFilter_cuda.cu:71-107,253-296.

    The result is divided by ``total_ant_size`` as the reference kernel does
    (MatchedFilter_cuda_kernel.cu:217-219); ``real``/``imag`` are only checked.
    """
    _check_common(real, imag, tx_pos, rx_pos, voxel)
    for t, n in ((grad_output_real, "grad_output_real"), (grad_output_imag, "grad_output_imag"),
                 (grad_real, "grad_real"), (grad_imag, "grad_imag")):
        _check(t.is_cuda, f"`{n}` must be a CUDA tensor.")
        _check(t.is_contiguous(), f"`{n}` must be contiguous.")
    _check(grad_output_real.dim() == 1, "`grad_output_real` must have shape [N].")
    _check(grad_output_imag.dim() == 1, "`grad_output_imag` must have shape [N].")
    _check(grad_real.dim() == 2, "`grad_real` must have shape [A, S].")
    _check(grad_imag.dim() == 2, "`grad_imag` must have shape [A, S].")
    A, K, N = int(total_ant_size), int(num_adc_samples), int(num_voxels)
    _check(grad_output_real.size(0) == N, "`grad_output_real` voxel dimension mismatch.")
    _check(grad_output_imag.size(0) == N, "`grad_output_imag` voxel dimension mismatch.")
    _check(grad_real.size(0) == A, "`grad_real` antenna dimension mismatch.")
    _check(grad_real.size(1) == K, "`grad_real` sample dimension mismatch.")
    _check(grad_imag.size(0) == A, "`grad_imag` antenna dimension mismatch.")
    _check(grad_imag.size(1) == K, "`grad_imag` sample dimension mismatch.")
    # The reference kernel reads positions of antennas [0, A) and voxels [0, N).
    _check(tx_pos.size(0) >= A and rx_pos.size(0) >= A, "`tx_pos` antenna dimension mismatch.")
    _check(voxel.size(0) == 3 and voxel.size(1) >= N, "`voxel` count mismatch.")
    dtype = _check_dtypes(grad_output_real, grad_output_imag, tx_pos, rx_pos, grad_real,
                          grad_imag, voxel)
    params = _params(adc_sample_rate, chirp_slope_scaled)
    with torch.cuda.device(grad_real.device):
        if dtype == torch.float16:
            f = [t.float() for t in (grad_output_real, grad_output_imag, tx_pos, rx_pos, voxel)]
            g_r = torch.empty((A, K), dtype=torch.float32, device=grad_real.device)
            g_i = torch.empty_like(g_r)
            _native.call("geraf_mf_backward_f32",
                         *map(_native.ptr, (f[0], f[1], f[2], f[3], g_r, g_i, f[4])),
                         A, K, N, A, params, _native.stream_of(g_r))
            grad_real.copy_(g_r)
            grad_imag.copy_(g_i)
            return
        name = "geraf_mf_backward_f32" if dtype == torch.float32 else "geraf_mf_backward_f64"
        _native.call(name, *map(_native.ptr, (grad_output_real, grad_output_imag, tx_pos, rx_pos,
                                             grad_real, grad_imag, voxel)),
                     A, K, N, A, params, _native.stream_of(grad_real))


def MatchedFilter_fft_forward(real, imag, tx_pos, rx_pos, output_real, output_imag, voxel,
                              total_ant_size: int, num_range_bins: int, fft_spacing: float,
                              adc_sample_rate: float, chirp_slope_scaled: float,
                              num_voxels: int) -> None:
    """In-place range-bin forward; reference: MatchedFilter_cuda.cu:212-250 (same
    checks as MatchedFilter_forward with num_range_bins as the sample count)."""
    _check_common(real, imag, tx_pos, rx_pos, voxel)
    for t, n in ((output_real, "output_real"), (output_imag, "output_imag")):
        _check(t.is_cuda, f"`{n}` must be a CUDA tensor.")
        _check(t.is_contiguous(), f"`{n}` must be contiguous.")
        _check(t.dim() == 1, f"`{n}` must have shape [N].")
    A, R, N = int(total_ant_size), int(num_range_bins), int(num_voxels)
    _check(tuple(real.shape) == (A, R), "`real` antenna dimension mismatch.")
    _check(tuple(imag.shape) == (A, R), "`imag` antenna dimension mismatch.")
    _check(tx_pos.size(0) == A and rx_pos.size(0) == A, "`tx_pos` antenna dimension mismatch.")
    _check(tuple(voxel.shape) == (3, N), "`voxel` must be laid out as [3, N].")
    _check(output_real.size(0) == N and output_imag.size(0) == N,
           "`output_real` voxel dimension mismatch.")
    dtype = _check_dtypes(real, imag, tx_pos, rx_pos, output_real, output_imag, voxel)
    _check(dtype != torch.float16, "range-bin matched filter supports float32/float64")
    params = _params(adc_sample_rate, chirp_slope_scaled)
    name = "geraf_mf_fft_forward_" + ("f32" if dtype == torch.float32 else "f64")
    with torch.cuda.device(real.device):
        _native.call(name, *map(_native.ptr, (real, imag, tx_pos, rx_pos, output_real,
                                             output_imag, voxel)),
                     A, R, N, _F32(fft_spacing), params, _native.stream_of(real))


def MatchedFilter_fft_backward_realimag(grad_output_real, grad_output_imag, real, imag, tx_pos,
                                        rx_pos, grad_real, grad_imag, voxel,
                                        total_ant_size: int, num_range_bins: int,
                                        fft_spacing: float, adc_sample_rate: float,
                                        chirp_slope_scaled: float, num_voxels: int) -> None:
    """In-place range-bin adjoint, divided by total_ant_size; reference:
    MatchedFilter_cuda.cu:298-343."""
    _check_common(real, imag, tx_pos, rx_pos, voxel)
    A, R, N = int(total_ant_size), int(num_range_bins), int(num_voxels)
    for t, n in ((grad_output_real, "grad_output_real"), (grad_output_imag, "grad_output_imag"),
                 (grad_real, "grad_real"), (grad_imag, "grad_imag")):
        _check(t.is_cuda, f"`{n}` must be a CUDA tensor.")
        _check(t.is_contiguous(), f"`{n}` must be contiguous.")
    _check(tuple(grad_output_real.shape) == (N,), "`grad_output_real` must have shape [N].")
    _check(tuple(grad_output_imag.shape) == (N,), "`grad_output_imag` must have shape [N].")
    _check(tuple(grad_real.shape) == (A, R), "`grad_real` must have shape [A, S].")
    _check(tuple(grad_imag.shape) == (A, R), "`grad_imag` must have shape [A, S].")
    dtype = _check_dtypes(grad_output_real, grad_output_imag, tx_pos, rx_pos, grad_real,
                          grad_imag, voxel)
    _check(dtype != torch.float16, "range-bin matched filter supports float32/float64")
    params = _params(adc_sample_rate, chirp_slope_scaled)
    name = "geraf_mf_fft_backward_" + ("f32" if dtype == torch.float32 else "f64")
    with torch.cuda.device(grad_real.device):
        _native.call(name, *map(_native.ptr, (grad_output_real, grad_output_imag, tx_pos, rx_pos,
                                             grad_real, grad_imag, voxel)),
                     A, R, N, A, _F32(fft_spacing), params, _native.stream_of(grad_real))


def get_compiler_version() -> str:
    return _native.version()


def get_cuda_version() -> str:
    return f"HIP {torch.version.hip}"


def has_cuda() -> bool:
    return True
