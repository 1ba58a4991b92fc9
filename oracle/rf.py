"""ctypes bindings of oracle/_build/librf_oracle.so (see rf_oracle.c).

TEST INFRASTRUCTURE ONLY - the checker, never the thing measured or shipped.
All functions take and return numpy arrays; inputs are copied to contiguous
float64 (float32 for the *_ref32 form) before the call.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "librf_oracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("f0", "slope", "fs", "c", "bias")]


_lib = None


def build() -> None:
    subprocess.run(["make", "-C", _HERE, "-s"], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        for fn in ("orc_mf_forward_f64", "orc_mf_forward_ref32", "orc_mf_backward_f64",
                   "orc_st_forward_f64", "orc_st_backward_f64", "orc_bilinear_forward_f64",
                   "orc_bilinear_backward_f64", "orc_mf_fft_forward_f64",
                   "orc_mf_fft_backward_f64"):
            getattr(lib, fn).restype = None
        _lib = lib
    return _lib


def _d(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _f(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32))


def _p(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _params(f0, slope, fs, c, bias) -> OrcParams:
    return OrcParams(float(f0), float(slope), float(fs), float(c), float(bias))


def mf_forward(real, imag, tx, rx, voxel_n3, f0, slope, fs, c, bias, form="f64"):
    """Matched filter on voxels [N,3]; returns complex out [N] (complex128, or
    complex64 for form='ref32')."""
    lib = _load()
    A, K = np.shape(real)
    N = np.shape(voxel_n3)[0]
    pr = _params(f0, slope, fs, c, bias)
    if form == "f64":
        cv = _d
        fn = lib.orc_mf_forward_f64
        odt = np.float64
    elif form == "ref32":
        cv = _f
        fn = lib.orc_mf_forward_ref32
        odt = np.float32
    else:
        raise ValueError(form)
    re, im, t, r = cv(real), cv(imag), cv(tx), cv(rx)
    vox = cv(np.asarray(voxel_n3).T)
    o_r = np.zeros(N, odt)
    o_i = np.zeros(N, odt)
    fn(_p(re), _p(im), _p(t), _p(r), _p(vox), _p(o_r), _p(o_i), ctypes.c_int64(A),
       ctypes.c_int64(K), ctypes.c_int64(N), ctypes.byref(pr))
    return o_r + 1j * o_i


def mf_backward(g_out, tx, rx, voxel_n3, K, f0, slope, fs, c, bias, norm):
    """Adjoint: complex g_out [N] -> complex grad [A,K] (FP64)."""
    lib = _load()
    g_out = np.asarray(g_out)
    gr, gi = _d(g_out.real), _d(g_out.imag)
    t, r = _d(tx), _d(rx)
    A = t.shape[0]
    N = gr.shape[0]
    vox = _d(np.asarray(voxel_n3).T)
    o_r = np.zeros((A, K))
    o_i = np.zeros((A, K))
    lib.orc_mf_backward_f64(_p(gr), _p(gi), _p(t), _p(r), _p(vox), _p(o_r), _p(o_i),
                            ctypes.c_int64(A), ctypes.c_int64(K), ctypes.c_int64(N),
                            ctypes.c_int64(norm), ctypes.byref(_params(f0, slope, fs, c, bias)))
    return o_r + 1j * o_i


def st_forward(normals, sigmas, points, tx, rx, inbounds, K, f0, slope, fs, c, bias):
    """[A,P] signal tracer forward; returns complex [A,K] (FP64)."""
    lib = _load()
    n, s, pt, t, r = map(_d, (normals, sigmas, points, tx, rx))
    inb = np.ascontiguousarray(np.asarray(inbounds, dtype=np.uint8))
    A, P = s.shape
    o_r = np.zeros((A, K))
    o_i = np.zeros((A, K))
    lib.orc_st_forward_f64(_p(n), _p(s), _p(pt), _p(t), _p(r), _p(inb), _p(o_r), _p(o_i),
                           ctypes.c_int64(A), ctypes.c_int64(P), ctypes.c_int64(K),
                           ctypes.byref(_params(f0, slope, fs, c, bias)))
    return o_r + 1j * o_i


def st_backward(grad_out, normals, sigmas, points, tx, rx, inbounds, f0, slope, fs, c, bias):
    """[A,P] signal tracer backward; complex grad_out [A,K] -> (grad_normals [P,3],
    grad_sigmas [A,P]) in FP64 (zeros for out-of-bounds points)."""
    lib = _load()
    g = np.asarray(grad_out)
    gr, gi = _d(g.real), _d(g.imag)
    n, s, pt, t, r = map(_d, (normals, sigmas, points, tx, rx))
    inb = np.ascontiguousarray(np.asarray(inbounds, dtype=np.uint8))
    A, P = s.shape
    K = gr.shape[1]
    gn = np.zeros((P, 3))
    gs = np.zeros((A, P))
    lib.orc_st_backward_f64(_p(gr), _p(gi), _p(n), _p(s), _p(pt), _p(t), _p(r), _p(inb),
                            _p(gn), _p(gs), ctypes.c_int64(A), ctypes.c_int64(P),
                            ctypes.c_int64(K), ctypes.byref(_params(f0, slope, fs, c, bias)))
    return gn, gs


def bilinear_forward(normals, sigmas, z, tz):
    """normals [R,S,3], sigmas [A,R,S], z [R,S], tz [R,T] -> (n [R,T,3], s [A,R,T], inb [R,T])."""
    lib = _load()
    n, s, zz, t = map(_d, (normals, sigmas, z, tz))
    A, R, S = s.shape
    T = t.shape[1]
    on = np.zeros((R, T, 3))
    os_ = np.zeros((A, R, T))
    inb = np.zeros((R, T), np.uint8)
    lib.orc_bilinear_forward_f64(_p(n), _p(s), _p(zz), _p(t), _p(on), _p(os_), _p(inb),
                                 ctypes.c_int64(A), ctypes.c_int64(R), ctypes.c_int64(S),
                                 ctypes.c_int64(T))
    return on, os_, inb.astype(bool)


def bilinear_backward(g_n, g_s, z, tz, S):
    """Transpose of bilinear_forward: (g_n [R,T,3], g_s [A,R,T]) -> (gn [R,S,3], gs [A,R,S])."""
    lib = _load()
    gn_in, gs_in, zz, t = map(_d, (g_n, g_s, z, tz))
    A, R, T = gs_in.shape
    gn = np.zeros((R, S, 3))
    gs = np.zeros((A, R, S))
    lib.orc_bilinear_backward_f64(_p(gn_in), _p(gs_in), _p(zz), _p(t), _p(gn), _p(gs),
                                  ctypes.c_int64(A), ctypes.c_int64(R), ctypes.c_int64(S),
                                  ctypes.c_int64(T))
    return gn, gs


def _fft_max_d(spacing, bins):
    return float(np.float32(np.float32(spacing) * np.float32(bins)))


def mf_fft_forward(real, imag, tx, rx, voxel_n3, spacing, f0, c, bias, slope=0.0, fs=1.0):
    """Range-bin matched filter: echoes [A,R] -> complex [N] (FP64)."""
    lib = _load()
    re, im, t, r = map(_d, (real, imag, tx, rx))
    A, R = re.shape
    vox = _d(np.asarray(voxel_n3).T)
    N = vox.shape[1]
    o_r = np.zeros(N)
    o_i = np.zeros(N)
    lib.orc_mf_fft_forward_f64(_p(re), _p(im), _p(t), _p(r), _p(vox), _p(o_r), _p(o_i),
                               ctypes.c_int64(A), ctypes.c_int64(R), ctypes.c_int64(N),
                               ctypes.c_double(spacing), ctypes.c_double(_fft_max_d(spacing, R)),
                               ctypes.byref(_params(f0, slope, fs, c, bias)))
    return o_r + 1j * o_i


def mf_fft_backward(g_out, tx, rx, voxel_n3, R, spacing, norm, f0, c, bias, slope=0.0, fs=1.0):
    """Adjoint of mf_fft_forward: complex [N] -> complex [A,R] (FP64)."""
    lib = _load()
    g = np.asarray(g_out)
    gr, gi, t, r = _d(g.real), _d(g.imag), _d(tx), _d(rx)
    A = t.shape[0]
    vox = _d(np.asarray(voxel_n3).T)
    N = vox.shape[1]
    o_r = np.zeros((A, R))
    o_i = np.zeros((A, R))
    lib.orc_mf_fft_backward_f64(_p(gr), _p(gi), _p(t), _p(r), _p(vox), _p(o_r), _p(o_i),
                                ctypes.c_int64(A), ctypes.c_int64(R), ctypes.c_int64(N),
                                ctypes.c_int64(norm), ctypes.c_double(spacing),
                                ctypes.c_double(_fft_max_d(spacing, R)),
                                ctypes.byref(_params(f0, slope, fs, c, bias)))
    return o_r + 1j * o_i
