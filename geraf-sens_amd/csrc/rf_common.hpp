// Shared device helpers for the GeRaF RF kernels (gfx950).
//
// Phase model shared by the matched filter and the signal tracer
// (MatchedFilter_cuda_kernel.cu:28-63, SignalTracer_cuda_kernel.cu:433-443):
//   tau      = (|tx - p| + |rx - p| + path_bias) / c          [1e-8 s]
//   phi_k    = 2*pi * (f0 + k * slope / fs) * tau              [rad]
// with f0 in 1e8 Hz, slope in 1e8 Hz/s, fs in Hz and c in 1e8 m/s.  All phase
// work here is done in cycles: cycles_k = f0*tau + k*(slope/fs)*tau.  The
// carrier term f0*tau is a few hundred cycles, so it is formed and reduced
// modulo 1 in FP64; everything after the reduction runs in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geraf_rf.h"

namespace geraf {

constexpr int kWave = 64;

// Host-side derived constants handed to every kernel by value.
struct PhaseConsts {
  double f0;     // chirp start frequency [1e8 Hz]
  double dfs;    // frequency step per ADC sample, slope / fs [1e8 Hz]
  double inv_c;  // 1 / c  [1e-8 s / m]
  double bias;   // path-length bias added before the phase [m]
};

inline PhaseConsts make_phase_consts(const geraf_fmcw_params& p) {
  PhaseConsts c;
  c.f0 = p.f0;
  c.dfs = p.slope / p.fs;
  c.inv_c = 1.0 / p.c;
  c.bias = p.path_bias;
  return c;
}

__device__ __forceinline__ double dist3_f64(double ax, double ay, double az, double bx,
                                            double by, double bz) {
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  return sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
}

// Fractional part of x in [-0.5, 0.5].
__device__ __forceinline__ double centered_frac(double x) { return x - rint(x); }

// (cos(2*pi*x), sign*sin(2*pi*x)) for |x| <= 0.5 (x in cycles).
__device__ __forceinline__ float2 phasor_cycles(float x, float sign) {
  float s, c;
  sincospif(2.0f * x, &s, &c);
  return make_float2(c, sign * s);
}

// Complex helpers on float2 = (re, im).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// Compensated (Kahan) accumulation of a complex value.
struct KahanC {
  float sr, si, cr, ci;
  __device__ __forceinline__ void init() { sr = si = cr = ci = 0.0f; }
  __device__ __forceinline__ void add(float2 v) {
    const float yr = v.x - cr;
    const float tr = sr + yr;
    cr = (tr - sr) - yr;
    sr = tr;
    const float yi = v.y - ci;
    const float ti = si + yi;
    ci = (ti - si) - yi;
    si = ti;
  }
  // Best FP64 estimate of the compensated sum.
  __device__ __forceinline__ double re() const { return (double)sr - (double)cr; }
  __device__ __forceinline__ double im() const { return (double)si - (double)ci; }
};

}  // namespace geraf
