// Softplus for the BF16 hidden layers of the SDF network, with its first and
// second derivatives, one pass each.  torch.nn.Softplus(beta, threshold)
// semantics (the activation of sdf_network.py:78-199, Softplus(beta=100)):
//   y  = x                          if beta x > threshold
//      = log1p(exp(beta x)) / beta  otherwise
//   dy/dx = s = 1 if beta x > threshold, else sigmoid(beta x) = e / (1 + e),
//   d^2y/dx^2 = beta s (1 - s) = beta e / (1 + e)^2   (0 above the threshold;
//   the second form has no cancellation where s rounds to 1 in FP32)
// Values are BF16 in memory (raw 16-bit patterns) and FP32 in registers;
// results are rounded to nearest even, as torch rounds FP32 to BF16.  s is
// recomputed from x in FP32 (as torch's softplus_backward does), never
// stored.  Each thread handles 8 consecutive elements with 16-byte accesses
// (a scalar tail covers n % 8 and unaligned pointers).
#include <hip/hip_runtime.h>

#include <stdint.h>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

__device__ __forceinline__ float bf16_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t f32_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

// log(1 + e) for e >= 0 to a few ulp: log(u) * e / (u - 1) with u = 1 + e
// corrects the rounding of u (Goldberg's method); e itself when u rounds to 1.
__device__ __forceinline__ float log1p_pos(float e) {
  const float u = 1.0f + e;
  if (u == 1.0f) return e;
  return __logf(u) * (e / (u - 1.0f));
}

struct SoftplusTerms {
  float y, s, c;  // value, slope, curvature / beta = s (1 - s)
};

__device__ __forceinline__ SoftplusTerms softplus_terms(float x, float beta, float threshold) {
  const float t = x * beta;
  if (t > threshold) return SoftplusTerms{x, 1.0f, 0.0f};
  const float e = __expf(t);
  const float r = 1.0f / (1.0f + e);
  return SoftplusTerms{log1p_pos(e) / beta, e * r, e * r * r};
}

constexpr int kVec = 8;

// Loads/stores of 8 BF16 values, vectorised when the address allows.
struct Vec8 {
  uint16_t v[kVec];
};
__device__ __forceinline__ Vec8 load8(const uint16_t* p, int64_t i, int64_t n) {
  Vec8 r;
  if (i + kVec <= n && (reinterpret_cast<uintptr_t>(p + i) & 15u) == 0) {
    const uint4 q = *reinterpret_cast<const uint4*>(p + i);
    const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      r.v[2 * j] = (uint16_t)(w[j] & 0xffffu);
      r.v[2 * j + 1] = (uint16_t)(w[j] >> 16);
    }
  } else {
#pragma unroll
    for (int j = 0; j < kVec; ++j) r.v[j] = i + j < n ? p[i + j] : (uint16_t)0;
  }
  return r;
}
__device__ __forceinline__ void store8(uint16_t* p, int64_t i, int64_t n, const Vec8& r) {
  if (i + kVec <= n && (reinterpret_cast<uintptr_t>(p + i) & 15u) == 0) {
    uint4 q;
    q.x = (uint32_t)r.v[0] | ((uint32_t)r.v[1] << 16);
    q.y = (uint32_t)r.v[2] | ((uint32_t)r.v[3] << 16);
    q.z = (uint32_t)r.v[4] | ((uint32_t)r.v[5] << 16);
    q.w = (uint32_t)r.v[6] | ((uint32_t)r.v[7] << 16);
    *reinterpret_cast<uint4*>(p + i) = q;
  } else {
#pragma unroll
    for (int j = 0; j < kVec; ++j)
      if (i + j < n) p[i + j] = r.v[j];
  }
}

__global__ __launch_bounds__(256) void softplus_kernel(const uint16_t* __restrict__ x,
                                                       uint16_t* __restrict__ y, int64_t n,
                                                       float beta, float threshold) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kVec;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kVec; i < n; i += stride) {
    const Vec8 a = load8(x, i, n);
    Vec8 o;
#pragma unroll
    for (int j = 0; j < kVec; ++j)
      o.v[j] = f32_bf16(softplus_terms(bf16_f32(a.v[j]), beta, threshold).y);
    store8(y, i, n, o);
  }
}

// gx = g * s(x)
__global__ __launch_bounds__(256) void softplus_grad_kernel(const uint16_t* __restrict__ g,
                                                            const uint16_t* __restrict__ x,
                                                            uint16_t* __restrict__ gx, int64_t n,
                                                            float beta, float threshold) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kVec;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kVec; i < n; i += stride) {
    const Vec8 a = load8(g, i, n), b = load8(x, i, n);
    Vec8 o;
#pragma unroll
    for (int j = 0; j < kVec; ++j)
      o.v[j] = f32_bf16(bf16_f32(a.v[j]) * softplus_terms(bf16_f32(b.v[j]), beta, threshold).s);
    store8(gx, i, n, o);
  }
}

// Adjoint of gx = g * s(x) for an incoming ggx:
//   d/dg = ggx * s(x),   d/dx = ggx * g * beta * e / (1 + e)^2
__global__ __launch_bounds__(256) void softplus_grad2_kernel(
    const uint16_t* __restrict__ ggx, const uint16_t* __restrict__ g,
    const uint16_t* __restrict__ x, uint16_t* __restrict__ out_g, uint16_t* __restrict__ out_x,
    int64_t n, float beta, float threshold) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x * kVec;
  for (int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * kVec; i < n; i += stride) {
    const Vec8 a = load8(ggx, i, n), b = load8(g, i, n), c = load8(x, i, n);
    Vec8 og, ox;
#pragma unroll
    for (int j = 0; j < kVec; ++j) {
      const float gg = bf16_f32(a.v[j]);
      const SoftplusTerms q = softplus_terms(bf16_f32(c.v[j]), beta, threshold);
      og.v[j] = f32_bf16(gg * q.s);
      ox.v[j] = f32_bf16(gg * bf16_f32(b.v[j]) * (beta * q.c));
    }
    store8(out_g, i, n, og);
    store8(out_x, i, n, ox);
  }
}

unsigned grid_for(int64_t n) {
  const int64_t chunks = (n + kVec - 1) / kVec;
  const int64_t blocks = (chunks + 255) / 256;
  return (unsigned)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536);
}

void check_softplus(int64_t n, float beta, float threshold) {
  if (n < 0) throw ArgError("n must be non-negative");
  if (!(beta > 0.0f) || !std::isfinite(beta)) throw ArgError("beta must be positive and finite");
  if (!std::isfinite(threshold)) throw ArgError("threshold must be finite");
}

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int geraf_softplus_bf16(const uint16_t* x, uint16_t* y, int64_t n, float beta,
                                   float threshold, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_softplus(n, beta, threshold);
  if (n == 0) return GERAF_RF_OK;
  require_ptrs({x, y}, true);
  hipLaunchKernelGGL(softplus_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y,
                     n, beta, threshold);
  return check_launch("softplus_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_softplus_grad_bf16(const uint16_t* g, const uint16_t* x, uint16_t* gx,
                                        int64_t n, float beta, float threshold,
                                        geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_softplus(n, beta, threshold);
  if (n == 0) return GERAF_RF_OK;
  require_ptrs({g, x, gx}, true);
  hipLaunchKernelGGL(softplus_grad_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream,
                     g, x, gx, n, beta, threshold);
  return check_launch("softplus_grad_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_softplus_grad2_bf16(const uint16_t* ggx, const uint16_t* g,
                                         const uint16_t* x, uint16_t* out_g, uint16_t* out_x,
                                         int64_t n, float beta, float threshold,
                                         geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_softplus(n, beta, threshold);
  if (n == 0) return GERAF_RF_OK;
  require_ptrs({ggx, g, x, out_g, out_x}, true);
  hipLaunchKernelGGL(softplus_grad2_kernel, dim3(grid_for(n)), dim3(256), 0,
                     (hipStream_t)stream, ggx, g, x, out_g, out_x, n, beta, threshold);
  return check_launch("softplus_grad2_kernel");
  GERAF_TRY_END
}
