// Transmittance calibration of the lensless renderer (gfx950).
//
// GeRaFStage1/2 shift each ray's transmittance by where the array "sees" the
// scene (rf_rendering.py:455-495,611-626,749-750):
//   * intersect_sphere: rays from the centre c of the antennas through every
//     sample p (SDF frame) meet the sphere |x| = radius at t1 <= t2 of
//     t^2 + b t + cc = 0, b = 2 c.dir, cc = c.c - radius^2,
//     dir = (p - c) / (|p - c| + 1e-8); t = t1 if t1 > 1e-5, else t2 if
//     t2 > 1e-5, else none; valid = (b^2 - 4 cc >= 0) and a t was found;
//     hit = c + t dir.
//   * the offset: prev_cdf = sigmoid(inv_s sdf(first sample of the ray)),
//     new_cdf = sigmoid(inv_s sdf(hit)) where valid, prev_cdf elsewhere; the
//     renderer uses T + (new_cdf - prev_cdf).
// The SDF evaluations between the two kernels are the caller's network calls.
// Both kernels are elementwise and HBM-bound: one thread per sample, FP32
// arithmetic in the reference's operation order, one rounding per operation
// (no FMA contraction), so the sphere hits equal the reference formula's
// FP32 evaluation.
#include <hip/hip_runtime.h>

#include <math.h>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

constexpr int kCalibBlock = 256;

__global__ __launch_bounds__(kCalibBlock) void intersect_sphere_kernel(
    const float* __restrict__ center, const float* __restrict__ points, float* __restrict__ hit,
    uint8_t* __restrict__ valid, int64_t M, float radius) {
  // one rounding per operation, as the reference's separate tensor ops: no
  // contraction into FMAs anywhere in this kernel
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * kCalibBlock + threadIdx.x;
  if (i >= M) return;
  const float cx = center[0], cy = center[1], cz = center[2];
  const float vx = points[3 * i] - cx, vy = points[3 * i + 1] - cy, vz = points[3 * i + 2] - cz;
  const float len = sqrtf((vx * vx + vy * vy) + vz * vz);
  const float den = len + 1e-8f;
  const float dx = vx / den, dy = vy / den, dz = vz / den;
  const float b = 2.0f * ((cx * dx + cy * dy) + cz * dz);
  const float cc = ((cx * cx + cy * cy) + cz * cz) - radius * radius;
  const float disc = b * b - 4.0f * cc;
  const float sq = sqrtf(fmaxf(disc, 0.0f));
  const float t1 = (-b - sq) * 0.5f;
  const float t2 = (-b + sq) * 0.5f;
  const bool has_t = t1 > 1e-5f || t2 > 1e-5f;
  const float t = t1 > 1e-5f ? t1 : t2;
  const bool ok = disc >= 0.0f && has_t;
  // an invalid sample's hit is never used; it gets c (t = 0) instead of the
  // reference's c + inf * dir, so a network evaluated on every hit stays finite
  const float tt = has_t ? t : 0.0f;
  hit[3 * i] = cx + tt * dx;
  hit[3 * i + 1] = cy + tt * dy;
  hit[3 * i + 2] = cz + tt * dz;
  valid[i] = ok ? 1 : 0;
}

__device__ __forceinline__ float sigmoidf_ref(float x) { return 1.0f / (1.0f + expf(-x)); }

__global__ __launch_bounds__(kCalibBlock) void calibration_offset_kernel(
    const float* __restrict__ sdf_hit, const float* __restrict__ sdf_first,
    const uint8_t* __restrict__ valid, const float* __restrict__ inv_s_ptr,
    float* __restrict__ offset, int64_t R, int64_t S) {
  const int64_t i = (int64_t)blockIdx.x * kCalibBlock + threadIdx.x;
  if (i >= R * S) return;
  const float inv_s = inv_s_ptr[0];
  const float prev = sigmoidf_ref(sdf_first[i / S] * inv_s);
  const float cdf = valid[i] ? sigmoidf_ref(sdf_hit[i] * inv_s) : prev;
  offset[i] = cdf - prev;
}

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int geraf_intersect_sphere_f32(const float* center, const float* points, float* hit,
                                          uint8_t* valid, int64_t M, float radius,
                                          geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0) throw ArgError("sizes must be non-negative");
  if (!std::isfinite(radius) || radius <= 0.0f) throw ArgError("radius must be positive");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({center, points, hit, valid}, true);
  const int64_t blocks = (M + kCalibBlock - 1) / kCalibBlock;
  if (blocks > INT32_MAX) throw ArgError("too many points");
  hipLaunchKernelGGL(intersect_sphere_kernel, dim3((unsigned)blocks), dim3(kCalibBlock), 0,
                     (hipStream_t)stream, center, points, hit, valid, M, radius);
  return check_launch("intersect_sphere_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_calibration_offset_f32(const float* sdf_hit, const float* sdf_first,
                                            const uint8_t* valid, const float* inv_s,
                                            float* offset, int64_t R, int64_t S,
                                            geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (R < 0 || S < 0) throw ArgError("sizes must be non-negative");
  if (R * S == 0) return GERAF_RF_OK;
  require_ptrs({sdf_hit, sdf_first, valid, inv_s, offset}, true);
  const int64_t blocks = (R * S + kCalibBlock - 1) / kCalibBlock;
  if (blocks > INT32_MAX) throw ArgError("too many samples");
  hipLaunchKernelGGL(calibration_offset_kernel, dim3((unsigned)blocks), dim3(kCalibBlock), 0,
                     (hipStream_t)stream, sdf_hit, sdf_first, valid, inv_s, offset, R, S);
  return check_launch("calibration_offset_kernel");
  GERAF_TRY_END
}
