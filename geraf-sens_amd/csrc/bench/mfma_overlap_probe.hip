// MFMA / VALU overlap probe for gfx950 (timing only).  Each wave runs, per
// iteration, 8 independent matrix products (v_mfma_f32_16x16x4f32, or
// v_mfma_f32_16x16x32_f16) and/or 64 independent FP32 FMAs, in four modes:
// matrix only, vector only, both in every wave, and half of the waves each.
// If matrix and vector work overlapped, "both" would approach the larger of
// the two and "half each" half of it; without overlap both approach the sum.
// This decides how the RF sums on the matrix cores are costed.
// Build: make -C geraf-sens_amd/csrc mfma-probe (-> geraf-sens_amd/_bin/mfma_overlap_probe);
// run the binary on the GPU.
//
// Every thread writes exactly out[gid]; every grid equals the buffer.
#include <hip/hip_runtime.h>
#include <stdio.h>

#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

// MODE 0: matrix only; 1: vector only; 2: both in every wave; 3: even waves
// matrix, odd waves vector.  F16: the 16x16x32 FP16 product instead of the
// 16x16x4 FP32 one.
template <int MODE, bool F16>
__global__ __launch_bounds__(256) void probe(float* out, float a, float b, int iters) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  f32x4 acc[8];
  for (int i = 0; i < 8; ++i) acc[i] = f32x4{0.f, 0.f, 0.f, 0.f};
  float v[16];
  for (int i = 0; i < 16; ++i) v[i] = a + lane * 1e-3f + i;
  const bool do_m = MODE == 0 || MODE == 2 || (MODE == 3 && (wave & 1) == 0);
  const bool do_v = MODE == 1 || MODE == 2 || (MODE == 3 && (wave & 1) == 1);
  const float x = a + lane, y = b - lane;
  f16x8 xh, yh;
  for (int i = 0; i < 8; ++i) {
    xh[i] = (_Float16)(x + i);
    yh[i] = (_Float16)(y - i);
  }
  for (int it = 0; it < iters; ++it) {
    if (do_m) {
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        if constexpr (F16)
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x32_f16(xh, yh, acc[i], 0, 0, 0);
        else
          acc[i] = __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, acc[i], 0, 0, 0);
      }
    }
    if (do_v) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = fmaf(v[i], b, a);
    }
  }
  float s = 0.f;
  for (int i = 0; i < 8; ++i) s += acc[i][0] + acc[i][1] + acc[i][2] + acc[i][3];
  for (int i = 0; i < 16; ++i) s += v[i];
  out[blockIdx.x * 256 + threadIdx.x] = s;
}

template <int MODE, bool F16>
float run(float* d, int blocks, int iters) {
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipLaunchKernelGGL((probe<MODE, F16>), dim3(blocks), dim3(256), 0, 0, d, 0.5f, 0.999f, iters);
  hipEventRecord(e0);
  for (int r = 0; r < 5; ++r)
    hipLaunchKernelGGL((probe<MODE, F16>), dim3(blocks), dim3(256), 0, 0, d, 0.5f, 0.999f, iters);
  hipEventRecord(e1);
  hipEventSynchronize(e1);
  float ms = 0.f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return ms / 5;
}

template <bool F16>
void report(float* d, int blocks, int iters) {
  const float m = run<0, F16>(d, blocks, iters), v = run<1, F16>(d, blocks, iters);
  const float both = run<2, F16>(d, blocks, iters), half = run<3, F16>(d, blocks, iters);
  printf("%s: matrix only %.3f ms, vector only %.3f ms, both in every wave %.3f ms (sum %.3f), "
         "half the waves each %.3f ms (half the sum %.3f)\n",
         F16 ? "16x16x32 f16" : "16x16x4 f32", m, v, both, m + v, half, 0.5f * (m + v));
}

int main() {
  const int blocks = 2048, iters = 2048;  // 8 waves per SIMD resident on 256 CUs
  float* d;
  CHECK(hipMalloc(&d, (size_t)blocks * 256 * sizeof(float)));
  report<false>(d, blocks, iters);
  report<true>(d, blocks, iters);
  CHECK(hipDeviceSynchronize());
  CHECK(hipFree(d));
  return 0;
}
