// VALU issue-rate probe for gfx950 (timing only): FP32 FMA streams with
// all-VGPR operands and with one SGPR operand, at 1-8 waves per SIMD and 1-4
// independent chains per lane.  It measures the issue ceiling that the
// matched-filter Horner loop (4 FMAs per term, half of them reading the echo
// from an SGPR) runs against.  Build: make -C geraf-sens_amd/csrc valu-probe
// (-> geraf-sens_amd/_bin/valu_issue_probe); run the binary on the GPU.
//
// Every thread writes exactly out[gid]; the largest grid equals the buffer.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <vector>

#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); return 1; } } while (0)

template <int ILP>
__global__ __launch_bounds__(256) void k_vgpr(float* out, int iters, float b0, float c0) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  float a[ILP];
  float b = b0 + 1e-7f * threadIdx.x, c = c0 - 1e-7f * threadIdx.x;
#pragma unroll
  for (int i = 0; i < ILP; ++i) a[i] = 1.0f + 1e-6f * (gid + i);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 32; ++r) {
#pragma unroll
      for (int i = 0; i < ILP; ++i) a[i] = fmaf(a[i], b, c);
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < ILP; ++i) s += a[i];
  out[gid] = s;
}

template <int ILP>
__global__ __launch_bounds__(256) void k_sgpr(float* out, int iters, float b0, float c0) {
  const int gid = blockIdx.x * blockDim.x + threadIdx.x;
  float a[ILP];
  float b = b0 + 1e-7f * threadIdx.x;
#pragma unroll
  for (int i = 0; i < ILP; ++i) a[i] = 1.0f + 1e-6f * (gid + i);
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int r = 0; r < 32; ++r) {
#pragma unroll
      for (int i = 0; i < ILP; ++i) a[i] = fmaf(a[i], b, c0);  // c0 stays in an SGPR
    }
  }
  float s = 0.0f;
#pragma unroll
  for (int i = 0; i < ILP; ++i) s += a[i];
  out[gid] = s;
}

int main() {
  const int blocks_per_cu[] = {1, 2, 4, 8};
  const int cus = 256, threads = 256, iters = 4096;
  float* out;
  const size_t max_threads = (size_t)cus * 8 * threads;
  CHECK(hipMalloc(&out, max_threads * sizeof(float)));
  hipEvent_t e0, e1;
  CHECK(hipEventCreate(&e0));
  CHECK(hipEventCreate(&e1));
  auto run = [&](const char* name, void (*kern)(float*, int, float, float), int ilp, int bpc) -> int {
    const int blocks = cus * bpc;  // blocks*threads <= max_threads by construction
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), 0, 0, out, 16, 0.999f, 1e-3f);
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0, 0));
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), 0, 0, out, iters, 0.999f, 1e-3f);
    CHECK(hipEventRecord(e1, 0));
    CHECK(hipEventSynchronize(e1));
    float ms;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    const double fmas = (double)blocks * threads * iters * 32.0 * ilp;
    printf("%-6s ILP=%d blocks/CU=%d (waves/SIMD=%d): %.3f ms  %.1f TFLOP/s  %.1f lane-FMA/clk/CU @2.4GHz\n",
           name, ilp, bpc, bpc, ms, 2 * fmas / ms / 1e9, fmas / (ms * 1e-3) / 2.4e9 / cus);
    return 0;
  };
  for (int b : blocks_per_cu) {
    if (run("vgpr", k_vgpr<1>, 1, b)) return 1;
    if (run("vgpr", k_vgpr<2>, 2, b)) return 1;
    if (run("vgpr", k_vgpr<4>, 4, b)) return 1;
    if (run("sgpr", k_sgpr<1>, 1, b)) return 1;
    if (run("sgpr", k_sgpr<2>, 2, b)) return 1;
    if (run("sgpr", k_sgpr<4>, 4, b)) return 1;
  }
  CHECK(hipFree(out));
  return 0;
}
