// Range-bin ("FFT") matched filter for gfx950: the nearest-range-bin
// backprojection of range-compressed echoes and its adjoint.
//
// Forward (MatchedFilter_cuda_kernel.cu:117-169), with d = |tx-v| + |rx-v| + bias:
//   skip if d < 0 or d > spacing * R;  bin = floor(d / spacing / 2);  skip if bin >= R
//   out[v] = (1/A) sum_a s[a, bin] * exp(-j*2*pi*f0*d/c)
// Adjoint (MatchedFilter_cuda_kernel.cu:222-279):
//   g[a, b] = (1/norm) sum_{v: bin(a,v) == b} gout[v] * exp(+j*2*pi*f0*d/c)
//
// The distance, bin and carrier phase are formed in FP64 (the carrier is
// reduced modulo one cycle before the FP32 sincospi).  The forward walks the
// antennas per voxel lane (wave-uniform positions) and gathers the echo bin.
// The adjoint gives each antenna one block that bins its voxels' contributions
// into FP64 accumulators in LDS, instead of the reference's loop over every
// voxel for every (antenna, bin) pair.
#include <hip/hip_runtime.h>

#include <math.h>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

struct BinConsts {
  double f0, inv_c, bias, spacing, max_d;
  int bins;
};

// Bin of distance d, or -1 when the reference would skip the pair.
__device__ __forceinline__ int range_bin(double d, const BinConsts& bc) {
  if (d < 0.0 || d > bc.max_d) return -1;
  const double b = floor(d / bc.spacing / 2.0);
  return (b < 0.0 || b >= (double)bc.bins) ? -1 : (int)b;
}

template <typename T>
__global__ __launch_bounds__(256) void mf_fft_forward_kernel(
    const T* __restrict__ real, const T* __restrict__ imag, const T* __restrict__ tx_pos,
    const T* __restrict__ rx_pos, T* __restrict__ out_real, T* __restrict__ out_imag,
    const T* __restrict__ voxel, int A, int N, BinConsts bc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
  double ar = 0.0, ai = 0.0;
  for (int a = 0; a < A; ++a) {
    const double d = dist3_f64(tx_pos[3 * a], tx_pos[3 * a + 1], tx_pos[3 * a + 2], px, py, pz) +
                     dist3_f64(rx_pos[3 * a], rx_pos[3 * a + 1], rx_pos[3 * a + 2], px, py, pz) +
                     bc.bias;
    const int b = range_bin(d, bc);
    if (b < 0) continue;
    double s, c;
    sincospi(2.0 * centered_frac(bc.f0 * d * bc.inv_c), &s, &c);
    const double sr = real[(size_t)a * bc.bins + b], si = imag[(size_t)a * bc.bins + b];
    ar += c * sr + s * si;  // s * exp(-j phi)
    ai += c * si - s * sr;
  }
  out_real[v] = (T)(ar / (double)A);
  out_imag[v] = (T)(ai / (double)A);
}

template <typename T>
__global__ __launch_bounds__(256) void mf_fft_backward_kernel(
    const T* __restrict__ gout_real, const T* __restrict__ gout_imag,
    const T* __restrict__ tx_pos, const T* __restrict__ rx_pos, T* __restrict__ grad_real,
    T* __restrict__ grad_imag, const T* __restrict__ voxel, int N, double inv_norm,
    BinConsts bc) {
  extern __shared__ double acc[];  // [bins][2]
  const int a = blockIdx.x;
  for (int i = threadIdx.x; i < 2 * bc.bins; i += blockDim.x) acc[i] = 0.0;
  __syncthreads();
  const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
  const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
  for (int v = threadIdx.x; v < N; v += blockDim.x) {
    const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
    const double d = dist3_f64(tx, ty, tz, px, py, pz) + dist3_f64(rx, ry, rz, px, py, pz) +
                     bc.bias;
    const int b = range_bin(d, bc);
    if (b < 0) continue;
    double s, c;
    sincospi(2.0 * centered_frac(bc.f0 * d * bc.inv_c), &s, &c);
    const double gr = gout_real[v], gi = gout_imag[v];
    atomicAdd(&acc[2 * b], c * gr - s * gi);  // gout * exp(+j phi)
    atomicAdd(&acc[2 * b + 1], s * gr + c * gi);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < bc.bins; b += blockDim.x) {
    grad_real[(size_t)a * bc.bins + b] = (T)(acc[2 * b] * inv_norm);
    grad_imag[(size_t)a * bc.bins + b] = (T)(acc[2 * b + 1] * inv_norm);
  }
}

BinConsts bin_consts(const geraf_fmcw_params* p, int64_t bins, double spacing) {
  check_params(p);
  if (!(spacing > 0.0) || !std::isfinite(spacing)) throw ArgError("fft_spacing must be > 0");
  BinConsts bc;
  bc.f0 = p->f0;
  bc.inv_c = 1.0 / p->c;
  bc.bias = p->path_bias;
  bc.spacing = spacing;
  // the reference compares with the FP32 product fft_spacing * num_range_bins
  bc.max_d = (double)((float)spacing * (float)bins);
  bc.bins = (int)bins;
  return bc;
}

template <typename T>
int fft_forward(const T* real, const T* imag, const T* tx, const T* rx, T* out_r, T* out_i,
                const T* voxel, int64_t A, int64_t bins, int64_t N, double spacing,
                const geraf_fmcw_params* p, geraf_stream_t stream) {
  check_mf_sizes(A, bins, N);
  const BinConsts bc = bin_consts(p, bins, spacing);
  if (N == 0) return GERAF_RF_OK;
  require_ptrs({out_r, out_i, voxel}, true);
  require_ptrs({real, imag, tx, rx}, A > 0);
  hipLaunchKernelGGL(mf_fft_forward_kernel<T>, dim3((unsigned)((N + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, real, imag, tx, rx, out_r, out_i, voxel, (int)A,
                     (int)N, bc);
  return check_launch("mf_fft_forward_kernel");
}

template <typename T>
int fft_backward(const T* gr, const T* gi, const T* tx, const T* rx, T* grad_r, T* grad_i,
                 const T* voxel, int64_t A, int64_t bins, int64_t N, int64_t norm,
                 double spacing, const geraf_fmcw_params* p, geraf_stream_t stream) {
  check_mf_sizes(A, bins, N);
  if (bins > 8192) throw ArgError("num_range_bins must be <= 8192");
  const BinConsts bc = bin_consts(p, bins, spacing);
  if (A == 0 || bins == 0) return GERAF_RF_OK;
  require_ptrs({tx, rx, grad_r, grad_i}, true);
  require_ptrs({gr, gi, voxel}, N > 0);
  hipLaunchKernelGGL(mf_fft_backward_kernel<T>, dim3((unsigned)A), dim3(256),
                     (size_t)bins * 2 * sizeof(double), (hipStream_t)stream, gr, gi, tx, rx,
                     grad_r, grad_i, voxel, (int)N, 1.0 / (double)norm, bc);
  return check_launch("mf_fft_backward_kernel");
}

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int geraf_mf_fft_forward_f32(const float* real, const float* imag, const float* tx_pos,
                                        const float* rx_pos, float* out_real, float* out_imag,
                                        const float* voxel, int64_t A, int64_t num_range_bins,
                                        int64_t N, double fft_spacing,
                                        const geraf_fmcw_params* params,
                                        geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return fft_forward(real, imag, tx_pos, rx_pos, out_real, out_imag, voxel, A, num_range_bins,
                     N, fft_spacing, params, stream);
  GERAF_TRY_END
}

extern "C" int geraf_mf_fft_forward_f64(const double* real, const double* imag,
                                        const double* tx_pos, const double* rx_pos,
                                        double* out_real, double* out_imag, const double* voxel,
                                        int64_t A, int64_t num_range_bins, int64_t N,
                                        double fft_spacing, const geraf_fmcw_params* params,
                                        geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return fft_forward(real, imag, tx_pos, rx_pos, out_real, out_imag, voxel, A, num_range_bins,
                     N, fft_spacing, params, stream);
  GERAF_TRY_END
}

extern "C" int geraf_mf_fft_backward_f32(const float* grad_out_real, const float* grad_out_imag,
                                         const float* tx_pos, const float* rx_pos,
                                         float* grad_real, float* grad_imag, const float* voxel,
                                         int64_t A, int64_t num_range_bins, int64_t N,
                                         int64_t norm_count, double fft_spacing,
                                         const geraf_fmcw_params* params,
                                         geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return fft_backward(grad_out_real, grad_out_imag, tx_pos, rx_pos, grad_real, grad_imag, voxel,
                      A, num_range_bins, N, norm_count, fft_spacing, params, stream);
  GERAF_TRY_END
}

extern "C" int geraf_mf_fft_backward_f64(const double* grad_out_real,
                                         const double* grad_out_imag, const double* tx_pos,
                                         const double* rx_pos, double* grad_real,
                                         double* grad_imag, const double* voxel, int64_t A,
                                         int64_t num_range_bins, int64_t N, int64_t norm_count,
                                         double fft_spacing, const geraf_fmcw_params* params,
                                         geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return fft_backward(grad_out_real, grad_out_imag, tx_pos, rx_pos, grad_real, grad_imag, voxel,
                      A, num_range_bins, N, norm_count, fft_spacing, params, stream);
  GERAF_TRY_END
}
