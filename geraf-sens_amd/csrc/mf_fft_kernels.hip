// Range-bin ("FFT") matched filter for gfx950: the nearest-range-bin
// backprojection of range-compressed echoes and its adjoint.
//
// Forward (MatchedFilter_cuda_kernel.cu:117-169), with d = |tx-v| + |rx-v| + bias:
//   skip if d < 0 or d > spacing * R;  bin = floor(d / spacing / 2);  skip if bin >= R
//   out[v] = (1/A) sum_a s[a, bin] * exp(-j*2*pi*f0*d/c)
// Adjoint (MatchedFilter_cuda_kernel.cu:222-279):
//   g[a, b] = (1/norm) sum_{v: bin(a,v) == b} gout[v] * exp(+j*2*pi*f0*d/c)
//
// The distance, bin and carrier-phase reduction are FP64 in both forms.  The
// FP32 form takes the phasor from the hardware sine and cosine of the
// FP64-reduced phase (about 1.3e-7 absolute, as the other FP32 kernels); the
// FP64 form uses the FP64 sincospi.  The forward walks the antennas per voxel
// lane (wave-uniform positions) and gathers the echo bin, summing in FP64.
// The adjoint gives each (group of up to 8 antennas, chunk of up to
// kChunkBins bins) one block that bins its voxels' contributions in LDS (one
// histogram per antenna, as many as fit in 64 KB), instead of the
// reference's loop over every voxel for every (antenna, bin) pair.  In the FP32 form the LDS
// accumulators are 64-bit fixed point with a power-of-two quantum chosen
// from the block's largest output-gradient component, so that no sum can
// overflow (|term| <= 2 max|g|, at most N terms, quantum ~4 N max|g| 2^-62):
// integer additions are exact, so the result does not depend on the order in
// which the lanes' atomic adds land (deterministic), and the quantisation is
// ~1e-12 of the largest gradient for N ~ 2^21.  The FP64 form (the
// validation form) adds FP64 values with LDS atomics (order-dependent at the
// 1e-16 level), as does the FP32 form for a non-finite output gradient.
#include <hip/hip_runtime.h>

#include <float.h>
#include <math.h>

#include <algorithm>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

constexpr int kFftThreads = 256;
constexpr int kChunkBins = 2048;  // adjoint LDS per block: 2048 x 2 x 8 B = 32 KB

struct BinConsts {
  double f0, inv_c, bias, spacing, max_d, inv_2spacing;
  int bins;
};

// Bin of distance d, or -1 when the reference would skip the pair.  The FP64
// form divides as the reference does; the FP32 form multiplies by the FP64
// reciprocal of 2 spacing (the two differ only when d / (2 spacing) lies
// within an FP64 ulp of an integer, far inside the reference's own FP32
// rounding of the bin).
template <typename T>
__device__ __forceinline__ int range_bin(double d, const BinConsts& bc) {
  if (d < 0.0 || d > bc.max_d) return -1;
  const double b = sizeof(T) == 8 ? floor(d / bc.spacing / 2.0) : floor(d * bc.inv_2spacing);
  return (b < 0.0 || b >= (double)bc.bins) ? -1 : (int)b;
}

// Round-trip distance plus bias, and the carrier phasor (cos, sin) of
// 2 pi f0 d / c: FP64 sincospi for the FP64 form, the hardware sine and cosine
// of the FP64-reduced phase for the FP32 form.
template <typename T>
__device__ __forceinline__ double fft_distance(double tx, double ty, double tz, double rx,
                                               double ry, double rz, double px, double py,
                                               double pz, double bias) {
  if constexpr (sizeof(T) == 8)
    return dist3_f64(tx, ty, tz, px, py, pz) + dist3_f64(rx, ry, rz, px, py, pz) + bias;
  else
    return round_trip_f64<false>(tx, ty, tz, rx, ry, rz, px, py, pz) + bias;
}
template <typename T>
__device__ __forceinline__ void fft_phasor(double d, const BinConsts& bc, double* c, double* s) {
  const double cyc = centered_frac(bc.f0 * d * bc.inv_c);
  if constexpr (sizeof(T) == 8) {
    sincospi(2.0 * cyc, s, c);
  } else {
    const float2 p = phasor_cycles_hw((float)cyc);
    *c = p.x;
    *s = p.y;
  }
}

template <typename T>
__global__ __launch_bounds__(kFftThreads) void mf_fft_forward_kernel(
    const T* __restrict__ real, const T* __restrict__ imag, const T* __restrict__ tx_pos,
    const T* __restrict__ rx_pos, T* __restrict__ out_real, T* __restrict__ out_imag,
    const T* __restrict__ voxel, int A, int N, BinConsts bc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
  double ar = 0.0, ai = 0.0;
  for (int a = 0; a < A; ++a) {
    const double d =
        fft_distance<T>(tx_pos[3 * a], tx_pos[3 * a + 1], tx_pos[3 * a + 2], rx_pos[3 * a],
                        rx_pos[3 * a + 1], rx_pos[3 * a + 2], px, py, pz, bc.bias);
    const int b = range_bin<T>(d, bc);
    if (b < 0) continue;
    double s, c;
    fft_phasor<T>(d, bc, &c, &s);
    const double sr = real[(size_t)a * bc.bins + b], si = imag[(size_t)a * bc.bins + b];
    ar += c * sr + s * si;  // s * exp(-j phi)
    ai += c * si - s * sr;
  }
  out_real[v] = (T)(ar / (double)A);
  out_imag[v] = (T)(ai / (double)A);
}

// Maximum over the block (kFftThreads threads), returned to every thread.
__device__ __forceinline__ float block_max(float m, float* red) {
  m = wave_max_dpp(m);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & (kWave - 1)) == 0) red[wave] = m;
  __syncthreads();
  float r = red[0];
#pragma unroll
  for (int w = 1; w < kFftThreads / kWave; ++w) r = fmaxf(r, red[w]);
  return r;
}

// AB antennas per block (their histograms side by side in LDS), so that
// each voxel's coordinates and gradient are read once per AB antennas.
template <typename T, int AB>
__global__ __launch_bounds__(kFftThreads) void mf_fft_backward_kernel(
    const T* __restrict__ gout_real, const T* __restrict__ gout_imag,
    const T* __restrict__ tx_pos, const T* __restrict__ rx_pos, T* __restrict__ grad_real,
    T* __restrict__ grad_imag, const T* __restrict__ voxel, int A, int N, double inv_norm,
    BinConsts bc) {
  // [AB][cap][2] accumulators, then the block reduction's four floats
  extern __shared__ double lds[];
  const int a0 = blockIdx.x * AB;
  const int na = min(AB, A - a0);
  const int b0 = blockIdx.y * kChunkBins;
  const int nb = min(kChunkBins, bc.bins - b0);
  const int cap = gridDim.y > 1 ? kChunkBins : bc.bins;  // accumulator pairs per antenna
  double* acc = lds;
  unsigned long long* iacc = reinterpret_cast<unsigned long long*>(lds);
  float* red = reinterpret_cast<float*>(lds + 2 * (size_t)AB * cap);
  // FP32 form: 64-bit fixed point with quantum 2^qe, from the largest
  // output-gradient component (block-uniform)
  bool fixed = false;
  int qe = 0;
  if constexpr (sizeof(T) == 4) {
    float m = 0.0f;
    for (int v = threadIdx.x; v < N; v += blockDim.x) {
      const float ar = fabsf((float)gout_real[v]), ai = fabsf((float)gout_imag[v]);
      // NaN and inf become inf (fmaxf alone would drop a NaN)
      m = fmaxf(m, (ar <= FLT_MAX && ai <= FLT_MAX) ? fmaxf(ar, ai) : INFINITY);
    }
    m = block_max(m, red);
    fixed = isfinite(m);  // NaN or inf: FP64 sums, as the FP64 form
    if (fixed && m > 0.0f) {
      int ex = 0;
      (void)frexp(4.0 * (double)m * (double)N, &ex);  // 4 m N < 2^ex
      qe = ex - 62;  // sums stay below 2^62 quanta
    }
  }
  for (int i = threadIdx.x; i < 2 * AB * cap; i += blockDim.x) acc[i] = 0.0;  // also iacc = 0
  __syncthreads();
  double tx[AB], ty[AB], tz[AB], rx[AB], ry[AB], rz[AB];
#pragma unroll
  for (int j = 0; j < AB; ++j) {
    const int a = a0 + min(j, na - 1);
    tx[j] = tx_pos[3 * a];
    ty[j] = tx_pos[3 * a + 1];
    tz[j] = tx_pos[3 * a + 2];
    rx[j] = rx_pos[3 * a];
    ry[j] = rx_pos[3 * a + 1];
    rz[j] = rx_pos[3 * a + 2];
  }
  for (int v = threadIdx.x; v < N; v += blockDim.x) {
    const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
    const double gr = gout_real[v], gi = gout_imag[v];
#pragma unroll
    for (int j = 0; j < AB; ++j) {
      if (j >= na) break;
      const double d = fft_distance<T>(tx[j], ty[j], tz[j], rx[j], ry[j], rz[j], px, py, pz,
                                       bc.bias);
      const int b = range_bin<T>(d, bc) - b0;
      if (b < 0 || b >= nb) continue;  // skipped pair, or a bin of another chunk
      double s, c;
      fft_phasor<T>(d, bc, &c, &s);
      const double tr = c * gr - s * gi, ti = s * gr + c * gi;  // gout * exp(+j phi)
      const int slot = 2 * (j * cap + b);
      if (fixed) {
        atomicAdd(&iacc[slot], (unsigned long long)__double2ll_rn(ldexp(tr, -qe)));
        atomicAdd(&iacc[slot + 1], (unsigned long long)__double2ll_rn(ldexp(ti, -qe)));
      } else {
        atomicAdd(&acc[slot], tr);
        atomicAdd(&acc[slot + 1], ti);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < na * nb; i += blockDim.x) {
    const int j = i / nb, b = i - j * nb;
    const int slot = 2 * (j * cap + b);
    double gr_b, gi_b;
    if (fixed) {
      gr_b = ldexp((double)(long long)iacc[slot], qe);
      gi_b = ldexp((double)(long long)iacc[slot + 1], qe);
    } else {
      gr_b = acc[slot];
      gi_b = acc[slot + 1];
    }
    const size_t o = (size_t)(a0 + j) * bc.bins + b0 + b;
    grad_real[o] = (T)(gr_b * inv_norm);
    grad_imag[o] = (T)(gi_b * inv_norm);
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
  bc.inv_2spacing = 1.0 / (2.0 * spacing);
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
  hipLaunchKernelGGL(mf_fft_forward_kernel<T>, dim3((unsigned)((N + kFftThreads - 1) / kFftThreads)),
                     dim3(kFftThreads), 0,
                     (hipStream_t)stream, real, imag, tx, rx, out_r, out_i, voxel, (int)A,
                     (int)N, bc);
  return check_launch("mf_fft_forward_kernel");
}

template <typename T>
int fft_backward(const T* gr, const T* gi, const T* tx, const T* rx, T* grad_r, T* grad_i,
                 const T* voxel, int64_t A, int64_t bins, int64_t N, int64_t norm,
                 double spacing, const geraf_fmcw_params* p, geraf_stream_t stream) {
  check_mf_sizes(A, bins, N);
  const BinConsts bc = bin_consts(p, bins, spacing);
  if (A == 0 || bins == 0) return GERAF_RF_OK;
  require_ptrs({tx, rx, grad_r, grad_i}, true);
  require_ptrs({gr, gi, voxel}, N > 0);
  const unsigned chunks = (unsigned)((bins + kChunkBins - 1) / kChunkBins);
  const size_t cap = chunks > 1 ? (size_t)kChunkBins : (size_t)bins;
  // antennas per block: as many histograms as fit in 64 KB of LDS (up to 8)
  const size_t per_antenna = cap * 2 * sizeof(double);
  const size_t budget = 65536 - (kFftThreads / kWave) * sizeof(float);
  int ab = 8;
  while (ab > 1 && (size_t)ab * per_antenna > budget) ab /= 2;
  const size_t lds = (size_t)ab * per_antenna + (kFftThreads / kWave) * sizeof(float);
  const dim3 grid((unsigned)((A + ab - 1) / ab), chunks);
#define GERAF_FFT_BWD(AB_)                                                                    \
  hipLaunchKernelGGL((mf_fft_backward_kernel<T, AB_>), grid, dim3(kFftThreads), lds,          \
                     (hipStream_t)stream, gr, gi, tx, rx, grad_r, grad_i, voxel, (int)A, (int)N, \
                     1.0 / (double)norm, bc)
  switch (ab) {
    case 8: GERAF_FFT_BWD(8); break;
    case 4: GERAF_FFT_BWD(4); break;
    case 2: GERAF_FFT_BWD(2); break;
    default: GERAF_FFT_BWD(1); break;
  }
#undef GERAF_FFT_BWD
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
