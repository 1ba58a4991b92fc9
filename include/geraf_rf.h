/*
 * geraf_rf.h - C ABI of libgeraf_rf.so, the MI355X (gfx950) implementation of
 * GeRaF's two RF hot paths: the FMCW matched filter (RFMatchedFilter) and the
 * differentiable signal tracer (RFSignalTracer).
 *
 * Conventions shared by every entry point
 *   - All pointers are device pointers (HBM) unless stated otherwise; the
 *     library never allocates, frees or synchronises caller memory.
 *   - Work is enqueued on `stream` (a hipStream_t; NULL = the legacy default
 *     stream) and returns without waiting for it.
 *   - Return value: GERAF_RF_OK, or an error code; geraf_rf_last_error()
 *     then gives a message (thread-local, valid until the next call).
 *   - Layouts follow the reference's pybind `_C` functions: planar real/imag
 *     arrays, antenna positions [A,3] row-major, voxels [3,N] (SoA).
 *
 * Phase model (reference: MatchedFilter_cuda_kernel.cu:28-63,
 * SignalTracer_cuda_kernel.cu:433-443):
 *     tau   = (|tx - p| + |rx - p| + path_bias) / c
 *     phi_k = 2*pi*(f0 + k*slope/fs)*tau
 * in the reference's scaled units (f0 in 1e8 Hz, slope = As_sci in 1e8 Hz/s,
 * fs in Hz, c in 1e8 m/s, distances in m).
 */
#ifndef GERAF_RF_H_
#define GERAF_RF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* geraf_stream_t; /* == hipStream_t */

enum {
  GERAF_RF_OK = 0,
  GERAF_RF_EINVAL = 1, /* bad size / pointer / parameter */
  GERAF_RF_EHIP = 2    /* a HIP launch or runtime call failed */
};

typedef struct geraf_fmcw_params {
  double f0;        /* chirp start frequency [1e8 Hz]. MF reference: 773.704 (hard-coded) */
  double slope;     /* chirp slope As_sci [1e8 Hz/s] */
  double fs;        /* ADC sample rate [Hz] */
  double c;         /* speed of light [1e8 m/s]. Reference: 2.9979 */
  double path_bias; /* metres added to the round-trip distance in the phase. Reference: 0.15 */
} geraf_fmcw_params;

/* Library identification (replaces get_compiler_version / get_cuda_version /
 * has_cuda of RFMatchedFilter/csrc/vision.cpp:70-79). Host-only, no GPU use. */
const char* geraf_rf_version(void);
const char* geraf_rf_last_error(void);

/* ---------------------------------------------------------------- matched filter */

/* Matched-filter backprojection, tuned gfx950 kernel.
 *   out[v] = (1/A) * sum_a sum_k (real + j*imag)[a,k] * exp(-j*phi_{a,k}(v))
 * FP32 bulk (two-level Horner recurrence along k, re-seeded every 32 samples),
 * FP64 distance and carrier-phase reduction per (voxel, antenna) pair, Kahan
 * summation across antennas.
 * Replaces _C.MatchedFilter_forward (RFMatchedFilter/csrc/vision.cpp:81-84;
 * MatchedFilter_cuda_kernel.cu:65-115). real/imag [A,K]; tx/rx [A,3];
 * voxel [3,N]; out_real/out_imag [N]. */
int geraf_mf_forward_f32(const float* real, const float* imag, const float* tx_pos,
                         const float* rx_pos, float* out_real, float* out_imag,
                         const float* voxel, int64_t A, int64_t K, int64_t N,
                         const geraf_fmcw_params* params, geraf_stream_t stream);

/* Same operator, plain FP32 form: one thread per voxel, per-term sincosf in
 * FP32 exactly as MatchedFilter_cuda_kernel.cu:87-111 computes it. Kept as the
 * reference-formula baseline; not the production path. */
int geraf_mf_forward_plain_f32(const float* real, const float* imag, const float* tx_pos,
                               const float* rx_pos, float* out_real, float* out_imag,
                               const float* voxel, int64_t A, int64_t K, int64_t N,
                               const geraf_fmcw_params* params, geraf_stream_t stream);

/* Same operator, plain FP64 form: one thread per voxel, per-term FP64 phase
 * and sincos. The GPU FP64 reference for full volumes; also serves float64
 * tensors in the drop-in. */
int geraf_mf_forward_f64(const double* real, const double* imag, const double* tx_pos,
                         const double* rx_pos, double* out_real, double* out_imag,
                         const double* voxel, int64_t A, int64_t K, int64_t N,
                         const geraf_fmcw_params* params, geraf_stream_t stream);

/* Adjoint of the matched filter (gradient into the echoes).
 *   grad[a,k] = (1/A) * sum_v (grad_out_real + j*grad_out_imag)[v] * exp(+j*phi_{a,k}(v))
 * written as grad_real/grad_imag [A,K]. `norm_count` is the divisor (the
 * reference divides by its antenna argument; the selected-rows variant of
 * faster_matched_filter_cuda.py:229-238 divides by the total count).
 * Replaces _C.MatchedFilter_backward_realimag (vision.cpp:89-92;
 * MatchedFilter_cuda_kernel.cu:171-220). */
int geraf_mf_backward_f32(const float* grad_out_real, const float* grad_out_imag,
                          const float* tx_pos, const float* rx_pos, float* grad_real,
                          float* grad_imag, const float* voxel, int64_t A, int64_t K,
                          int64_t N, int64_t norm_count, const geraf_fmcw_params* params,
                          geraf_stream_t stream);

int geraf_mf_backward_f64(const double* grad_out_real, const double* grad_out_imag,
                          const double* tx_pos, const double* rx_pos, double* grad_real,
                          double* grad_imag, const double* voxel, int64_t A, int64_t K,
                          int64_t N, int64_t norm_count, const geraf_fmcw_params* params,
                          geraf_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GERAF_RF_H_ */
