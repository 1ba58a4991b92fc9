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

/* Same operator (_C.MatchedFilter_forward, vision.cpp:81-84;
 * MatchedFilter_cuda_kernel.cu:65-115) on the matrix cores, for
 * 1 <= K <= 2048 (faster than geraf_mf_forward_f32 from K ~ 300 on).  With
 * k = 16*alpha + beta, each antenna's sum is a complex matrix product over
 * beta on v_mfma_f32_16x16x32_f16 (the complex product folded into the
 * depth), every real operand split into FP16 hi + lo and each real product
 * taken as three (FP32-accurate products, FP32 accumulation); phases as in
 * geraf_mf_forward_f32 (FP64 distances and reduction).  workspace: device memory of workspace_bytes >=
 * geraf_mf_forward_mma_workspace_bytes(A, K), 256-byte aligned, owned by the
 * caller (the echo in operand order); nothing is allocated.  K > 2048 is an
 * argument error.  geraf_mf_forward_mma_workspace_bytes returns -1 for
 * negative sizes. */
int64_t geraf_mf_forward_mma_workspace_bytes(int64_t A, int64_t K);
int geraf_mf_forward_f32_mma(const float* real, const float* imag, const float* tx_pos,
                             const float* rx_pos, float* out_real, float* out_imag,
                             const float* voxel, int64_t A, int64_t K, int64_t N,
                             const geraf_fmcw_params* params, void* workspace,
                             int64_t workspace_bytes, geraf_stream_t stream);

/* Size of the device-resident scratch of geraf_mf_forward_f32_mma_noalloc
 * (48 MiB per GPU that loads the library, held until the process ends). */
#define GERAF_MF_MMA_SCRATCH_BYTES (48u << 20)

/* geraf_mf_forward_f32_mma without a workspace argument and without any
 * allocation: the echo's operand buffer is a static device array of
 * GERAF_MF_MMA_SCRATCH_BYTES, so geraf_mf_forward_mma_workspace_bytes(A, K)
 * must not exceed that (4,096 A ceil(K / 256) + 4 A bytes, each term rounded
 * up to 256: A <= 6,141 at K <= 512, A <= 2,457 at K = 1,025..1,280); larger
 * sizes return GERAF_RF_EINVAL with a message.  Same
 * kernels, same results.  The scratch is shared by every call on one GPU:
 * calls that overlap on one GPU (different streams or host threads) race.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). */
int geraf_mf_forward_f32_mma_noalloc(const float* real, const float* imag, const float* tx_pos,
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
 * MatchedFilter_cuda_kernel.cu:171-220). Deterministic (fixed-order sums). */
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

/* Range-bin ("FFT") matched filter on range-compressed echoes real/imag
 * [A, num_range_bins]: with d = |tx-v| + |rx-v| + path_bias, pairs with d < 0
 * or d > fft_spacing*num_range_bins are skipped; bin = floor(d/fft_spacing/2)
 * (skipped if >= num_range_bins);
 *   out[v] = (1/A) sum_a s[a,bin] * exp(-j*2*pi*f0*d/c)
 * (params->slope and params->fs are unused). Distance, bin and phase
 * reduction in FP64; the f32 form takes the phasor of the reduced phase from
 * the hardware sine and cosine (~1.3e-7), the f64 form from FP64 sincospi.
 * Replaces _C.MatchedFilter_fft_forward (vision.cpp:85-88;
 * MatchedFilter_cuda_kernel.cu:117-169). */
int geraf_mf_fft_forward_f32(const float* real, const float* imag, const float* tx_pos,
                             const float* rx_pos, float* out_real, float* out_imag,
                             const float* voxel, int64_t A, int64_t num_range_bins, int64_t N,
                             double fft_spacing, const geraf_fmcw_params* params,
                             geraf_stream_t stream);
int geraf_mf_fft_forward_f64(const double* real, const double* imag, const double* tx_pos,
                             const double* rx_pos, double* out_real, double* out_imag,
                             const double* voxel, int64_t A, int64_t num_range_bins, int64_t N,
                             double fft_spacing, const geraf_fmcw_params* params,
                             geraf_stream_t stream);

/* Adjoint of the range-bin matched filter into grad_real/grad_imag
 * [A, num_range_bins], divided by norm_count.  The f32 form sums each bin
 * exactly in 64-bit fixed point (deterministic: the result does not depend
 * on the order of the additions); the f64 form sums in FP64 in whatever order
 * the additions land (1e-16-level differences between runs).
 * Replaces _C.MatchedFilter_fft_backward_realimag (vision.cpp:93-96;
 * MatchedFilter_cuda_kernel.cu:222-279). */
int geraf_mf_fft_backward_f32(const float* grad_out_real, const float* grad_out_imag,
                              const float* tx_pos, const float* rx_pos, float* grad_real,
                              float* grad_imag, const float* voxel, int64_t A,
                              int64_t num_range_bins, int64_t N, int64_t norm_count,
                              double fft_spacing, const geraf_fmcw_params* params,
                              geraf_stream_t stream);
int geraf_mf_fft_backward_f64(const double* grad_out_real, const double* grad_out_imag,
                              const double* tx_pos, const double* rx_pos, double* grad_real,
                              double* grad_imag, const double* voxel, int64_t A,
                              int64_t num_range_bins, int64_t N, int64_t norm_count,
                              double fft_spacing, const geraf_fmcw_params* params,
                              geraf_stream_t stream);

/* ---------------------------------------------------------------- signal tracer */

/* [A,P] signal tracer forward:
 *   out[a,k] = (1/P) sum_{p: inbounds, spec >= 1e-6, inc.n <= 0}
 *              sigma[a,p] * spec_ap / d_ap^2 * exp(+j*phi_{a,k}(p))
 * with the specular geometry of SignalTracer_cuda_kernel.cu:172-197 (normals
 * used as given), d = |tx-p| + |rx-p| (no bias in the amplitude).
 * normals/points [P,3]; sigmas [A,P]; tx/rx [A,3]; inbounds [P] (bool, 1 byte);
 * out_real/out_imag [A,K]. params->f0 is the tracer's fc_start.
 * Replaces _C.SignalTracer_forward (RFSignalTracer/csrc/vision.cpp;
 * SignalTracer_cuda_kernel.cu:351-450). */
int geraf_st_forward_f32(const float* normals, const float* sigmas, const float* points,
                         const float* tx_pos, const float* rx_pos, const uint8_t* inbounds,
                         float* out_real, float* out_imag, int64_t A, int64_t P, int64_t K,
                         const geraf_fmcw_params* params, geraf_stream_t stream);
int geraf_st_forward_f64(const double* normals, const double* sigmas, const double* points,
                         const double* tx_pos, const double* rx_pos, const uint8_t* inbounds,
                         double* out_real, double* out_imag, int64_t A, int64_t P, int64_t K,
                         const geraf_fmcw_params* params, geraf_stream_t stream);

/* [A,P] signal tracer backward: gradients into normals [P,3] and sigmas [A,P]
 * from grad_out_real/imag [A,K]. Every output element is written; points with
 * inbounds == false get zeros (the reference leaves them uninitialised).
 * Replaces _C.SignalTracer_backward_normal_sigma
 * (SignalTracer_cuda_kernel.cu:452-587). */
int geraf_st_backward_f32(const float* grad_out_real, const float* grad_out_imag,
                          const float* normals, const float* sigmas, const float* points,
                          const float* tx_pos, const float* rx_pos, const uint8_t* inbounds,
                          float* grad_normals, float* grad_sigmas, int64_t A, int64_t P,
                          int64_t K, const geraf_fmcw_params* params, geraf_stream_t stream);
int geraf_st_backward_f64(const double* grad_out_real, const double* grad_out_imag,
                          const double* normals, const double* sigmas, const double* points,
                          const double* tx_pos, const double* rx_pos, const uint8_t* inbounds,
                          double* grad_normals, double* grad_sigmas, int64_t A, int64_t P,
                          int64_t K, const geraf_fmcw_params* params, geraf_stream_t stream);

/* Linear resampling of per-ray normals [R,S,3] and sigmas [A,R,S] from source
 * depths z_vals [R,S] to target depths tgt_z_vals [R,T] (binary search for the
 * bracket; targets outside [z_lo, z_hi] give 0 and inbounds = false).
 * Requires S >= 2. Replaces _C.BilinearRaysample_forward
 * (SignalTracer_cuda_kernel.cu:199-279). */
int geraf_bilinear_forward_f32(const float* normals, const float* sigmas, const float* z_vals,
                               const float* tgt_z_vals, float* out_normals, float* out_sigmas,
                               uint8_t* out_inbounds, int64_t A, int64_t R, int64_t S,
                               int64_t T, geraf_stream_t stream);
int geraf_bilinear_forward_f64(const double* normals, const double* sigmas,
                               const double* z_vals, const double* tgt_z_vals,
                               double* out_normals, double* out_sigmas, uint8_t* out_inbounds,
                               int64_t A, int64_t R, int64_t S, int64_t T,
                               geraf_stream_t stream);

/* Transpose of the resampler, written without atomics (deterministic); every
 * element of grad_normals [R,S,3] and grad_sigmas [A,R,S] is written.
 * Replaces _C.BilinearRaysample_backward (SignalTracer_cuda_kernel.cu:281-349). */
int geraf_bilinear_backward_f32(const float* grad_out_normals, const float* grad_out_sigmas,
                                const float* z_vals, const float* tgt_z_vals,
                                float* grad_normals, float* grad_sigmas, int64_t A, int64_t R,
                                int64_t S, int64_t T, geraf_stream_t stream);
int geraf_bilinear_backward_f64(const double* grad_out_normals, const double* grad_out_sigmas,
                                const double* z_vals, const double* tgt_z_vals,
                                double* grad_normals, double* grad_sigmas, int64_t A, int64_t R,
                                int64_t S, int64_t T, geraf_stream_t stream);

/* ---------------------------------------------------------------- fused lensless renderer */

/* Per-ray lensless renderer: NeuS-style compositing of SDF samples and the
 * monostatic FMCW signal of each ray, in one kernel. For ray r (antenna at
 * origin[r], unit direction dirs[r]) and sample s < n_valid[r]:
 *   alpha, T     as rf_rendering.py:46-52,145-169 (raw gradients in iter_cos)
 *   Tc = T + trans_offset; sigma = max(refl*power*Tc, 1e-6)*alpha*Tc
 *                  (rf_rendering.py:749-759, sdf_network.py:475-481)
 *   signal       SignalTracer_cuda_kernel.cu:351-450 with tx = rx = origin,
 *                normal = grad/max(|grad|,1e-12), point = origin + z*dir
 *   out[r,k] = (1/n_valid[r]) sum_s amp_s exp(+j*phi_{s,k})
 * sdf, dists, z, refl, trans_offset: [R,S]; grad: [R,S,3]; origin, dirs: [R,3];
 * n_valid: [R] int32; inv_s, power: device scalars [1]; out_real/out_imag
 * [R,K]; sigma_out [R,S] (optional); tile_carry [R, ceil(S/64)] (optional):
 * the transmittance at the start of each 64-sample tile, for the backward.
 * refl, trans_offset, n_valid, sigma_out, tile_carry may be NULL (1, 0, S,
 * not written, not written). Limits: S <= 2048, K <= 2048.
 * No counterpart in the reference: it fuses the torch glue of
 * rf_rendering.py:740-780 with the tracer op for the per-ray case. */
int geraf_lensless_forward_f32(const float* sdf, const float* grad, const float* refl,
                               const float* dists, const float* z, const float* origin,
                               const float* dirs, const float* trans_offset,
                               const int32_t* n_valid, const float* inv_s, const float* power,
                               float* out_real, float* out_imag, float* sigma_out,
                               float* tile_carry, int64_t R, int64_t S, int64_t K,
                               float cos_anneal_ratio, const geraf_fmcw_params* params,
                               geraf_stream_t stream);

/* Adjoint of geraf_lensless_forward_f32: gradients into sdf, grad (the raw
 * SDF gradients, through iter_cos and the normalised normal), refl, and per-ray
 * partial sums of the inv_s and power gradients ([R] each; the caller sums
 * them). Recomputes the compositing instead of reading stored intermediates.
 * grad_refl may be NULL. Samples past n_valid get zeros. tile_carry: the
 * forward's tile_carry output for the same inputs, or NULL; with it the
 * kernel walks the samples once (without it, twice: the first walk rebuilds
 * the carries); the results are the same bit for bit. Limits as the forward
 * (S <= 2048, K <= 2048). */
int geraf_lensless_backward_f32(const float* grad_out_real, const float* grad_out_imag,
                                const float* sdf, const float* grad, const float* refl,
                                const float* dists, const float* z, const float* origin,
                                const float* dirs, const float* trans_offset,
                                const int32_t* n_valid, const float* inv_s, const float* power,
                                float* grad_sdf, float* grad_grad, float* grad_refl,
                                float* grad_inv_s_per_ray, float* grad_power_per_ray,
                                const float* tile_carry, int64_t R, int64_t S, int64_t K,
                                float cos_anneal_ratio, const geraf_fmcw_params* params,
                                geraf_stream_t stream);

/* Sphere hits of the transmittance calibration (GeRaFStage1.intersect_sphere,
 * rf_rendering.py:455-495): for each of M points p [M,3] (SDF frame) and the
 * antennas' centre c (center: device pointer to 3 floats, e.g. the mean of the
 * normalised receiver positions), dir = (p - c)/(|p - c| + 1e-8); the ray
 * c + t*dir meets |x| = radius at t1 <= t2; t = t1 if t1 > 1e-5, else t2 if
 * t2 > 1e-5, else none.  valid[i] = 1 if the discriminant is >= 0 and a t was
 * found; hit [M,3] = c + t*dir (= c where no t was found, where the reference
 * has c + inf*dir; such hits are never used).  FP32, torch's operation order. */
int geraf_intersect_sphere_f32(const float* center, const float* points, float* hit,
                               uint8_t* valid, int64_t M, float radius, geraf_stream_t stream);

/* Transmittance offset of the calibration (_calibrate_transmission,
 * rf_rendering.py:611-626, used at :749-750): with prev = sigmoid(inv_s *
 * sdf_first[r]) and new = valid[r,s] ? sigmoid(inv_s * sdf_hit[r,s]) : prev,
 * offset[r,s] = new - prev (the fused renderer's trans_offset).  sdf_hit,
 * valid, offset [R,S]; sdf_first [R] (the SDF at each ray's first sample);
 * inv_s: device scalar [1]. */
int geraf_calibration_offset_f32(const float* sdf_hit, const float* sdf_first,
                                 const uint8_t* valid, const float* inv_s, float* offset,
                                 int64_t R, int64_t S, geraf_stream_t stream);

/* Fast matched filter: the operator of geraf_mf_forward_f32 (and of
 * _C.MatchedFilter_forward, vision.cpp:70-74; MatchedFilter_cuda_kernel.cu:65-115)
 * evaluated by per-tile expansions in the delay offset.  It has no counterpart
 * in the reference.  voxel [3, N] must be ordered so that every tile of
 * tile_voxels (128 or 512) consecutive voxels is compact in space; tile_info
 * [ceil(N/tile_voxels), 4] holds each tile's centre (x, y, z) and radius r
 * (every voxel of the tile within r of the centre).  With
 * Z = pi*(K-1)*(slope/fs)*2*r_max/c, nterms = N_T bounds the truncation by
 * sum_k |s[a,k]| * Z^N_T / N_T! * e^Z per antenna (rounded up to one of 4, 6,
 * 8, 10, 12, 14, 16, 20, 24, 32; at most 32).  workspace: device memory of
 * workspace_bytes >= geraf_mf_tiled_workspace_bytes for A=1 (the antennas are
 * then processed in chunks that fit); the caller owns it, nothing is
 * allocated.  geraf_mf_tiled_workspace_bytes returns -1 for invalid sizes. */
int64_t geraf_mf_tiled_workspace_bytes(int64_t A, int64_t K, int64_t N, int32_t nterms,
                                       int32_t tile_voxels);
int geraf_mf_forward_tiled_f32(const float* real, const float* imag, const float* tx_pos,
                               const float* rx_pos, float* out_real, float* out_imag,
                               const float* voxel, const float* tile_info, int64_t A, int64_t K,
                               int64_t N, int32_t nterms, int32_t tile_voxels, void* workspace,
                               int64_t workspace_bytes, const geraf_fmcw_params* params,
                               geraf_stream_t stream);
/* Adjoint of geraf_mf_forward_tiled_f32 by the same expansion (the operator
 * of geraf_mf_backward_f32, i.e. _C.MatchedFilter_backward_realimag,
 * vision.cpp:75-80; MatchedFilter_cuda_kernel.cu:171-220), with the same
 * truncation bound; no counterpart in the reference.  grad_out_real/imag [N]
 * in the tile order of voxel; grad [A, K] divided by norm_count. */
int geraf_mf_backward_tiled_f32(const float* grad_out_real, const float* grad_out_imag,
                                const float* tx_pos, const float* rx_pos, float* grad_real,
                                float* grad_imag, const float* voxel, const float* tile_info,
                                int64_t A, int64_t K, int64_t N, int64_t norm_count,
                                int32_t nterms, int32_t tile_voxels, void* workspace,
                                int64_t workspace_bytes, const geraf_fmcw_params* params,
                                geraf_stream_t stream);

/* ---------------------------------------------------------------- MLP activation (BF16) */

/* Softplus of the SDF network's BF16 hidden layers and its derivatives
 * (torch.nn.Softplus(beta, threshold) semantics; the network uses beta = 100,
 * threshold = 20, sdf_network.py:78-199). x, y, g, ... are n raw BF16 values
 * (16-bit patterns); arithmetic is FP32, results are rounded to nearest even.
 * With t = beta x and s = (t > threshold ? 1 : sigmoid(t)):
 *   geraf_softplus_bf16:       y = (t > threshold ? x : log1p(exp(t)) / beta)
 *   geraf_softplus_grad_bf16:  gx = g s
 *   geraf_softplus_grad2_bf16: out_g = ggx s,  out_x = ggx g beta s (1 - s)
 *                              (the adjoint of the second, for double backward)
 * No counterpart in the reference's C++: they replace torch's softplus
 * kernels (forward, backward and its double backward) for these layers. */
int geraf_softplus_bf16(const uint16_t* x, uint16_t* y, int64_t n, float beta, float threshold,
                        geraf_stream_t stream);
int geraf_softplus_grad_bf16(const uint16_t* g, const uint16_t* x, uint16_t* gx, int64_t n,
                             float beta, float threshold, geraf_stream_t stream);
int geraf_softplus_grad2_bf16(const uint16_t* ggx, const uint16_t* g, const uint16_t* x,
                              uint16_t* out_g, uint16_t* out_x, int64_t n, float beta,
                              float threshold, geraf_stream_t stream);

/* ---------------------------------------------------------------- SDF network products (BF16) */

/* The layer products of the SDF network (sdf_network.py:78-199) on the BF16
 * matrix cores, for the hand-written forward, normals and backward of
 * geraf.models.networks.fused_sdf.  No counterpart in the reference's C++:
 * they replace the network's torch layers (GEMMs and the element-wise work
 * around them).  BF16 values are raw 16-bit patterns; accumulation is FP32.
 *
 * geraf_mlp_pack_weights: W [N,K] FP32 (w[n*ldw + k], or w[k*ldw + n] when
 * transpose != 0) times scale, rounded to BF16, into the operand order the
 * products read (geraf_mlp_pack_bytes(N, K) bytes, 16-byte aligned).
 * N <= 272, K <= 288, and K <= 256 when N > 256 (the padded weights stay in
 * LDS); geraf_mlp_pack_bytes returns -1 outside that.
 *
 * Slope codes: where the network stores a softplus pre-activation z, the
 * training epilogue writes instead its FP16 code q = copysign(m, beta z),
 * m = sigmoid(-|beta z|) (from the FP32 accumulator); the slope is
 * s = 1 - |q| for q >= +0 and |q| for q <= -0 (sign bit), the curvature
 * s (1 - s) = |q| (1 - |q|).
 *
 * geraf_mlp_gemm_bf16: y[M,N] = epilogue(prologue(x)[M,K] . W^T + bias):
 *   prologue 0: x; 1: x * s(x2) (x2 the slope codes, element by element);
 *   epilogue (z = the product + bias, bias may be NULL):
 *     0: y = z;  1: y = softplus(z);  2: y = softplus(z), y2 = code(z) (FP16);
 *     3: y = relu(z);  4: y = z, y2 = z * e1 * beta * s(e0) (1 - s(e0));
 *     5: y = z * s(e0) + e1  (e0 slope codes, e1 BF16);
 *     6: y = z * s(e0);
 *     7: y = z * s(e0), y2 = z * e1 * beta * (1 - s(e0));
 *     8: y = z * s(e0), y2 = z * e1 * beta * s(e0) (1 - s(e0)).
 *   Softplus(beta, threshold) as torch's, evaluated as (max(t,0) +
 *   log1p(exp(-|t|))) / beta, t = beta z, which equals torch's linear branch
 *   to an FP32 ulp for t > threshold; threshold must be >= 15.  Outputs are
 *   written in groups of 8 columns: columns N .. 8 ceil(N/8) - 1 of y and y2
 *   receive unspecified values.  Leading dimensions are in elements: 0
 *   broadcasts one row (inputs only), otherwise a multiple of 8 covering the
 *   columns rounded up to 8.  All pointers 16-byte aligned.
 *
 * geraf_mlp_wgrad_bf16: dw[N,K] (+)= prologue(g)^T prologue(x) over the M
 * points (g [M,N], x [M,K], prologues as above; beta and threshold are
 * unused, the codes carry the slopes), FP32, row stride lddw;
 * db[N] (+)= the column sums of prologue(g) when db is not NULL; with
 * accumulate != 0 the results are added to dw / db.  N, K <= 512 (a
 * prologue needs N <= 256 or N > 272).  The
 * workspace (geraf_mlp_wgrad_workspace_bytes, caller-owned) holds per-block
 * partial sums, added in a fixed order (deterministic).
 *
 * geraf_mlp_positional_encoding: out[p*ldo + col0 + i] = gamma(x_p * scale)_i
 * (i < 3 + 6L: x, then per band l the 3 sines and the 3 cosines of
 * 2^l x, the Embedder order of sdf_network.py:19-75), BF16; zeros up to
 * pad_to columns (3 + 6L <= pad_to <= 64, L <= 10).
 * geraf_mlp_pe_gradient: n[p] = d sdf / d x from d sdf / d gamma = ga + gb
 * (two BF16 [M, 3+6L] parts), FP32 [M,3].
 * geraf_mlp_pe_gradient_adjoint: its transpose, d loss / d gamma = J dn
 * (BF16, into out at col0, zero-padded up to pad_to).
 *
 * geraf_mlp_sdf_chain_bf16: the SDF network's eight hidden layers in one
 * kernel (gamma with 5 <= L <= 10 bands, d = 3 + 6L features, of the points x
 * [M,3] times scale; 256-wide softplus layers, layer 3 of width 256 - d,
 * the skip [h_3, gamma] into layer 4), activations carried in registers.
 * wpacked holds the hidden layers' packed weights back to back, each as
 * geraf_mlp_pack_weights writes it (W_0 [256, d], W_1 .. W_7, W_3 [256 - d,
 * 256], W_4 times the skip's scale): geraf_mlp_sdf_chain_pack_bytes(L)
 * bytes.  bias is [8][256] FP32 (layer 3's past its width unused).  With
 * sdf not NULL: sdf[p] = bf16(w) . h_7 + b in FP32, w_sdf = [w, b] (257
 * FP32: the output layer's sdf row and bias).  With h and q not NULL
 * (training):
 * h_l (BF16) and the slope codes q_l (FP16) of layer l at h + l *
 * layer_stride_h and q + l * layer_stride_q, rows of ldh / ldq elements,
 * 256 columns (h_3's last d columns gamma, q_3's there the code of slope
 * 1), bit for bit what geraf_mlp_gemm_bf16 with epilogue 2 and the encoding
 * write layer by layer.
 *
 * geraf_mlp_chain_adjoint_bf16: `steps` chained 256 x 256 layer products
 * (the networks' reverse sweeps), each output kept as the next product's
 * operand: y_0 = (prologue(x) . W_0^T) * s(e0_0) [+ e1_0], y_s = (y_{s-1} .
 * W_s^T) * s(e0_s) [+ e1_s], as geraf_mlp_gemm_bf16 with epilogue 6 (e1
 * NULL) or 5 computes each, bit for bit; steps from mask_steps on take no
 * slope (y_s = z [+ e1_s], e0 unread).  The slopes s are the softplus codes'
 * (mask_kind 0) or the ReLU masks a > 0 of BF16 activations (mask_kind 1,
 * which needs the prologue and no e1).  wpacked: the steps' packed weights
 * back to back, each geraf_mlp_pack_bytes(256, 256) bytes (pad a narrower
 * input with zero weights).  x [M, ldx] (ldx 0 broadcasts a row); x2 the
 * slopes for the prologue x * s(x2) (times x_scale[p] when x_scale, FP32
 * [M], is not NULL) or NULL (not with e1); x_out, when not NULL, receives
 * the first operand as multiplied.  Step s reads e0 + s * step_e0 (and e1
 * + s * step_e1) and writes y + s * step_y (strides in elements, may be
 * negative); y may equal e1. */
int64_t geraf_mlp_pack_bytes(int64_t N, int64_t K);
int geraf_mlp_pack_weights(const float* w, int64_t ldw, int32_t transpose, float scale,
                           int64_t N, int64_t K, void* packed, geraf_stream_t stream);
int geraf_mlp_gemm_bf16(const uint16_t* x, int64_t ldx, const uint16_t* x2, int64_t ldx2,
                        const void* wpacked, const float* bias, uint16_t* y, int64_t ldy,
                        uint16_t* y2, int64_t ldy2, const uint16_t* e0, int64_t lde0,
                        const uint16_t* e1, int64_t lde1, int64_t M, int64_t N, int64_t K,
                        int32_t prologue, int32_t epilogue, float beta, float threshold,
                        geraf_stream_t stream);
int64_t geraf_mlp_wgrad_workspace_bytes(int64_t M, int64_t N, int64_t K);
int geraf_mlp_wgrad_bf16(const uint16_t* g, int64_t ldg, const uint16_t* g2, int64_t ldg2,
                         int32_t prologue_g, const uint16_t* x, int64_t ldx, const uint16_t* x2,
                         int64_t ldx2, int32_t prologue_x, float* dw, int64_t lddw, float* db,
                         int32_t accumulate, int64_t M, int64_t N, int64_t K, float beta,
                         float threshold, void* workspace, int64_t workspace_bytes,
                         geraf_stream_t stream);
int geraf_mlp_positional_encoding(const float* x, int64_t M, int32_t L, float scale,
                                  uint16_t* out, int64_t ldo, int32_t col0, int32_t pad_to,
                                  geraf_stream_t stream);
int geraf_mlp_pe_gradient(const float* x, int64_t M, int32_t L, float scale, const uint16_t* ga,
                          int64_t lda, const uint16_t* gb, int64_t ldb, float* n,
                          geraf_stream_t stream);
int geraf_mlp_pe_gradient_adjoint(const float* x, int64_t M, int32_t L, float scale,
                                  const float* dn, uint16_t* out, int64_t ldo, int32_t col0,
                                  int32_t pad_to, geraf_stream_t stream);
int64_t geraf_mlp_sdf_chain_pack_bytes(int32_t L);
int geraf_mlp_chain_adjoint_bf16(const uint16_t* x, int64_t ldx, const uint16_t* x2,
                                 int64_t ldx2, const float* x_scale, int32_t mask_kind,
                                 const void* wpacked, int32_t steps, int32_t mask_steps,
                                 const uint16_t* e0, int64_t lde0, int64_t step_e0,
                                 const uint16_t* e1, int64_t lde1, int64_t step_e1, uint16_t* y,
                                 int64_t ldy, int64_t step_y, uint16_t* x_out, int64_t ldx_out,
                                 int64_t M, geraf_stream_t stream);
int geraf_mlp_sdf_chain_bf16(const float* x, int64_t M, int32_t L, float scale,
                             const void* wpacked, const float* bias, const float* w_sdf,
                             float* sdf, uint16_t* h, int64_t ldh,
                             int64_t layer_stride_h, uint16_t* q, int64_t ldq,
                             int64_t layer_stride_q, float beta, float threshold,
                             geraf_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* GERAF_RF_H_ */
