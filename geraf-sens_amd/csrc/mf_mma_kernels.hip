// Matched-filter forward on the matrix cores (gfx950), FP32-accurate
// products from a three-term FP16 split.
//
// Forward operator (MatchedFilter_cuda_kernel.cu:65-115):
//   out[v] = (1/A) sum_a sum_k s[a,k] exp(-j 2 pi (f0 + k dfs) tau_av).
// With k = 16 alpha + beta (beta < 16) the exponential factors:
//   exp(-j 2 pi (f0 + k dfs) tau) = X[alpha,v] Y[beta,v],
//   X[alpha,v] = exp(-j 2 pi (f0 tau + 16 alpha dfs tau)),  Y[beta,v] = exp(-j 2 pi beta dfs tau),
// so per antenna the inner sum is a complex matrix product followed by a
// contraction over alpha:
//   T[alpha,v] = sum_beta S[alpha,beta] Y[beta,v],  out[v] += sum_alpha X[alpha,v] T[alpha,v].
// T runs on v_mfma_f32_16x16x32_f16 with the complex product folded into the
// depth: depth 32 = 16 beta x (Re, Im) of Y, and two real products
//   Re T = sum_beta (Sr Yr - Si Yi),  Im T = sum_beta (Si Yr + Sr Yi)
// with A operands [Sr, -Si] and [Si, Sr] against one B operand [Yr, Yi].
// Each real operand is split into FP16 hi + lo (hi = fp16(x),
// lo = fp16(x - hi)) and a real product is taken as hi*hi + hi*lo + lo*hi
// (three instructions into one FP32 accumulator; the dropped lo*lo is ~2^-22
// relative), so the products carry the ~22 bits of FP32 operands: 6
// instructions per 16-row tile (256 frequencies) and 16 voxels.
//
// Per voxel there are 16 Y values (four per lane: the group's first phasor
// and three complex products) and 16 NT rows of T; the MFMA rows of tile m
// are assigned so that lane j + 16 g receives the consecutive rows
// alpha = 4 NT g + 4 m + r (r < 4), contracted by a two-level Horner in the
// 16- and 64-frequency steps against one X phasor per lane.
//
// The echo operands (hi/lo of [Sr, -Si] and [Si, Sr], scaled per antenna by
// a power of two so that max |s| lies in [0.5, 1) and FP16 neither overflows
// nor loses the small entries) are laid out once in operand order by
// mf_mma_prep_kernel.  The main kernel gives each wave 64 voxels (four groups
// of 16).  Per antenna, each lane forms its own voxel's delay in FP64 (FP64
// distances, as the other forms) and publishes to the wave through LDS the
// voxel's step phasors exp(-j 2 pi {1, 2, 16, 64} dfs tau) and, for each
// lane group g, its first Y phasor Y[4 g] and its X phasor X[4 NT g] (the
// hardware sine and cosine for the powers 1, 4, 16, 64 and 64 NT and for the
// carrier, the others by products).  Antennas accumulate per lane in FP64;
// the four lanes of a voxel are summed at the end.
//
// Operand layout of the instruction (checked on the GPU with an in-bounds
// probe): A[i][d] from lane i + 16 (d / 8), element d % 8; B[d][j] from lane
// j + 16 (d / 8), element d % 8; D[4 q + r][j] is element r of lane j + 16 q.
// Only the pairing of A's and B's depth index matters, and both use
// d = 8 (lane / 16) + element.
#include <hip/hip_runtime.h>

#include <math.h>

#include <mutex>
#include <string>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

constexpr int kMmaTile = 256;    // frequencies per tile: 16 rows of 16
constexpr int kMmaMaxTiles = 8;  // K <= 2048
constexpr int kMmaWaves = 4;     // waves per block, 64 voxels each
constexpr int kOpsPerTile = 4;   // [Sr, -Si] hi, lo; [Si, Sr] hi, lo

// One antenna row, tile m of NT: lane i + 16 g holds row i of the tile,
// alpha = 4 NT (i / 4) + 4 m + i % 4, depth elements e = 0..7 for
// beta = 4 g + e / 2, (Re, Im) = e % 2: [Sr, -Si] and [Si, Sr] of
// s[a, 16 alpha + beta], scaled, as FP16 hi and lo.
// One wave per (antenna, tile); the scale (a power of two) is per antenna.
__global__ __launch_bounds__(64) void mf_mma_prep_kernel(const float* __restrict__ real,
                                                          const float* __restrict__ imag,
                                                          int K, int NT, h8* __restrict__ ops,
                                                          float* __restrict__ unscale) {
  const int a = blockIdx.x;
  const int m = blockIdx.y;
  const int lane = threadIdx.x;
  const float* er = real + (size_t)a * K;
  const float* ei = imag + (size_t)a * K;
  float mx = 0.0f;
  for (int k = lane; k < K; k += kWave) mx = fmaxf(mx, fmaxf(fabsf(er[k]), fabsf(ei[k])));
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
  // 2^e > mx >= 2^(e-1); the operands are s * 2^-e (exact), the results are
  // multiplied back by 2^e
  int e = 0;
  if (mx > 0.0f && isfinite(mx)) (void)frexpf(mx, &e);
  const float down = ldexpf(1.0f, -e);
  if (m == 0 && lane == 0) unscale[a] = ldexpf(1.0f, e);
  const int i = lane & 15, g = lane >> 4;
  const int alpha = 4 * NT * (i >> 2) + 4 * m + (i & 3);
  h8 reh, rel, imh, iml;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int k = 16 * alpha + 4 * g + q;
    const float xr = k < K ? er[k] * down : 0.0f;
    const float xi = k < K ? ei[k] * down : 0.0f;
    // [Sr, -Si] against [Yr, Yi] is Re T; [Si, Sr] is Im T
    const float re0 = xr, re1 = -xi, im0 = xi, im1 = xr;
    reh[2 * q] = (_Float16)re0;
    reh[2 * q + 1] = (_Float16)re1;
    imh[2 * q] = (_Float16)im0;
    imh[2 * q + 1] = (_Float16)im1;
    rel[2 * q] = (_Float16)(re0 - (float)reh[2 * q]);
    rel[2 * q + 1] = (_Float16)(re1 - (float)reh[2 * q + 1]);
    iml[2 * q] = (_Float16)(im0 - (float)imh[2 * q]);
    iml[2 * q + 1] = (_Float16)(im1 - (float)imh[2 * q + 1]);
  }
  h8* out = ops + ((size_t)a * NT + m) * kOpsPerTile * kWave;
  out[0 * kWave + lane] = reh;
  out[1 * kWave + lane] = rel;
  out[2 * kWave + lane] = imh;
  out[3 * kWave + lane] = iml;
}

#ifndef GERAF_MF_MMA_GROUP_BARRIER
#define GERAF_MF_MMA_GROUP_BARRIER 1
#endif
#ifndef GERAF_MF_MMA_WPE
#define GERAF_MF_MMA_WPE 1
#endif

template <int NT>
__global__ __launch_bounds__(64 * kMmaWaves)
__attribute__((amdgpu_waves_per_eu(GERAF_MF_MMA_WPE))) void mf_forward_mma_kernel(
    const h8* __restrict__ ops, const float* __restrict__ unscale,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    float* __restrict__ out_real, float* __restrict__ out_imag, const float* __restrict__ voxel,
    int A, int N, PhaseConsts pc) {
  // per voxel of the wave: the step phasors w1, w2 = exp(-j 2 pi {1, 2} dfs
  // tau) and w16, w64; per lane group g of the voxel: its first Y phasor
  // Y[4 g] and its X phasor X[4 NT g]
  __shared__ float4 par[kMmaWaves][6][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int v0 = (blockIdx.x * kMmaWaves + wave) * kWave;
  if (v0 >= N) return;  // wave-uniform; only wave-level synchronisation below
  const int j = lane & 15, g = lane >> 4;
  const int vo = min(v0 + lane, N - 1);
  const double px = voxel[vo], py = voxel[(size_t)N + vo], pz = voxel[2 * (size_t)N + vo];

  // per-lane sums over the antennas in FP64 (one conversion and one FMA per
  // component and antenna; more accurate than compensated FP32 and cheaper)
  double acc_r[4], acc_i[4];
#pragma unroll
  for (int G = 0; G < 4; ++G) acc_r[G] = acc_i[G] = 0.0;

  for (int a = 0; a < A; ++a) {
    const h8* op = ops + (size_t)a * NT * kOpsPerTile * kWave + lane;
    h8 reh[NT], rel[NT], imh[NT], iml[NT];
#pragma unroll
    for (int m = 0; m < NT; ++m) {
      reh[m] = op[(m * kOpsPerTile + 0) * kWave];
      rel[m] = op[(m * kOpsPerTile + 1) * kWave];
      imh[m] = op[(m * kOpsPerTile + 2) * kWave];
      iml[m] = op[(m * kOpsPerTile + 3) * kWave];
    }
    const double up = (double)unscale[a];
    const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
    const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
    double d;
    if (monostatic(tx_pos, rx_pos, a))
      d = round_trip_f64<true>(tx, ty, tz, rx, ry, rz, px, py, pz);
    else
      d = round_trip_f64<false>(tx, ty, tz, rx, ry, rz, px, py, pz);
    const double tau = (d + pc.bias) * pc.inv_c;
    const SplitCycles c1 = split_cycles13(pc.dfs * tau);
    {
      // lane group g takes beta = 4 g .. 4 g + 3 and the rows from
      // alpha = 4 NT g (frequency 64 NT g) on
      const float2 w1 = cis_neg_reduced(frac_multiple(1.0f, c1));
      const float2 w16 = cis_neg_reduced(frac_multiple(16.0f, c1));
      const float2 w64 = cis_neg_reduced(frac_multiple(64.0f, c1));
      const float2 w4 = cis_neg_reduced(frac_multiple(4.0f, c1));
      const float2 wx = cis_neg_reduced(frac_multiple((float)(64 * NT), c1));
      const float2 x0 = cis_neg_reduced((float)centered_frac(pc.f0 * tau));
      const float2 w2 = csq(w1), w8 = csq(w4), w12 = cmul(w8, w4);
      const float2 x1 = cmul(x0, wx), x2 = cmul(x1, wx), x3 = cmul(x2, wx);
      par[wave][0][lane] = make_float4(w1.x, w1.y, w2.x, w2.y);
      par[wave][1][lane] = make_float4(w16.x, w16.y, w64.x, w64.y);
      par[wave][2][lane] = make_float4(1.0f, 0.0f, x0.x, x0.y);
      par[wave][3][lane] = make_float4(w4.x, w4.y, x1.x, x1.y);
      par[wave][4][lane] = make_float4(w8.x, w8.y, x2.x, x2.y);
      par[wave][5][lane] = make_float4(w12.x, w12.y, x3.x, x3.y);
    }
    wave_sync();
    // groups unrolled (the accumulators stay in registers), one at a time: the
    // scheduling barrier keeps the compiler from interleaving them, which
    // would cost registers (occupancy) and gain nothing
#pragma unroll
    for (int G = 0; G < 4; ++G) {
      const float4 q1 = par[wave][0][16 * G + j];
      const float4 q2 = par[wave][1][16 * G + j];
      const float4 p = par[wave][2 + g][16 * G + j];
      const float2 w1 = make_float2(q1.x, q1.y), w2 = make_float2(q1.z, q1.w);
      const float2 w16 = make_float2(q2.x, q2.y), w64 = make_float2(q2.z, q2.w);
      // Y[4 g + u], u = 0..3: the group's first phasor and three products, as
      // (Re, Im) pairs
      float2 y[4];
      y[0] = make_float2(p.x, p.y);
      y[1] = cmul(y[0], w1);
      y[2] = cmul(y[0], w2);
      y[3] = cmul(y[1], w2);
      h8 yh, yl;
      split4c(y, yh, yl);
      // every tile's products first (the Horner below then waits on none)
      f32x4v tr[NT], ti[NT];
#pragma unroll
      for (int m = 0; m < NT; ++m) {
        tr[m] = ti[m] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
        tr[m] = mma3(reh[m], rel[m], yh, yl, tr[m]);
        ti[m] = mma3(imh[m], iml[m], yh, yl, ti[m]);
      }
      // h = sum_m w64^m sum_r w16^r T[4 NT g + 4 m + r], tiles from the last
      float2 h;
#pragma unroll
      for (int m = NT - 1; m >= 0; --m) {
        float2 t = make_float2(tr[m][3], ti[m][3]);
#pragma unroll
        for (int r = 2; r >= 0; --r) t = cmul_add(t, w16, make_float2(tr[m][r], ti[m][r]));
        h = (m == NT - 1) ? t : cmul_add(h, w64, t);
      }
      const float2 part = cmul(make_float2(p.z, p.w), h);
      acc_r[G] = fma((double)part.x, up, acc_r[G]);
      acc_i[G] = fma((double)part.y, up, acc_i[G]);
      if (GERAF_MF_MMA_GROUP_BARRIER) __builtin_amdgcn_sched_barrier(0);
    }
    wave_sync();
  }

  const double inv_a = A > 0 ? 1.0 / (double)A : 0.0;
#pragma unroll
  for (int G = 0; G < 4; ++G) {
    double re = acc_r[G], im = acc_i[G];
    re += __shfl_xor(re, 16, kWave);
    im += __shfl_xor(im, 16, kWave);
    re += __shfl_xor(re, 32, kWave);
    im += __shfl_xor(im, 32, kWave);
    const int v = v0 + 16 * G + j;
    if (g == 0 && v < N) {
      out_real[v] = (float)(re * inv_a);
      out_imag[v] = (float)(im * inv_a);
    }
  }
}

int mma_tiles(int64_t K) { return (int)((K + kMmaTile - 1) / kMmaTile); }

size_t mma_ops_bytes(int64_t A, int64_t K) {
  return (size_t)A * (size_t)mma_tiles(K) * kOpsPerTile * kWave * sizeof(h8);
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int64_t geraf_mf_forward_mma_workspace_bytes(int64_t A, int64_t K) {
  if (A < 0 || K < 0) return -1;
  return (int64_t)(align256(mma_ops_bytes(A, K)) + align256((size_t)A * sizeof(float)));
}

namespace geraf {
namespace {

// Enqueues the prep and main kernels with `workspace` as the echo's operand
// buffer (at least geraf_mf_forward_mma_workspace_bytes(A, K), 256-byte
// aligned).  Shared by the caller-workspace and the static-scratch entry
// points.
int mma_forward(const float* real, const float* imag, const float* tx_pos, const float* rx_pos,
                float* out_real, float* out_imag, const float* voxel, int64_t A, int64_t K,
                int64_t N, const geraf_fmcw_params* params, void* workspace,
                int64_t workspace_bytes, hipStream_t stream) {
  const unsigned blocks = (unsigned)((N + 64 * kMmaWaves - 1) / (64 * kMmaWaves));
  const PhaseConsts pc = make_phase_consts(*params);
  if (A == 0 || K == 0) {
    // no antennas or no samples: the sum is empty; the kernel writes zeros
    hipLaunchKernelGGL((mf_forward_mma_kernel<1>), dim3(blocks), dim3(64 * kMmaWaves), 0,
                       stream, nullptr, nullptr, tx_pos, rx_pos, out_real, out_imag, voxel, 0,
                       (int)N, pc);
    return check_launch("mf_forward_mma_kernel");
  }
  require_ptrs({real, imag, tx_pos, rx_pos, workspace}, true);
  if (workspace_bytes < geraf_mf_forward_mma_workspace_bytes(A, K))
    throw ArgError("workspace too small (see geraf_mf_forward_mma_workspace_bytes)");
  if ((reinterpret_cast<uintptr_t>(workspace) & 255u) != 0)
    throw ArgError("workspace must be 256-byte aligned");
  const int NT = mma_tiles(K);
  h8* ops = reinterpret_cast<h8*>(workspace);
  float* unscale =
      reinterpret_cast<float*>(reinterpret_cast<char*>(workspace) + align256(mma_ops_bytes(A, K)));
  hipLaunchKernelGGL(mf_mma_prep_kernel, dim3((unsigned)A, (unsigned)NT), dim3(kWave), 0, stream,
                     real, imag, (int)K, NT, ops, unscale);
  const int rc = check_launch("mf_mma_prep_kernel");
  if (rc != GERAF_RF_OK) return rc;
#define GERAF_MMA_LAUNCH(T)                                                                   \
  hipLaunchKernelGGL((mf_forward_mma_kernel<T>), dim3(blocks), dim3(64 * kMmaWaves), 0, stream, \
                     ops, unscale, tx_pos, rx_pos, out_real, out_imag, voxel, (int)A, (int)N, pc)
  switch (NT) {
    case 1: GERAF_MMA_LAUNCH(1); break;
    case 2: GERAF_MMA_LAUNCH(2); break;
    case 3: GERAF_MMA_LAUNCH(3); break;
    case 4: GERAF_MMA_LAUNCH(4); break;
    case 5: GERAF_MMA_LAUNCH(5); break;
    case 6: GERAF_MMA_LAUNCH(6); break;
    case 7: GERAF_MMA_LAUNCH(7); break;
    default: GERAF_MMA_LAUNCH(8); break;
  }
#undef GERAF_MMA_LAUNCH
  return check_launch("mf_forward_mma_kernel");
}

void check_mma_args(int64_t A, int64_t K, int64_t N, const geraf_fmcw_params* params) {
  check_mf_sizes(A, K, N);
  check_params(params);
  if (K > (int64_t)kMmaTile * kMmaMaxTiles) throw ArgError("K must be <= 2048 for this form");
}

}  // namespace

// Device-resident scratch of geraf_mf_forward_f32_mma_noalloc: the echo in
// operand order, as the caller's workspace holds it in
// geraf_mf_forward_f32_mma.  Each GPU that loads the library holds one copy
// (GERAF_MF_MMA_SCRATCH_BYTES) until the process ends.  In host code the
// array's name is not a device address: the copy of the current device is
// found once with hipGetSymbolAddress and only that pointer is used.
static __device__ __align__(256) unsigned char scratch[GERAF_MF_MMA_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The address of the current device's copy of `scratch`, looked up on the
// first call on that device (with that device current) and kept.
void* scratch_of_current_device() {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) throw HipError(std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) throw ArgError("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  return g_scratch_ptr[dev];
}

}  // namespace
}  // namespace geraf

extern "C" int geraf_mf_forward_f32_mma(const float* real, const float* imag,
                                        const float* tx_pos, const float* rx_pos,
                                        float* out_real, float* out_imag, const float* voxel,
                                        int64_t A, int64_t K, int64_t N,
                                        const geraf_fmcw_params* params, void* workspace,
                                        int64_t workspace_bytes, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_mma_args(A, K, N, params);
  if (N == 0) return GERAF_RF_OK;
  require_ptrs({out_real, out_imag, voxel}, true);
  return mma_forward(real, imag, tx_pos, rx_pos, out_real, out_imag, voxel, A, K, N, params,
                     workspace, workspace_bytes, (hipStream_t)stream);
  GERAF_TRY_END
}

extern "C" int geraf_mf_forward_f32_mma_noalloc(const float* real, const float* imag,
                                                const float* tx_pos, const float* rx_pos,
                                                float* out_real, float* out_imag,
                                                const float* voxel, int64_t A, int64_t K,
                                                int64_t N, const geraf_fmcw_params* params,
                                                geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_mma_args(A, K, N, params);
  const int64_t need = geraf_mf_forward_mma_workspace_bytes(A, K);
  if (need > (int64_t)GERAF_MF_MMA_SCRATCH_BYTES)
    throw ArgError("A * ceil(K / 256) too large for the static scratch: the echo needs " +
                   std::to_string(need) + " bytes, the scratch holds " +
                   std::to_string((int64_t)GERAF_MF_MMA_SCRATCH_BYTES) +
                   " (use geraf_mf_forward_f32_mma with a workspace)");
  if (N == 0) return GERAF_RF_OK;
  require_ptrs({out_real, out_imag, voxel}, true);
  void* ws = (A == 0 || K == 0) ? nullptr : scratch_of_current_device();
  return mma_forward(real, imag, tx_pos, rx_pos, out_real, out_imag, voxel, A, K, N, params, ws,
                     (int64_t)GERAF_MF_MMA_SCRATCH_BYTES, (hipStream_t)stream);
  GERAF_TRY_END
}
