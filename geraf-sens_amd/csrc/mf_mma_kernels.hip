// Matched-filter forward on the matrix cores (gfx950), FP32-accurate
// products from a three-term FP16 split.
//
// Forward operator (MatchedFilter_cuda_kernel.cu:65-115):
//   out[v] = (1/A) sum_a sum_k s[a,k] exp(-j 2 pi (f0 + k dfs) tau_av).
// With k = 32 alpha + beta (beta < 32) the exponential factors:
//   exp(-j 2 pi (f0 + k dfs) tau) = X[alpha,v] Y[beta,v],
//   X[alpha,v] = exp(-j 2 pi (f0 tau + 32 alpha dfs tau)),  Y[beta,v] = exp(-j 2 pi beta dfs tau),
// so per antenna the inner sum is a complex matrix product followed by a
// contraction over alpha:
//   T[alpha,v] = sum_beta S[alpha,beta] Y[beta,v],  out[v] += sum_alpha X[alpha,v] T[alpha,v].
// T runs on v_mfma_f32_16x16x32_f16: 16 rows of alpha (one 512-frequency
// tile), 16 voxels, depth 32 = beta.  Each real operand is split into FP16
// hi + lo (hi = fp16(x), lo = fp16(x - hi)) and a real product is taken as
// hi*hi + hi*lo + lo*hi (three instructions into one FP32 accumulator; the
// dropped lo*lo is ~2^-22 relative), so the products carry the ~22 bits of
// FP32 operands.  A complex product is four real ones (Re = Sr Yr - Si Yi,
// Im = Sr Yi + Si Yr): 12 instructions per tile and 16 voxels.
//
// The echo operands (hi/lo of Re s, Im s and -Im s, scaled per antenna by a
// power of two so that max |s| lies in [0.5, 1) and FP16 neither overflows
// nor loses the small entries) are laid out once in operand order by
// mf_mma_prep_kernel.  The main kernel gives each wave 64 voxels (four groups
// of 16).  Per antenna, each lane forms its own voxel's delay in FP64 (FP64
// distances, as the other forms) and publishes the reduced phase terms to
// the wave through LDS, with the voxel's step phasors exp(-j 2 pi {1, 2, 4, 32}
// dfs tau).  Per group, each lane then forms its eight Y phasors (beta =
// 8 g .. 8 g + 7 for lane j + 16 g) from one hardware phasor and a rotation
// tree, splits them, issues the 12 products and contracts its four rows of T
// with X (one phasor, Horner in the 32-frequency step).  Antennas
// accumulate per lane in FP64; the four lanes of a voxel are summed at the
// end.
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

constexpr int kMmaTile = 512;    // frequencies per tile: 16 rows of 32
constexpr int kMmaMaxTiles = 4;  // K <= 2048
constexpr int kMmaWaves = 4;     // waves per block, 64 voxels each
constexpr int kOpsPerTile = 6;   // Re hi, Re lo, Im hi, Im lo, -Im hi, -Im lo

// One antenna row, tile m: lane i + 16 g holds s[a, 512 m + 32 i + 8 g + e],
// e = 0..7, scaled, as FP16 hi and lo of the real and imaginary parts (and of
// the negated imaginary part, the operand of Re = Sr Yr - Si Yi).
// One wave per (antenna, tile); the scale (a power of two) is per antenna.
__global__ __launch_bounds__(64) void mf_mma_prep_kernel(const float* __restrict__ real,
                                                          const float* __restrict__ imag,
                                                          int K, int MT, h8* __restrict__ ops,
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
  h8 rh, rl, ih, il;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int k = kMmaTile * m + 32 * i + 8 * g + q;
    const float xr = k < K ? er[k] * down : 0.0f;
    const float xi = k < K ? ei[k] * down : 0.0f;
    rh[q] = (_Float16)xr;
    rl[q] = (_Float16)(xr - (float)rh[q]);
    ih[q] = (_Float16)xi;
    il[q] = (_Float16)(xi - (float)ih[q]);
  }
  h8* out = ops + ((size_t)a * MT + m) * kOpsPerTile * kWave;
  out[0 * kWave + lane] = rh;
  out[1 * kWave + lane] = rl;
  out[2 * kWave + lane] = ih;
  out[3 * kWave + lane] = il;
  out[4 * kWave + lane] = -ih;
  out[5 * kWave + lane] = -il;
}

template <int MT>
__global__ __launch_bounds__(64 * kMmaWaves) void mf_forward_mma_kernel(
    const h8* __restrict__ ops, const float* __restrict__ unscale,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    float* __restrict__ out_real, float* __restrict__ out_imag, const float* __restrict__ voxel,
    int A, int N, PhaseConsts pc) {
  // per voxel of the wave: step cycles (hi, lo) and carrier cycles; the step
  // phasors w1, w32 = exp(-j 2 pi {1, 32} dfs tau); w2 = w1^2, w4 = w2^2
  __shared__ float4 par[kMmaWaves][3][kWave];
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
    const h8* op = ops + (size_t)a * MT * kOpsPerTile * kWave + lane;
    h8 rh[MT], rl[MT], ih[MT], il[MT], nh[MT], nl[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      rh[m] = op[(m * kOpsPerTile + 0) * kWave];
      rl[m] = op[(m * kOpsPerTile + 1) * kWave];
      ih[m] = op[(m * kOpsPerTile + 2) * kWave];
      il[m] = op[(m * kOpsPerTile + 3) * kWave];
      nh[m] = op[(m * kOpsPerTile + 4) * kWave];
      nl[m] = op[(m * kOpsPerTile + 5) * kWave];
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
      const float2 w1 = cis_neg(frac_multiple(1.0f, c1));
      const float2 w32 = cis_neg(frac_multiple(32.0f, c1));
      const float2 w2 = cmul(w1, w1), w4 = cmul(w2, w2);
      par[wave][0][lane] = make_float4(c1.hi, c1.lo, (float)centered_frac(pc.f0 * tau), 0.0f);
      par[wave][1][lane] = make_float4(w1.x, w1.y, w32.x, w32.y);
      par[wave][2][lane] = make_float4(w2.x, w2.y, w4.x, w4.y);
    }
    wave_sync();
    // groups unrolled (the accumulators stay in registers), one at a time: the
    // scheduling barrier keeps the compiler from interleaving them, which
    // would cost registers (occupancy) and gain nothing
#pragma unroll
    for (int G = 0; G < 4; ++G) {
      const float4 p = par[wave][0][16 * G + j];
      const float4 q1 = par[wave][1][16 * G + j];
      const float4 q2 = par[wave][2][16 * G + j];
      const SplitCycles c = {p.x, p.y};
      const float2 w1 = make_float2(q1.x, q1.y), st = make_float2(q1.z, q1.w);
      const float2 w2 = make_float2(q2.x, q2.y), w4 = make_float2(q2.z, q2.w);
      // Y[8 g + e] for e = 0..7: one phasor and a rotation tree (depth <= 3)
      const float2 y0 = cis_neg(frac_multiple((float)(8 * g), c));
      float2 y[8];
      y[0] = y0;
      y[1] = cmul(y0, w1);
      y[2] = cmul(y0, w2);
      y[3] = cmul(y[1], w2);
      y[4] = cmul(y0, w4);
      y[5] = cmul(y[1], w4);
      y[6] = cmul(y[2], w4);
      y[7] = cmul(y[3], w4);
      h8 yrh, yrl, yih, yil;
      split8(y, yrh, yrl, yih, yil);
      // rows alpha = 16 m + 4 g + r of this lane: X[alpha] = X[16 m + 4 g] st^r
      // with st = w32; sum_r X T_r by Horner in st
      float2 part;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        f32x4v tr = {0.0f, 0.0f, 0.0f, 0.0f}, ti = {0.0f, 0.0f, 0.0f, 0.0f};
        tr = mma3(rh[m], rl[m], yrh, yrl, tr);
        tr = mma3(nh[m], nl[m], yih, yil, tr);
        ti = mma3(rh[m], rl[m], yih, yil, ti);
        ti = mma3(ih[m], il[m], yrh, yrl, ti);
        float2 h = make_float2(tr[3], ti[3]);
#pragma unroll
        for (int r = 2; r >= 0; --r) {
          const float2 t = cmul(h, st);
          h = make_float2(t.x + tr[r], t.y + ti[r]);
        }
        const float2 x = cis_neg(p.z + frac_multiple((float)(32 * (16 * m + 4 * g)), c));
        const float2 xh = cmul(x, h);
        if (m == 0) {
          part = xh;  // not 0 + xh: the compiler may not drop an add of +0
        } else {
          part.x += xh.x;
          part.y += xh.y;
        }
      }
      acc_r[G] = fma((double)part.x, up, acc_r[G]);
      acc_i[G] = fma((double)part.y, up, acc_i[G]);
      __builtin_amdgcn_sched_barrier(0);
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
  const int MT = mma_tiles(K);
  h8* ops = reinterpret_cast<h8*>(workspace);
  float* unscale =
      reinterpret_cast<float*>(reinterpret_cast<char*>(workspace) + align256(mma_ops_bytes(A, K)));
  hipLaunchKernelGGL(mf_mma_prep_kernel, dim3((unsigned)A, (unsigned)MT), dim3(kWave), 0, stream,
                     real, imag, (int)K, MT, ops, unscale);
  const int rc = check_launch("mf_mma_prep_kernel");
  if (rc != GERAF_RF_OK) return rc;
#define GERAF_MMA_LAUNCH(T)                                                                   \
  hipLaunchKernelGGL((mf_forward_mma_kernel<T>), dim3(blocks), dim3(64 * kMmaWaves), 0, stream, \
                     ops, unscale, tx_pos, rx_pos, out_real, out_imag, voxel, (int)A, (int)N, pc)
  switch (MT) {
    case 1: GERAF_MMA_LAUNCH(1); break;
    case 2: GERAF_MMA_LAUNCH(2); break;
    case 3: GERAF_MMA_LAUNCH(3); break;
    default: GERAF_MMA_LAUNCH(4); break;
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
    throw ArgError("A * ceil(K / 512) too large for the static scratch: the echo needs " +
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
