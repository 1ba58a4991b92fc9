// Matched-filter kernels for gfx950 (MI355X).
//
// Forward operator (MatchedFilter_cuda_kernel.cu:65-115):
//   out[v] = (1/A) * sum_a sum_k s[a,k] * exp(-j*2*pi*(f0 + k*dfs)*tau_av)
// Adjoint (MatchedFilter_cuda_kernel.cu:171-220):
//   g[a,k] = (1/norm) * sum_v gout[v] * exp(+j*2*pi*(f0 + k*dfs)*tau_av)
//
// Tuned forward: one voxel per lane, the wave walks the antennas in lock step,
// so the echo row s[a,:] and the antenna positions are wave-uniform and are
// read through the scalar data cache straight into SGPRs (no VGPR or LDS
// traffic for the 8 bytes a term needs).  Per (voxel, antenna) pair:
//   * distance, tau and the carrier cycles f0*tau are formed in FP64 and
//     reduced modulo one cycle; the three phasors (carrier, per-sample step
//     w = exp(-j*2*pi*dfs*tau) and the 32-sample step W = w^32) come from FP32
//     sincospi of reduced arguments;
//   * sum_k s_k w^k is evaluated by a two-level Horner scheme: 32-term inner
//     Horner in w, outer Horner in W over the 32-sample blocks.  A Horner step
//     acc = acc*w + s is 4 FP32 FMAs, i.e. the whole complex multiply-add per
//     term, and the power of w never exceeds 31 (re-seeded every 32 samples),
//     which bounds the |w|^k magnitude drift and the angle error growth;
//   * the pair's contribution base*S is added with Kahan compensation.
// Small volumes split the antennas over the waves of a block; the partial
// sums are combined in FP64 through LDS in a fixed order (deterministic).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <math.h>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

constexpr int kHornerBlock = 32;

// Voxels per lane and waves per block of the large-volume launch.
#ifndef GERAF_MF_VPL
#define GERAF_MF_VPL 2
#endif
#ifndef GERAF_MF_WAVES
#define GERAF_MF_WAVES 4
#endif

// ---------------------------------------------------------------- tuned forward

// VPL voxels per lane (tiles of 64 voxels, VPL apart): the VPL Horner chains
// share every echo value read into SGPRs and give the scheduler independent
// FMAs to interleave.
template <int WAVES, int VPL>
__global__ __launch_bounds__(64 * WAVES) void mf_forward_horner_kernel(
    const float* __restrict__ real, const float* __restrict__ imag,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    float* __restrict__ out_real, float* __restrict__ out_imag,
    const float* __restrict__ voxel, int A, int K, int N, int groups, PhaseConsts pc) {
  __shared__ double partial[WAVES][VPL][kWave][2];

  const int lane = threadIdx.x & (kWave - 1);
  // The wave index is uniform across the wave; saying so lets the antenna
  // loop bounds, positions and echo rows live in SGPRs (scalar loads).
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int group = wave % groups;  // antenna group of this wave
  const int tile = wave / groups;   // voxel tile group of this wave
  const int tiles_per_block = WAVES / groups;

  int v[VPL];
  bool valid[VPL];
  double px[VPL], py[VPL], pz[VPL];
#pragma unroll
  for (int i = 0; i < VPL; ++i) {
    v[i] = ((blockIdx.x * tiles_per_block + tile) * VPL + i) * kWave + lane;
    valid[i] = v[i] < N;
    const int vc = valid[i] ? v[i] : (N - 1);
    px[i] = voxel[vc];
    py[i] = voxel[(size_t)N + vc];
    pz[i] = voxel[2 * (size_t)N + vc];
  }

  const int ants_per_group = (A + groups - 1) / groups;
  const int a_begin = group * ants_per_group;
  const int a_end = min(A, a_begin + ants_per_group);

  const int nb = K / kHornerBlock;        // full 32-sample blocks
  const int rem = K - nb * kHornerBlock;  // samples in the partial top block

  KahanC acc[VPL];
#pragma unroll
  for (int i = 0; i < VPL; ++i) acc[i].init();

  for (int a = a_begin; a < a_end; ++a) {
    const double tx = tx_pos[3 * a + 0], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
    const double rx = rx_pos[3 * a + 0], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
    PathPhasors ph[VPL];
    auto phasors = [&](auto mono) {
#pragma unroll
      for (int i = 0; i < VPL; ++i) {
        const double d =
            round_trip_f64<decltype(mono)::value>(tx, ty, tz, rx, ry, rz, px[i], py[i], pz[i]);
        ph[i] = path_phasors((d + pc.bias) * pc.inv_c, pc, -1.0f);
      }
    };
    if (monostatic(tx_pos, rx_pos, a))
      phasors(std::true_type{});
    else
      phasors(std::false_type{});

    // a*K < 2^31 (checked on the host); readfirstlane keeps the row offset,
    // and with it every echo address below, in SGPRs.
    const int row = __builtin_amdgcn_readfirstlane(a * K);
    const float* __restrict__ er = real + row;
    const float* __restrict__ ei = imag + row;

    float Sr[VPL], Si[VPL];
#pragma unroll
    for (int i = 0; i < VPL; ++i) Sr[i] = Si[i] = 0.0f;
    if (rem) {  // partial top block, k = K-1 .. 32*nb
      const int k0 = nb * kHornerBlock;
#pragma unroll
      for (int i = 0; i < VPL; ++i) {
        Sr[i] = er[K - 1];
        Si[i] = ei[K - 1];
      }
      for (int k = K - 2; k >= k0; --k) {
        const float xr = er[k], xi = ei[k];
#pragma unroll
        for (int i = 0; i < VPL; ++i) {
          const float u1 = fmaf(Sr[i], ph[i].w.x, xr);
          const float u2 = fmaf(Si[i], ph[i].w.x, xi);
          const float nr = fmaf(-Si[i], ph[i].w.y, u1);
          Si[i] = fmaf(Sr[i], ph[i].w.y, u2);
          Sr[i] = nr;
        }
      }
    }
    for (int b = nb - 1; b >= 0; --b) {
      const float* __restrict__ xr = er + b * kHornerBlock;
      const float* __restrict__ xi = ei + b * kHornerBlock;
      float ir[VPL], ii[VPL];
#pragma unroll
      for (int i = 0; i < VPL; ++i) {
        ir[i] = xr[kHornerBlock - 1];
        ii[i] = xi[kHornerBlock - 1];
      }
#pragma unroll
      for (int j = kHornerBlock - 2; j >= 0; --j) {
#pragma unroll
        for (int i = 0; i < VPL; ++i) {
          const float u1 = fmaf(ir[i], ph[i].w.x, xr[j]);
          const float u2 = fmaf(ii[i], ph[i].w.x, xi[j]);
          const float nr = fmaf(-ii[i], ph[i].w.y, u1);
          ii[i] = fmaf(ir[i], ph[i].w.y, u2);
          ir[i] = nr;
        }
      }
      // S = S*W + inner
#pragma unroll
      for (int i = 0; i < VPL; ++i) {
        const float u1 = fmaf(Sr[i], ph[i].W.x, ir[i]);
        const float u2 = fmaf(Si[i], ph[i].W.x, ii[i]);
        const float nr = fmaf(-Si[i], ph[i].W.y, u1);
        Si[i] = fmaf(Sr[i], ph[i].W.y, u2);
        Sr[i] = nr;
      }
    }
#pragma unroll
    for (int i = 0; i < VPL; ++i) acc[i].add(cmul(ph[i].base, make_float2(Sr[i], Si[i])));
  }

  const double inv_a = 1.0 / (double)A;
  if (groups == 1) {
#pragma unroll
    for (int i = 0; i < VPL; ++i)
      if (valid[i]) {
        out_real[v[i]] = (float)(acc[i].re() * inv_a);
        out_imag[v[i]] = (float)(acc[i].im() * inv_a);
      }
    return;
  }
#pragma unroll
  for (int i = 0; i < VPL; ++i) {
    partial[wave][i][lane][0] = acc[i].re();
    partial[wave][i][lane][1] = acc[i].im();
  }
  __syncthreads();
  if (group == 0) {
#pragma unroll
    for (int i = 0; i < VPL; ++i) {
      if (!valid[i]) continue;
      double sr = 0.0, si = 0.0;
      for (int g = 0; g < groups; ++g) {
        sr += partial[wave + g][i][lane][0];
        si += partial[wave + g][i][lane][1];
      }
      out_real[v[i]] = (float)(sr * inv_a);
      out_imag[v[i]] = (float)(si * inv_a);
    }
  }
}

// ---------------------------------------------------------------- plain FP32 form

// Reference formula: FP32 geometry, FP32 phase and a sincosf per term
// (MatchedFilter_cuda_kernel.cu:87-114).
__global__ __launch_bounds__(256) void mf_forward_plain_f32_kernel(
    const float* __restrict__ real, const float* __restrict__ imag,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    float* __restrict__ out_real, float* __restrict__ out_imag,
    const float* __restrict__ voxel, int A, int K, int N, float f0, float slope, float fs,
    float c, float bias) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const float px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
  const float two_pi = 6.28318530717958647692f;
  float acc_r = 0.0f, acc_i = 0.0f;
  for (int a = 0; a < A; ++a) {
    const float dtx = tx_pos[3 * a] - px, dty = tx_pos[3 * a + 1] - py,
                dtz = tx_pos[3 * a + 2] - pz;
    const float drx = rx_pos[3 * a] - px, dry = rx_pos[3 * a + 1] - py,
                drz = rx_pos[3 * a + 2] - pz;
    const float dist = sqrtf(dtx * dtx + dty * dty + dtz * dtz) +
                       sqrtf(drx * drx + dry * dry + drz * drz) + bias;
    const float tau = dist / c;
    const float base_phase = two_pi * f0 * tau;
    for (int k = 0; k < K; ++k) {
      const float t = (float)k / fs;
      const float phase = base_phase + two_pi * slope * t * tau;
      const float cs = cosf(phase), ns = -sinf(phase);
      const float sr = real[(size_t)a * K + k], si = imag[(size_t)a * K + k];
      acc_r += cs * sr - ns * si;
      acc_i += cs * si + ns * sr;
    }
  }
  out_real[v] = acc_r / (float)A;
  out_imag[v] = acc_i / (float)A;
}

// ---------------------------------------------------------------- plain FP64 form

__global__ __launch_bounds__(256) void mf_forward_f64_kernel(
    const double* __restrict__ real, const double* __restrict__ imag,
    const double* __restrict__ tx_pos, const double* __restrict__ rx_pos,
    double* __restrict__ out_real, double* __restrict__ out_imag,
    const double* __restrict__ voxel, int A, int K, int N, PhaseConsts pc) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
  double acc_r = 0.0, acc_i = 0.0;
  for (int a = 0; a < A; ++a) {
    const double d = dist3_f64(tx_pos[3 * a], tx_pos[3 * a + 1], tx_pos[3 * a + 2], px, py, pz) +
                     dist3_f64(rx_pos[3 * a], rx_pos[3 * a + 1], rx_pos[3 * a + 2], px, py, pz);
    const double tau = (d + pc.bias) * pc.inv_c;
    const double fr0 = centered_frac(pc.f0 * tau);
    const double cyc1 = pc.dfs * tau;
    for (int k = 0; k < K; ++k) {
      double s, c;
      sincospi(2.0 * centered_frac(fr0 + (double)k * cyc1), &s, &c);
      const double sr = real[(size_t)a * K + k], si = imag[(size_t)a * K + k];
      // s * exp(-j*phi) = (sr + j si)(c - j s)
      acc_r += c * sr + s * si;
      acc_i += c * si - s * sr;
    }
  }
  out_real[v] = acc_r / (double)A;
  out_imag[v] = acc_i / (double)A;
}

// ---------------------------------------------------------------- adjoint

// Block = (antenna a, a group of up to kAdjWaves 32-sample chunks).  The
// voxels are streamed in tiles of one voxel per thread: first every thread
// forms its voxel's per-(a, v) terms once -- round-trip distance (FP64 from an
// FP32 rsqrt seed), delay tau, the output gradient times the carrier phasor
// exp(+j 2 pi f0 tau), the per-sample step phasor and the FP64 step in cycles
// -- and stages them in LDS; then each wave takes one chunk (and, when the
// block has fewer chunks than waves, a share of the tile's voxels), seeds the
// chunk's first phasor from the FP64-reduced phase and rotates it through the
// chunk's 32 samples into 32 private complex accumulators.  The per-lane
// accumulators are reduced across the wave and the voxel shares in a fixed
// order at the end.
constexpr int kAdjChunk = 32;
constexpr int kAdjWaves = 8;
constexpr int kAdjTile = 64 * kAdjWaves;

struct AdjStage {
  float2 cg[kAdjTile];      // output gradient times the carrier phasor
  float2 w[kAdjTile];       // per-sample step phasor
  double cyc1[kAdjTile];    // per-sample step [cycles]
};

__global__ __launch_bounds__(kAdjTile) void mf_backward_kernel(
    const float* __restrict__ gout_real, const float* __restrict__ gout_imag,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    float* __restrict__ grad_real, float* __restrict__ grad_imag,
    const float* __restrict__ voxel, int A, int K, int N, double inv_norm, PhaseConsts pc,
    int chunks_per_block) {
  __shared__ AdjStage st;
  __shared__ float red[kAdjWaves][kAdjChunk][2];
  const int a = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int cpb = chunks_per_block;
  const int shares = kAdjWaves / cpb;  // voxel shares per chunk
  const int my_chunk = blockIdx.y * cpb + wave % cpb;
  const int share = wave / cpb;
  const int k0 = my_chunk * kAdjChunk;
  const bool wave_ok = wave < cpb * shares && k0 < K;

  const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
  const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
  const double k0d = (double)k0;

  float pr[kAdjChunk], pi[kAdjChunk];
#pragma unroll
  for (int j = 0; j < kAdjChunk; ++j) pr[j] = pi[j] = 0.0f;

  for (int base = 0; base < N; base += kAdjTile) {
    {
      const int v = base + threadIdx.x;
      float2 cg = make_float2(0.0f, 0.0f), w = make_float2(1.0f, 0.0f);
      double cyc1 = 0.0;
      if (v < N) {
        const float gr = gout_real[v], gi = gout_imag[v];
        const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
        const double d =
            dist3_f64_fast(tx, ty, tz, px, py, pz) + dist3_f64_fast(rx, ry, rz, px, py, pz);
        const double tau = (d + pc.bias) * pc.inv_c;
        cyc1 = pc.dfs * tau;
        cg = cmul(make_float2(gr, gi), phasor_cycles((float)centered_frac(pc.f0 * tau), 1.0f));
        w = phasor_cycles((float)centered_frac(cyc1), 1.0f);
      }
      st.cg[threadIdx.x] = cg;
      st.w[threadIdx.x] = w;
      st.cyc1[threadIdx.x] = cyc1;
    }
    __syncthreads();
    if (wave_ok) {
      const int n = min(kAdjTile, N - base);
      for (int i = share * kWave + lane; i < n; i += shares * kWave) {
        const float2 w = st.w[i];
        float2 z = cmul(st.cg[i], phasor_cycles((float)centered_frac(k0d * st.cyc1[i]), 1.0f));
#pragma unroll
        for (int j = 0; j < kAdjChunk; ++j) {
          pr[j] += z.x;
          pi[j] += z.y;
          z = cmul(z, w);
        }
      }
    }
    __syncthreads();
  }

#pragma unroll
  for (int j = 0; j < kAdjChunk; ++j) {
    const float sr = wave_sum(pr[j]);
    const float si = wave_sum(pi[j]);
    if (lane == 0) {
      red[wave][j][0] = sr;
      red[wave][j][1] = si;
    }
  }
  __syncthreads();
  // thread (chunk c of the block, sample j) sums its chunk's voxel shares
  const int c = threadIdx.x / kAdjChunk;
  const int j = threadIdx.x % kAdjChunk;
  if (c < cpb) {
    const int k = (blockIdx.y * cpb + c) * kAdjChunk + j;
    if (k < K) {
      double sr = 0.0, si = 0.0;
      for (int s = 0; s < shares; ++s) {
        sr += red[s * cpb + c][j][0];
        si += red[s * cpb + c][j][1];
      }
      grad_real[(size_t)a * K + k] = (float)(sr * inv_norm);
      grad_imag[(size_t)a * K + k] = (float)(si * inv_norm);
    }
  }
}

// Matrix-core adjoint (128 < K <= 1024).  Block = antenna a.  With
// k = 256 m + 16 al + b (b < 16)
//   grad[a, k] = sum_v X[16 m + al, v] Y[v, b],
//   X = cg_v exp(+j 2 pi 16 (16 m + al) cyc1_v),  Y = exp(+j 2 pi b cyc1_v),
// cg_v = g_v exp(+j 2 pi f0 tau_v), one 16 x 16 output tile per m,
// on v_mfma_f32_16x16x32_f16 with FP16 hi/lo operands (depth 32 voxels per
// instruction): A[i][d] = X[16 m + i, v_d]
// and B[d][j] = Y[v_d, j] with d = 8 (lane / 16) + element, i.e. lane i + 16 g
// forms X and Y for the eight voxels 8 g .. 8 g + 7 of its 32-voxel chunk.
// Each real operand is split into hi = fp16(x) and lo = fp16(x - hi), a real
// product is hi*hi + hi*lo + lo*hi into one FP32 accumulator, a complex
// product four real ones (12 instructions per 16 x 16 tile and 32 voxels).
//
// Each wave streams its own 64-voxel bases (no block-wide synchronisation
// until the final reduction).  Per base, lane l first forms voxel l's terms
// and writes a table of phasor factors to the wave's LDS:
//   Y[v, b]  = exp(+j 2 pi b c_v)               = ya[b & 3] * yb[b >> 2],
//   X[al, v] = cg_v exp(+j 2 pi 16 al c_v)      = xa[al & 3] * xb[al >> 2],
//   ya = w^{0,1,2,3}, yb = w^{0,4,8,12}, xa = cg w^{0,16,32,48}, xb = w^{0,64,128,192},
// with c_v the per-frequency step in cycles, w = exp(+j 2 pi c_v) (the powers
// 1, 4, 16, 64 and 256 from the hardware sine and cosine, the others by one
// squaring and one product), cg_v the output gradient times the carrier
// phasor and rot = w^256 the step to the next 256-frequency tile.  The
// consumers then form every operand with one complex product of two table
// entries instead of a sine and a cosine each.  cg is scaled per wave base by
// a power of two (max |cg| in [0.5, 1)) so FP16 neither overflows nor drops
// small gradients, and the base's sums are scaled back before the FP64 flush.
#ifndef GERAF_ADJ16_WPE
#define GERAF_ADJ16_WPE 3  // waves per SIMD the register allocation aims for
#endif
constexpr int kAdjFactors = 9;              // table entries per voxel
constexpr int kAdjFactorStride = kWave + 1;  // float4 per entry (odd: bank spread)
constexpr int kAdjTableFloat4 = kAdjFactors * kAdjFactorStride;  // per wave
constexpr int kAdj16Waves = 4;                     // waves per block (one antenna)
constexpr int kAdj16Threads = kWave * kAdj16Waves;
constexpr int kAdjRedPerWave = 8 * kWave;  // (Re, Im) x 4 rows per lane, one tile

template <int MT>
__global__ __launch_bounds__(kAdj16Threads)
__attribute__((amdgpu_waves_per_eu(MT >= 4 ? 2 : GERAF_ADJ16_WPE))) void mf_backward_mma16_kernel(
    const float* __restrict__ gout_real, const float* __restrict__ gout_imag,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    float* __restrict__ grad_real, float* __restrict__ grad_imag,
    const float* __restrict__ voxel, int A, int K, int N, double inv_norm, PhaseConsts pc) {
  // the waves' factor tables; after the voxel loop, the cross-wave reduction
  __shared__ float4 lds4[kAdj16Waves * kAdjTableFloat4];
  static_assert(sizeof(lds4) >= sizeof(double) * kAdj16Waves * kAdjRedPerWave,
                "reduction space");
  const int a = blockIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = lane & 15, g = lane >> 4;
  const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
  const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
  const bool mono = monostatic(tx_pos, rx_pos, a);
  float4* tbl = lds4 + wave * kAdjTableFloat4;
  // this lane's entries: ya/xa of index i & 3, yb/xb of index i >> 2
  const float4* tp = tbl + (i & 3) * kAdjFactorStride;
  const float4* tq = tbl + (4 + (i >> 2)) * kAdjFactorStride;
  const float4* tr = tbl + 8 * kAdjFactorStride;

  double sr[MT][4], si[MT][4];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) sr[m][v] = si[m][v] = 0.0;

  for (int base = wave * kWave; base < N; base += kAdj16Threads) {
    double up;
    {
      const int v = base + lane;
      float2 cg = make_float2(0.0f, 0.0f);
      SplitCycles c1 = {0.0f, 0.0f};
      if (v < N) {
        const float gr = gout_real[v], gi = gout_imag[v];
        const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
        const double d = mono ? round_trip_f64<true>(tx, ty, tz, rx, ry, rz, px, py, pz)
                              : round_trip_f64<false>(tx, ty, tz, rx, ry, rz, px, py, pz);
        const double tau = (d + pc.bias) * pc.inv_c;
        cg = cmul(make_float2(gr, gi), phasor_cycles_hw((float)centered_frac(pc.f0 * tau)));
        c1 = split_cycles(pc.dfs * tau);
      }
      const float mx = wave_max_dpp(fmaxf(fabsf(cg.x), fabsf(cg.y)));
      // cg * 2^-e with 2^e > max |cg| (exact); the base's sums are scaled back
      int e = 0;
      if (mx > 0.0f && isfinite(mx)) (void)frexpf(mx, &e);
      const float down = ldexpf(1.0f, -e);
      up = ldexp(1.0, e);
      cg = make_float2(cg.x * down, cg.y * down);
      const float2 w1 = phasor_cycles_hw(frac_multiple(1.0f, c1));
      const float2 w2 = csq(w1), w3 = cmul(w2, w1);
      const float2 w4 = phasor_cycles_hw(frac_multiple(4.0f, c1));
      const float2 w8 = csq(w4), w12 = cmul(w8, w4);
      const float2 w16 = phasor_cycles_hw(frac_multiple(16.0f, c1));
      const float2 w32 = csq(w16), w48 = cmul(w32, w16);
      const float2 w64 = phasor_cycles_hw(frac_multiple(64.0f, c1));
      const float2 w128 = csq(w64), w192 = cmul(w128, w64);
      const float2 x1 = cmul(cg, w16), x2 = cmul(cg, w32), x3 = cmul(cg, w48);
      tbl[0 * kAdjFactorStride + lane] = make_float4(1.0f, 0.0f, cg.x, cg.y);
      tbl[1 * kAdjFactorStride + lane] = make_float4(w1.x, w1.y, x1.x, x1.y);
      tbl[2 * kAdjFactorStride + lane] = make_float4(w2.x, w2.y, x2.x, x2.y);
      tbl[3 * kAdjFactorStride + lane] = make_float4(w3.x, w3.y, x3.x, x3.y);
      tbl[4 * kAdjFactorStride + lane] = make_float4(1.0f, 0.0f, 1.0f, 0.0f);
      tbl[5 * kAdjFactorStride + lane] = make_float4(w4.x, w4.y, w64.x, w64.y);
      tbl[6 * kAdjFactorStride + lane] = make_float4(w8.x, w8.y, w128.x, w128.y);
      tbl[7 * kAdjFactorStride + lane] = make_float4(w12.x, w12.y, w192.x, w192.y);
      if constexpr (MT > 1) {
        const float2 rot = phasor_cycles_hw(frac_multiple(256.0f, c1));
        tbl[8 * kAdjFactorStride + lane] = make_float4(rot.x, rot.y, 0.0f, 0.0f);
      }
    }
    wave_sync();
    // Re = sum Xr Yr - sum Xi Yi from two accumulators (no negated operand),
    // Im = sum (Xr Yi + Xi Yr)
    f32x4v prr[MT], pii[MT], pim[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) prr[m] = pii[m] = pim[m] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
    // both chunks unrolled: the first products of each base then start from
    // a zero accumulator operand instead of registers zeroed per base
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int v0 = 32 * c + 8 * g;  // this lane's eight voxels (of the base)
      float2 x[8];
      h8 yrh, yrl, yih, yil;
      {
        float2 y[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const float4 P = tp[v0 + q], Q = tq[v0 + q];
          y[q] = cmul(make_float2(P.x, P.y), make_float2(Q.x, Q.y));
          x[q] = cmul(make_float2(P.z, P.w), make_float2(Q.z, Q.w));
        }
        split8(y, yrh, yrl, yih, yil);
      }
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        if (m > 0) {
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const float4 R = tr[v0 + q];
            x[q] = cmul(x[q], make_float2(R.x, R.y));
          }
        }
        h8 xrh, xrl, xih, xil;
        split8(x, xrh, xrl, xih, xil);
        prr[m] = mma3(xrh, xrl, yrh, yrl, prr[m]);
        pii[m] = mma3(xih, xil, yih, yil, pii[m]);
        pim[m] = mma3(xrh, xrl, yih, yil, pim[m]);
        pim[m] = mma3(xih, xil, yrh, yrl, pim[m]);
      }
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        sr[m][v] += (double)(prr[m][v] - pii[m][v]) * up;
        si[m][v] += (double)pim[m][v] * up;
      }
    wave_sync();  // the table is rewritten by the next base
  }

  // reduce the waves in a fixed order, one 256-frequency tile at a time:
  // entry t = 8 lane + 2 v + (0: Re, 1: Im) is row 4 (lane / 16) + v, column
  // lane % 16 of the tile, i.e. frequency 64 (lane / 16) + 16 v + lane % 16
  __syncthreads();
  double (*red)[kAdjRedPerWave] = reinterpret_cast<double (*)[kAdjRedPerWave]>(lds4);
#pragma unroll
  for (int m = 0; m < MT; ++m) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      red[wave][8 * lane + 2 * v] = sr[m][v];
      red[wave][8 * lane + 2 * v + 1] = si[m][v];
    }
    __syncthreads();
#pragma unroll
    for (int t = threadIdx.x; t < kAdjRedPerWave; t += kAdj16Threads) {
      double acc = 0.0;
#pragma unroll
      for (int w = 0; w < kAdj16Waves; ++w) acc += red[w][t];
      const int tl = t >> 3, te = t & 7;
      const int k = 256 * m + 64 * (tl >> 4) + 16 * (te >> 1) + (tl & 15);
      if (k < K) {
        float* out = (te & 1) ? grad_imag : grad_real;
        out[(size_t)a * K + k] = (float)(acc * inv_norm);
      }
    }
    __syncthreads();
  }
}

// Plain FP64 adjoint: one thread per (a, k), FP64 phase and sincos per voxel.
__global__ __launch_bounds__(256) void mf_backward_f64_kernel(
    const double* __restrict__ gout_real, const double* __restrict__ gout_imag,
    const double* __restrict__ tx_pos, const double* __restrict__ rx_pos,
    double* __restrict__ grad_real, double* __restrict__ grad_imag,
    const double* __restrict__ voxel, int A, int K, int N, double inv_norm, PhaseConsts pc) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)A * K) return;
  const int a = (int)(idx / K);
  const int k = (int)(idx % K);
  const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
  const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
  double acc_r = 0.0, acc_i = 0.0;
  for (int v = 0; v < N; ++v) {
    const double px = voxel[v], py = voxel[(size_t)N + v], pz = voxel[2 * (size_t)N + v];
    const double d = dist3_f64(tx, ty, tz, px, py, pz) + dist3_f64(rx, ry, rz, px, py, pz);
    const double tau = (d + pc.bias) * pc.inv_c;
    double s, c;
    sincospi(2.0 * centered_frac(centered_frac(pc.f0 * tau) + (double)k * (pc.dfs * tau)), &s,
             &c);
    const double gr = gout_real[v], gi = gout_imag[v];
    acc_r += c * gr - s * gi;
    acc_i += s * gr + c * gi;
  }
  grad_real[idx] = acc_r * inv_norm;
  grad_imag[idx] = acc_i * inv_norm;
}


// Waves wanted per launch: 8 per SIMD of the 256-CU part, so that two of them
// are ready to dual-issue most of the time.
constexpr long long kWantWaves = 8LL * 1024;

// Antenna groups per voxel unit (a power of two, at most `waves` and A) so
// that units * groups reaches kWantWaves.
int pick_groups(int A, long long units, int waves) {
  int g = 1;
  while (g < waves && g * 2 <= A && units * g < kWantWaves) g *= 2;
  return g;
}

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int geraf_mf_forward_f32(const float* real, const float* imag, const float* tx_pos,
                                    const float* rx_pos, float* out_real, float* out_imag,
                                    const float* voxel, int64_t A, int64_t K, int64_t N,
                                    const geraf_fmcw_params* params, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_mf_sizes(A, K, N);
  check_params(params);
  if (N == 0) return GERAF_RF_OK;
  require_ptrs({real, imag, tx_pos, rx_pos}, A * K > 0 || A > 0);
  require_ptrs({out_real, out_imag, voxel}, true);
  if (A * K >= (INT64_C(1) << 31)) throw ArgError("A*K must be < 2^31");
  const PhaseConsts pc = make_phase_consts(*params);
  const int n = (int)N;
  const int tiles = (n + kWave - 1) / kWave;
  constexpr int V = GERAF_MF_VPL;
  const int units = (tiles + V - 1) / V;  // waves of V voxel tiles each
  if (units < kWantWaves && A > 1) {
    // volume too small to fill the part: split the antennas over the waves of
    // a block as well
    constexpr int W = 16;
    const int g = pick_groups((int)A, units, W);
    const int units_per_block = W / g;
    const int blocks = (units + units_per_block - 1) / units_per_block;
    hipLaunchKernelGGL((mf_forward_horner_kernel<W, V>), dim3(blocks), dim3(64 * W), 0,
                       (hipStream_t)stream, real, imag, tx_pos, rx_pos, out_real, out_imag,
                       voxel, (int)A, (int)K, n, g, pc);
  } else {
    constexpr int W = GERAF_MF_WAVES;
    const int blocks = (tiles + W * V - 1) / (W * V);
    hipLaunchKernelGGL((mf_forward_horner_kernel<W, V>), dim3(blocks), dim3(64 * W), 0,
                       (hipStream_t)stream, real, imag, tx_pos, rx_pos, out_real, out_imag,
                       voxel, (int)A, (int)K, n, 1, pc);
  }
  return check_launch("mf_forward_horner_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mf_forward_plain_f32(const float* real, const float* imag,
                                          const float* tx_pos, const float* rx_pos,
                                          float* out_real, float* out_imag, const float* voxel,
                                          int64_t A, int64_t K, int64_t N,
                                          const geraf_fmcw_params* params,
                                          geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_mf_sizes(A, K, N);
  check_params(params);
  if (N == 0) return GERAF_RF_OK;
  require_ptrs({real, imag, tx_pos, rx_pos}, A > 0);
  require_ptrs({out_real, out_imag, voxel}, true);
  const int blocks = (int)((N + 255) / 256);
  hipLaunchKernelGGL(mf_forward_plain_f32_kernel, dim3(blocks), dim3(256), 0,
                     (hipStream_t)stream, real, imag, tx_pos, rx_pos, out_real, out_imag, voxel,
                     (int)A, (int)K, (int)N, (float)params->f0, (float)params->slope,
                     (float)params->fs, (float)params->c, (float)params->path_bias);
  return check_launch("mf_forward_plain_f32_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mf_forward_f64(const double* real, const double* imag,
                                    const double* tx_pos, const double* rx_pos,
                                    double* out_real, double* out_imag, const double* voxel,
                                    int64_t A, int64_t K, int64_t N,
                                    const geraf_fmcw_params* params, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_mf_sizes(A, K, N);
  check_params(params);
  if (N == 0) return GERAF_RF_OK;
  require_ptrs({real, imag, tx_pos, rx_pos}, A > 0);
  require_ptrs({out_real, out_imag, voxel}, true);
  const int blocks = (int)((N + 255) / 256);
  hipLaunchKernelGGL(mf_forward_f64_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                     real, imag, tx_pos, rx_pos, out_real, out_imag, voxel, (int)A, (int)K,
                     (int)N, make_phase_consts(*params));
  return check_launch("mf_forward_f64_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mf_backward_f32(const float* grad_out_real, const float* grad_out_imag,
                                     const float* tx_pos, const float* rx_pos,
                                     float* grad_real, float* grad_imag, const float* voxel,
                                     int64_t A, int64_t K, int64_t N, int64_t norm_count,
                                     const geraf_fmcw_params* params, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_mf_sizes(A, K, N);
  check_params(params);
  if (A == 0 || K == 0) return GERAF_RF_OK;
  require_ptrs({tx_pos, rx_pos, grad_real, grad_imag}, true);
  require_ptrs({grad_out_real, grad_out_imag, voxel}, N > 0);
  if (K > 128 && K <= 1024) {
    const PhaseConsts pc = make_phase_consts(*params);
    const double inv_norm = 1.0 / (double)norm_count;
    hipStream_t st = (hipStream_t)stream;
#define GERAF_MF_ADJ(MT_)                                                                      \
  hipLaunchKernelGGL(mf_backward_mma16_kernel<MT_>, dim3((unsigned)A), dim3(kAdj16Threads), 0, \
                     st,                                                                     \
                     grad_out_real, grad_out_imag, tx_pos, rx_pos, grad_real, grad_imag,     \
                     voxel, (int)A, (int)K, (int)N, inv_norm, pc)
    if (K <= 256)
      GERAF_MF_ADJ(1);
    else if (K <= 512)
      GERAF_MF_ADJ(2);
    else
      GERAF_MF_ADJ(4);
#undef GERAF_MF_ADJ
    return check_launch("mf_backward_mma16_kernel");
  }
  // chunks per block: a power of two dividing kAdjWaves, at most the chunk count
  const int chunks = (int)((K + kAdjChunk - 1) / kAdjChunk);
  int cpb = kAdjWaves;
  while (cpb > 1 && cpb / 2 >= chunks) cpb /= 2;
  const dim3 grid((unsigned)A, (unsigned)((chunks + cpb - 1) / cpb));
  if (grid.y > 65535u) throw ArgError("too many samples per antenna");
  hipLaunchKernelGGL(mf_backward_kernel, grid, dim3(kAdjTile), 0, (hipStream_t)stream,
                     grad_out_real, grad_out_imag, tx_pos, rx_pos, grad_real, grad_imag, voxel,
                     (int)A, (int)K, (int)N, 1.0 / (double)norm_count,
                     make_phase_consts(*params), cpb);
  return check_launch("mf_backward_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mf_backward_f64(const double* grad_out_real, const double* grad_out_imag,
                                     const double* tx_pos, const double* rx_pos,
                                     double* grad_real, double* grad_imag, const double* voxel,
                                     int64_t A, int64_t K, int64_t N, int64_t norm_count,
                                     const geraf_fmcw_params* params, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_mf_sizes(A, K, N);
  check_params(params);
  if (A == 0 || K == 0) return GERAF_RF_OK;
  require_ptrs({tx_pos, rx_pos, grad_real, grad_imag}, true);
  require_ptrs({grad_out_real, grad_out_imag, voxel}, N > 0);
  const long long total = A * K;
  hipLaunchKernelGGL(mf_backward_f64_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256),
                     0, (hipStream_t)stream, grad_out_real, grad_out_imag, tx_pos, rx_pos,
                     grad_real, grad_imag, voxel, (int)A, (int)K, (int)N,
                     1.0 / (double)norm_count, make_phase_consts(*params));
  return check_launch("mf_backward_f64_kernel");
  GERAF_TRY_END
}
