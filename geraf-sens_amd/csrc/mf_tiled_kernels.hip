// Fast matched filter by per-tile expansions: the operator of
// MatchedFilter_cuda_kernel.cu:65-115,
//   out[v] = (1/A) sum_a sum_k s[a,k] exp(-j 2 pi (f0 + k d) tau_av),  d = slope/fs,
// evaluated to a stated truncation bound with O(N_T) work per voxel-antenna
// pair instead of O(K).
//
// Voxels come in tiles of 512 or 128 (the caller orders them so that a tile
// is compact in space and picks the size: smaller tiles need fewer terms and
// cost more per-tile moments).  A tile with centre c and radius r sees every antenna's
// delays within H = 2 r / c of the centre delay tau_ac (triangle inequality).
// With k = k_c + kappa, k_c = (K-1)/2, u = (tau_av - tau_ac)/H in [-1, 1] and
// xi_k = kappa / k_c in [-1, 1]:
//   sum_k s_k e^{-j2pi(f0+kd)tau} = e^{-j2pi(f0 + k_c d) tau} sum_n C_n u^n,
//   C_n = (-j)^n Z^n / n! sum_k s_k e^{-j2pi kappa d tau_ac} xi_k^n,
//   Z = pi (K-1) d H,
// and truncating after N_T terms leaves at most sum_k |s_k| Z^N_T/N_T! e^Z
// (the host picks N_T from the largest tile).
//
// Kernels: per antenna chunk, mf_tile_moments_mma_kernel (the sums over k as
// matrix products: the xi^n table times the rotated echo of 16 tiles; the
// phasor e^{-j2pi kappa d tau_ac} is a rotation re-seeded every 8 samples
// from the FP64-reduced phase), then mf_tile_eval_kernel (block = tile, two
// voxels per lane, the tile's coefficients of one antenna are wave-uniform
// SGPR operands; per pair: FP64 distances and phase reduction, one phasor
// from the hardware sine and cosine, an N_T-term real Horner with complex
// coefficients, FP64 accumulation over the antennas), adding each chunk's
// sums into FP64 accumulators; a last kernel scales by 1/A.
//
// The adjoint (MatchedFilter_cuda_kernel.cu:171-220) is the transpose of the
// same expansion:
//   g[a,k] = (1/norm) sum_t e^{+j2pi kappa d tau_ac} sum_n (j Z xi_k)^n / n! M_n(a,t),
//   M_n(a,t) = sum_{v in t} gout_v e^{+j2pi(f0 + k_c d) tau_av} u_av^n,
// computed by mf_tile_adj_moments_kernel (one wave per tile, eight voxels per
// lane, wave sums of the moments by recursive halving) and
// mf_tile_adj_eval_mma_kernel (one wave per antenna and 64 frequencies: the
// sums over n as matrix products of the xi^n table and the tiles' Z^n M'_n,
// then the phasor-weighted sum over the tiles on the vector units).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <math.h>

#include <algorithm>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

constexpr int kTileVox = 512;  // the larger tile: 4 waves x 64 lanes x 2 voxels
constexpr int kEvalWaves = 4;
constexpr int kMomentWaves = 4;

struct TileGeom {
  double cx, cy, cz;  // centre
  double H;           // delay half-spread 2 r / c
};

__device__ __forceinline__ TileGeom tile_geom(const float4* tile_info, int t, double inv_c) {
  const float4 ti = tile_info[t];
  TileGeom g;
  g.cx = ti.x;
  g.cy = ti.y;
  g.cz = ti.z;
  g.H = 2.0 * (double)ti.w * inv_c;
  return g;
}

__global__ void fill_zero_kernel(double* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// The forward moments on the matrix cores: S_n(t) = sum_k xi_k^n b_k(t),
// b_k(t) = s_k e^{-j 2 pi (k - kc) d tau_c(t)}, for 16 tiles per wave and
// one antenna, on v_mfma_f32_16x16x32_f16 with FP16 hi/lo operands: A = the
// powers xi_k^n (16 moments x 32 frequencies, from a table in operand
// order), B = b (32 frequencies x 16 tiles; lane (j, g) forms b_k of tile j
// for k = 32 ks + 8 g .. + 7 from one phasor seed and seven rotations), the
// real and imaginary parts as two real products of three instructions: per
// pair (frequency, tile) ~15 vector instructions instead of the 8 + 2 N_T
// FMAs of a per-lane sum.  The echo row is scaled by a power of two (max |s|
// in [0.5, 1)) so FP16 keeps its small entries; the moments are scaled back.
constexpr int kMomTiles = 16;
constexpr int kMomChunk = 1024;  // echo samples staged in LDS at a time (8 KB)

__device__ __forceinline__ int moment_blocks(int nt) { return (nt + 15) / 16; }

// xi_ops[((ks * NB + blk) * 2 + h) * 64 + lane], NB = ceil(NT / 16): the A
// operand of frequency step ks and moment block blk, hi (h = 0) or lo;
// element e of lane (i, g) is xi_k^n for n = 16 blk + i, k = 32 ks + 8 g + e
// (0 past K or NT).
__global__ void xi_ops_kernel(h8* __restrict__ ops, int K, int NT, double kc) {
  const int NB = moment_blocks(NT);
  const int steps = (K + 31) / 32;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= steps * NB * kWave) return;
  const int lane = idx & (kWave - 1), blk = (idx / kWave) % NB, ks = idx / (kWave * NB);
  const int n = 16 * blk + (lane & 15);
  h8 hi, lo;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = 32 * ks + 8 * (lane >> 4) + e;
    double x = 0.0;
    if (k < K && n < NT) {
      const double xi = kc > 0.0 ? ((double)k - kc) / kc : 0.0;
      x = 1.0;
      for (int i = 0; i < n; ++i) x *= xi;
    }
    const _Float16 h = (_Float16)x;
    hi[e] = h;
    lo[e] = (_Float16)(x - (double)h);
  }
  ops[((size_t)(ks * NB + blk) * 2 + 0) * kWave + lane] = hi;
  ops[((size_t)(ks * NB + blk) * 2 + 1) * kWave + lane] = lo;
}

template <int NT>
__global__ __launch_bounds__(64 * kMomentWaves) void mf_tile_moments_mma_kernel(
    const float* __restrict__ real, const float* __restrict__ imag,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    const float4* __restrict__ tile_info, const h8* __restrict__ xi_ops, int a0, int na, int K,
    int ntiles, PhaseConsts pc, double kc, float2* __restrict__ coef,
    double* __restrict__ tau_c) {
  constexpr int NB = (NT + 15) / 16;
  __shared__ float4 echo[kMomChunk / 2];  // (re, im) pairs of the chunk, zero past K
  __shared__ float wmax[kMomentWaves];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int tg = blockIdx.x * kMomentWaves + wave;  // group of 16 tiles
  const int ai = blockIdx.y;                        // antenna within the chunk (< na)
  const int a = a0 + ai;
  const int j = lane & 15, g = lane >> 4;
  const int t = tg * kMomTiles + j;
  const bool tile_ok = t < ntiles;  // waves past the last tile still join the barriers
  const TileGeom geo = tile_geom(tile_info, tile_ok ? t : ntiles - 1, pc.inv_c);
  const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
  const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
  const double tac = (dist3_f64_fast(tx, ty, tz, geo.cx, geo.cy, geo.cz) +
                      dist3_f64_fast(rx, ry, rz, geo.cx, geo.cy, geo.cz) + pc.bias) *
                     pc.inv_c;
  const double step = pc.dfs * tac;  // cycles per sample at the centre delay
  const float2 w = cis_neg_reduced((float)centered_frac(step));  // one sample forward

  // the echo row's scale: max |s| of the row into [0.5, 1)
  const float* __restrict__ er = real + (size_t)a * K;
  const float* __restrict__ ei = imag + (size_t)a * K;
  float mx = 0.0f;
  for (int k = threadIdx.x; k < K; k += 64 * kMomentWaves)
    mx = fmaxf(mx, fmaxf(fabsf(er[k]), fabsf(ei[k])));
  mx = wave_max_dpp(mx);
  if (lane == 0) wmax[wave] = mx;
  __syncthreads();
#pragma unroll
  for (int w2 = 0; w2 < kMomentWaves; ++w2) mx = fmaxf(mx, wmax[w2]);
  int ex = 0;
  if (mx > 0.0f && isfinite(mx)) (void)frexpf(mx, &ex);
  ex = min(max(ex, -120), 120);  // both scales stay finite normal numbers
  const float down = ldexpf(1.0f, -ex);

  f32x4v Sr[NB], Si[NB];
#pragma unroll
  for (int blk = 0; blk < NB; ++blk) Sr[blk] = Si[blk] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
  for (int c0 = 0; c0 < K; c0 += kMomChunk) {
    const int steps = (min(kMomChunk, K - c0) + 31) / 32;
    if (c0 > 0) __syncthreads();  // the previous chunk's reads are done
    for (int i = threadIdx.x; i < 16 * steps; i += 64 * kMomentWaves) {
      const int k = c0 + 2 * i;
      float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      if (k < K) v.x = er[k], v.y = ei[k];
      if (k + 1 < K) v.z = er[k + 1], v.w = ei[k + 1];
      echo[i] = v;
    }
    __syncthreads();
    for (int ks = 0; ks < steps; ++ks) {
      const int kl = 32 * ks + 8 * g;  // first sample of the lane, within the chunk
      // e^{-j 2 pi (k - kc) d tau_c} at k = c0 + kl from the FP64-reduced
      // phase, times the row's scale
      float2 ph = phasor_cycles_hw((float)centered_frac((kc - (double)(c0 + kl)) * step));
      ph = make_float2(ph.x * down, ph.y * down);
      float2 b[8];
#pragma unroll
      for (int e = 0; e < 8; e += 2) {
        const float4 v = echo[kl / 2 + e / 2];
        b[e] = make_float2(fmaf(v.x, ph.x, -v.y * ph.y), fmaf(v.x, ph.y, v.y * ph.x));
        ph = cmul(ph, w);
        b[e + 1] = make_float2(fmaf(v.z, ph.x, -v.w * ph.y), fmaf(v.z, ph.y, v.w * ph.x));
        ph = cmul(ph, w);
      }
      h8 brh, brl, bih, bil;
      split8(b, brh, brl, bih, bil);
      const int kg = c0 / 32 + ks;  // frequency step of the whole row
#pragma unroll
      for (int blk = 0; blk < NB; ++blk) {
        const h8 ah = xi_ops[((size_t)(kg * NB + blk) * 2 + 0) * kWave + lane];
        const h8 al = xi_ops[((size_t)(kg * NB + blk) * 2 + 1) * kWave + lane];
        Sr[blk] = mma3(ah, al, brh, brl, Sr[blk]);
        Si[blk] = mma3(ah, al, bih, bil, Si[blk]);
      }
    }
  }
  if (!tile_ok) return;
  // lane (j, g) holds S_n of tile j for n = 16 blk + 4 g + r.
  // C_n = (-j)^n Z^n / n! S_n, times the row's scale back
  const float Z = (float)(3.14159265358979323846 * (double)(K - 1) * pc.dfs * geo.H);
  const float up = ldexpf(1.0f, ex);
  float sc[NB][4];
  {
    float cur = up;  // up Z^n / n!, n = 0, 1, ...
#pragma unroll
    for (int n = 0; n < 16 * NB; ++n) {
      if (((n >> 2) & 3) == g) sc[n >> 4][n & 3] = cur;
      cur *= Z / (float)(n + 1);
    }
  }
  float2* __restrict__ out = coef + ((size_t)ai * ntiles + t) * NT;
#pragma unroll
  for (int blk = 0; blk < NB; ++blk)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 16 * blk + 4 * g + r;
      if (n < NT) {
        const float xr = Sr[blk][r] * sc[blk][r], xi = Si[blk][r] * sc[blk][r];
        float2 c;
        switch (r) {  // n & 3 == r
          case 0: c = make_float2(xr, xi); break;
          case 1: c = make_float2(xi, -xr); break;
          case 2: c = make_float2(-xr, -xi); break;
          default: c = make_float2(-xi, xr); break;
        }
        out[n] = c;
      }
    }
  if (g == 0) tau_c[(size_t)ai * ntiles + t] = tac;
}

// TW waves per tile (tiles of 128 TW voxels); a block of kEvalWaves waves
// takes kEvalWaves / TW tiles.
template <int NT, int TW>
__global__ __launch_bounds__(64 * kEvalWaves) void mf_tile_eval_kernel(
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    const float* __restrict__ voxel, const float4* __restrict__ tile_info,
    const float2* __restrict__ coef, const double* __restrict__ tau_c, int a0, int na, int N,
    int ntiles, PhaseConsts pc, double fcen, double* __restrict__ acc_re,
    double* __restrict__ acc_im, int per_group, size_t acc_stride) {
  constexpr int VPL = 2;
  constexpr int TV = kWave * VPL * TW;  // voxels per tile
  // antenna group of this block: antennas [ai_lo, ai_hi) of the chunk, into
  // its own accumulator slice
  const int ai_lo = blockIdx.y * per_group;
  const int ai_hi = min(na, ai_lo + per_group);
  acc_re += blockIdx.y * acc_stride;
  acc_im += blockIdx.y * acc_stride;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = blockIdx.x * (kEvalWaves / TW) + wave / TW;
  if (t >= ntiles) return;  // wave-uniform; no block-level synchronisation here
  const int wt = wave % TW;  // wave within the tile
  const TileGeom g = tile_geom(tile_info, t, pc.inv_c);
  const double invH = g.H > 0.0 ? 1.0 / g.H : 0.0;

  int v[VPL];
  bool valid[VPL];
  double px[VPL], py[VPL], pz[VPL];
#pragma unroll
  for (int i = 0; i < VPL; ++i) {
    v[i] = t * TV + (wt * VPL + i) * kWave + lane;
    valid[i] = v[i] < N;
    const int vc = valid[i] ? v[i] : N - 1;
    px[i] = voxel[vc];
    py[i] = voxel[(size_t)N + vc];
    pz[i] = voxel[2 * (size_t)N + vc];
  }
  double acc_r[VPL], acc_i[VPL];  // FP64 sums over the antennas
#pragma unroll
  for (int i = 0; i < VPL; ++i) acc_r[i] = acc_i[i] = 0.0;

  for (int ai = ai_lo; ai < ai_hi; ++ai) {
    const int a = a0 + ai;
    const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
    const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
    const size_t ct = __builtin_amdgcn_readfirstlane((unsigned)ai) * (size_t)ntiles + t;
    const double tac = tau_c[ct];
    const float2* __restrict__ cf = coef + ct * NT;
    auto pairs = [&](auto mono) {
#pragma unroll
      for (int i = 0; i < VPL; ++i) {
        const double tau = (round_trip_f64<decltype(mono)::value>(tx, ty, tz, rx, ry, rz, px[i],
                                                                  py[i], pz[i]) +
                            pc.bias) *
                           pc.inv_c;
        const float u = (float)((tau - tac) * invH);
        float Pr = cf[NT - 1].x, Pi = cf[NT - 1].y;
#pragma unroll
        for (int n = NT - 2; n >= 0; --n) {
          Pr = fmaf(Pr, u, cf[n].x);
          Pi = fmaf(Pi, u, cf[n].y);
        }
        const float2 cs = phasor_cycles_hw((float)centered_frac(fcen * tau));
        const float2 t = cmul(make_float2(cs.x, -cs.y), make_float2(Pr, Pi));
        acc_r[i] += (double)t.x;
        acc_i[i] += (double)t.y;
      }
    };
    if (monostatic(tx_pos, rx_pos, a))
      pairs(std::true_type{});
    else
      pairs(std::false_type{});
  }
#pragma unroll
  for (int i = 0; i < VPL; ++i) {
    if (valid[i]) {
      acc_re[v[i]] += acc_r[i];
      acc_im[v[i]] += acc_i[i];
    }
  }
}

__global__ void mf_tile_finish_kernel(const double* __restrict__ acc_re,
                                      const double* __restrict__ acc_im,
                                      float* __restrict__ out_re, float* __restrict__ out_im,
                                      int N, double inv_a, int groups, size_t acc_stride) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= N) return;
  double r = 0.0, i = 0.0;
  for (int g = 0; g < groups; ++g) {
    r += acc_re[g * acc_stride + v];
    i += acc_im[g * acc_stride + v];
  }
  out_re[v] = (float)(r * inv_a);
  out_im[v] = (float)(i * inv_a);
}

// ------------------------------------------------------------------ adjoint

// Wave totals of moment pair q (moments 2q, 2q + 1, re and im) by recursive
// halving: a permlane32 swap joins (Pr_m, Pi_m) in one register (lanes < 32
// a partial of Pr_m over lanes i, i + 32; lanes >= 32 of Pi_m), a permlane16
// swap joins two of those (row 0: Pr_2q, row 1: Pr_2q+1, row 2: Pi_2q, row
// 3: Pi_2q+1, each over 4 lanes), and four DPP steps sum each row: about 10
// instructions for the 4 totals, where a full-wave reduction per value
// takes about 11.  Every lane of row rho ends with its row's total.
__device__ __forceinline__ float wave_moment_pair_total(float pr0, float pi0, float pr1,
                                                        float pi1) {
  const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(pr0), __float_as_uint(pi0),
                                                  false, false);
  const auto b = __builtin_amdgcn_permlane32_swap(__float_as_uint(pr1), __float_as_uint(pi1),
                                                  false, false);
  const float h0 = __uint_as_float(a[0]) + __uint_as_float(a[1]);
  const float h1 = __uint_as_float(b[0]) + __uint_as_float(b[1]);
  const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(h0), __float_as_uint(h1),
                                                  false, false);
  float v = __uint_as_float(r[0]) + __uint_as_float(r[1]);
  v += dpp_mov<dpp::kQuadXor1>(0.0f, v);
  v += dpp_mov<dpp::kQuadXor2>(0.0f, v);
  v += dpp_mov<dpp::kRowHalfMirror>(0.0f, v);
  v += dpp_mov<dpp::kRowMirror>(0.0f, v);
  return v;
}

constexpr float inv_factorial(int n) {
  double f = 1.0;
  for (int i = 2; i <= n; ++i) f *= i;
  return (float)(1.0 / f);
}

// One wave per tile of kWave * AVPL voxels (AVPL voxels per lane).
template <int NT, int AVPL>
__global__ __launch_bounds__(256) void mf_tile_adj_moments_kernel(
    const float* __restrict__ gout_re, const float* __restrict__ gout_im,
    const float* __restrict__ tx_pos, const float* __restrict__ rx_pos,
    const float* __restrict__ voxel, const float4* __restrict__ tile_info, int a0, int na,
    int N, int ntiles, PhaseConsts pc, double fcen, float2* __restrict__ mom,
    double* __restrict__ tau_c, int per_group) {
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int t = blockIdx.x * 4 + wave;
  const int ai_lo = blockIdx.y * per_group;
  const int ai_hi = min(na, ai_lo + per_group);
  if (t >= ntiles) return;  // wave-uniform
  const TileGeom g = tile_geom(tile_info, t, pc.inv_c);
  const double invH = g.H > 0.0 ? 1.0 / g.H : 0.0;
  double px[AVPL], py[AVPL], pz[AVPL];
  float cr[AVPL], ci[AVPL];
#pragma unroll
  for (int i = 0; i < AVPL; ++i) {
    const int v = t * (kWave * AVPL) + i * kWave + lane;
    const bool ok = v < N;
    const int vc = ok ? v : N - 1;
    px[i] = voxel[vc];
    py[i] = voxel[(size_t)N + vc];
    pz[i] = voxel[2 * (size_t)N + vc];
    cr[i] = ok ? gout_re[vc] : 0.0f;
    ci[i] = ok ? gout_im[vc] : 0.0f;
  }
  for (int ai = ai_lo; ai < ai_hi; ++ai) {
    const int a = a0 + ai;
    const double tx = tx_pos[3 * a], ty = tx_pos[3 * a + 1], tz = tx_pos[3 * a + 2];
    const double rx = rx_pos[3 * a], ry = rx_pos[3 * a + 1], rz = rx_pos[3 * a + 2];
    const double tac = (dist3_f64_fast(tx, ty, tz, g.cx, g.cy, g.cz) +
                        dist3_f64_fast(rx, ry, rz, g.cx, g.cy, g.cz) + pc.bias) *
                       pc.inv_c;
    float Pr[NT], Pi[NT];
#pragma unroll
    for (int n = 0; n < NT; ++n) Pr[n] = Pi[n] = 0.0f;
    auto pairs = [&](auto mono) {
#pragma unroll
      for (int i = 0; i < AVPL; ++i) {
        const double tau = (round_trip_f64<decltype(mono)::value>(tx, ty, tz, rx, ry, rz, px[i],
                                                                  py[i], pz[i]) +
                            pc.bias) *
                           pc.inv_c;
        const float u = (float)((tau - tac) * invH);
        const float2 p = cmul(make_float2(cr[i], ci[i]),
                              phasor_cycles_hw((float)centered_frac(fcen * tau)));
        float q = 1.0f;  // u^n
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          Pr[n] = fmaf(p.x, q, Pr[n]);
          Pi[n] = fmaf(p.y, q, Pi[n]);
          q *= u;
        }
      }
    };
    if (monostatic(tx_pos, rx_pos, a))
      pairs(std::true_type{});
    else
      pairs(std::false_type{});
    // M'_n = j^n / n! * sum over the tile.  The first lane of row rho holds
    // the total of component rho >> 1 (re, im) of moment 2 q + (rho & 1); as
    // j^k, k = n (+1 for im), it goes to component k & 1, negated for k & 2.
    static_assert(NT % 2 == 0, "moment count must be even");
    const int rho = lane >> 4;
    float* __restrict__ out = reinterpret_cast<float*>(mom + ((size_t)ai * ntiles + t) * NT);
#pragma unroll
    for (int q = 0; q < NT / 2; ++q) {
      const float tot = wave_moment_pair_total(Pr[2 * q], Pi[2 * q], Pr[2 * q + 1], Pi[2 * q + 1]);
      const int n = 2 * q + (rho & 1);
      const int k = n + (rho >> 1);
      const float f = (rho & 1) ? inv_factorial(2 * q + 1) : inv_factorial(2 * q);
      const float val = (k & 2) ? -tot * f : tot * f;
      if ((lane & 15) == 0) out[2 * n + (k & 1)] = val;
    }
    if (lane == 0) tau_c[(size_t)ai * ntiles + t] = tac;
  }
}

// The adjoint evaluation on the matrix cores.  Per antenna and tile t,
//   Q[k, t] = sum_n xi_k^n Z_t^n M'_n(a, t)      (M'_n = j^n / n! M_n),
// a real (frequencies x N_T) by complex (N_T x tiles) product on
// v_mfma_f32_16x16x32_f16 with FP16 hi/lo operands (contraction over the
// terms, zero-padded to 32): A = xi_k^n from a table in operand order, B =
// the tile's scaled moments, built per 16 tiles (lane (j, g): tile j, terms
// 8 g .. 8 g + 7), each tile's column scaled by a power of two (its largest
// entry in [0.5, 1)) and scaled back through the phasor.  Then
//   g[a, k] = sum_t e^{+j 2 pi (k - kc) d tau_c(t)} Q[k, t]
// on the vector units: lane (j, q) holds k = k_w + 16 kb + 4 q + r of tile j
// (kb, r = 0..3) and rotates one phasor seed per tile along those 16
// frequencies; products are summed in FP32 over kAdjFlush tile blocks and
// then in FP64, and the 16 lanes of a row (its tiles) are summed at the end.
// One wave per (antenna, 64 frequencies); the waves of a block (up to four)
// share an antenna and each tile block's B fragments and phase steps, built
// once per block (wave w of W builds tile blocks w, w + W, ...) in
// double-buffered LDS.
constexpr int kAdjKB = 4;      // 16-frequency blocks per wave (one per 4 lanes of a row)
constexpr int kAdjFlush = 8;   // tile blocks summed in FP32 before each FP64 add

// xi_adj_ops[(kb * 2 + h) * 64 + lane]: the A operand of frequency block kb,
// hi (h = 0) or lo; element e of lane (i, g) is xi_k^n for k = 16 kb + i,
// n = 8 g + e (0 past K or NT).
__global__ void xi_adj_ops_kernel(h8* __restrict__ ops, int K, int NT, double kc) {
  const int nkb = (K + 15) / 16;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nkb * kWave) return;
  const int lane = idx & (kWave - 1), kb = idx / kWave;
  const int k = 16 * kb + (lane & 15);
  const double xi = kc > 0.0 ? ((double)k - kc) / kc : 0.0;
  h8 hi, lo;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int n = 8 * (lane >> 4) + e;
    double x = 0.0;
    if (k < K && n < NT) {
      x = 1.0;
      for (int i = 0; i < n; ++i) x *= xi;
    }
    const _Float16 h = (_Float16)x;
    hi[e] = h;
    lo[e] = (_Float16)(x - (double)h);
  }
  ops[(size_t)(kb * 2 + 0) * kWave + lane] = hi;
  ops[(size_t)(kb * 2 + 1) * kWave + lane] = lo;
}

// Maximum of the four lanes j, j + 16, j + 32, j + 48, in each of them.
__device__ __forceinline__ float column_max4(float x) {
  const auto s32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x),
                                                    false, false);
  x = fmaxf(__uint_as_float(s32[0]), __uint_as_float(s32[1]));
  const auto s16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x),
                                                    false, false);
  return fmaxf(__uint_as_float(s16[0]), __uint_as_float(s16[1]));
}

// One tile block's operands, shared by the block's four waves: the B
// fragments (hi and lo of the real and imaginary parts, per lane), and per
// tile the phase step d tau_c, the one- and sixteen-sample rotations and the
// power of two that undoes the column's scale.
struct AdjTileOps {
  h8 b[4][kWave];
  double step[16];
  float2 w1[16], w16[16];
  float up[16];
};
constexpr int kAdjWaves = 4;  // waves per block: one antenna, 4 x 64 frequencies

template <int NT>
__global__ __launch_bounds__(64 * kAdjWaves) void mf_tile_adj_eval_mma_kernel(
    const float2* __restrict__ mom, const double* __restrict__ tau_c,
    const float4* __restrict__ tile_info, const h8* __restrict__ xi_ops, int a0, int na, int K,
    int ntiles, PhaseConsts pc, double kc, double inv_norm, float* __restrict__ grad_re,
    float* __restrict__ grad_im) {
  static_assert(NT <= 32, "one contraction step of 32 terms");
  static_assert(kAdjKB == 4, "the 16 lanes of a row write the 16 frequencies (kb, r)");
  // [stage][tile block of the group]: wave w builds entry w of the next
  // stage while all four read the current one
  __shared__ AdjTileOps ops[2][kAdjWaves];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int W = blockDim.x / kWave;  // waves of the block, 1..kAdjWaves
  const int nkw = (K + 16 * kAdjKB - 1) / (16 * kAdjKB);  // 64-frequency ranges
  const int nbk = (nkw + W - 1) / W;                     // blocks per antenna
  const int ai = blockIdx.x / nbk;                       // < na (grid = na * nbk)
  const int kw = ((blockIdx.x - ai * nbk) * W + wave) * 16 * kAdjKB;
  const bool active = kw < K;  // wave-uniform; idle waves still build and join barriers
  const int a = a0 + ai;
  const int j = lane & 15, q = lane >> 4;
  const int nkb = (K + 15) / 16;
  h8 ah[kAdjKB], al[kAdjKB];
#pragma unroll
  for (int kb = 0; kb < kAdjKB; ++kb) {
    const int b = kw / 16 + kb;  // uniform
    if (b < nkb) {
      ah[kb] = xi_ops[(size_t)(b * 2 + 0) * kWave + lane];
      al[kb] = xi_ops[(size_t)(b * 2 + 1) * kWave + lane];
    } else {
      ah[kb] = al[kb] = h8{0, 0, 0, 0, 0, 0, 0, 0};
    }
  }
  const double zfac = 3.14159265358979323846 * (double)(K - 1) * pc.dfs;
  const double k0 = (double)(kw + 4 * q) - kc;  // the lane's first frequency, centred
  const int ntb = (ntiles + 15) / 16;
  const int ngroups = (ntb + W - 1) / W;  // W tile blocks per group

  // the inputs of tile block tb (tiles past the last one clamped; their B
  // columns are zeroed in build)
  auto load = [&](int tb, float2 (&m)[8], float& rad, double& tac) {
    const int t = min(tb * 16 + j, ntiles - 1);
    const size_t ct = (size_t)ai * ntiles + t;
#pragma unroll
    for (int e = 0; e < 8; ++e) m[e] = mom[ct * NT + min(8 * q + e, NT - 1)];
    rad = tile_info[t].w;
    tac = tau_c[ct];
  };
  // B = Z^n M'_n for n = 8 q + e (each tile's column scaled into [0.5, 1)),
  // and the tile's phase steps
  auto build = [&](int tb, const float2 (&m)[8], float rad, double tac, AdjTileOps& o) {
    const bool ok = tb * 16 + j < ntiles;
    float2 v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = (ok && 8 * q + e < NT) ? m[e] : make_float2(0.0f, 0.0f);
    const float Z = (float)(zfac * 2.0 * (double)rad * pc.inv_c);
    const float z2 = Z * Z, z4 = z2 * z2, z8 = z4 * z4;
    float zp = q == 0 ? 1.0f : (q == 1 ? z8 : (q == 2 ? z8 * z8 : z8 * z8 * z8));
    float mx = 0.0f;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      v[e] = make_float2(v[e].x * zp, v[e].y * zp);
      mx = fmaxf(mx, fmaxf(fabsf(v[e].x), fabsf(v[e].y)));
      zp *= Z;
    }
    mx = column_max4(mx);
    int ex = 0;
    if (mx > 0.0f && isfinite(mx)) (void)frexpf(mx, &ex);
    ex = min(max(ex, -120), 120);
    const float down = ldexpf(1.0f, -ex);
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = make_float2(v[e].x * down, v[e].y * down);
    h8 brh, brl, bih, bil;
    split8(v, brh, brl, bih, bil);
    o.b[0][lane] = brh;
    o.b[1][lane] = brl;
    o.b[2][lane] = bih;
    o.b[3][lane] = bil;
    if (q == 0) {
      // polynomial sine and cosine: these phasors set the accuracy of all
      // the products of the tile
      const double step = pc.dfs * tac;
      o.step[j] = step;
      o.w1[j] = phasor_cycles((float)centered_frac(step), 1.0f);
      o.w16[j] = phasor_cycles((float)centered_frac(16.0 * step), 1.0f);
      o.up[j] = ldexpf(1.0f, ex);
    }
  };

  float2 part[kAdjKB][4];
  double accr[kAdjKB][4], acci[kAdjKB][4];
#pragma unroll
  for (int kb = 0; kb < kAdjKB; ++kb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      part[kb][r] = make_float2(0.0f, 0.0f);
      accr[kb][r] = acci[kb][r] = 0.0;
    }
  float2 nm[8];
  float nrad;
  double ntac;
  load(wave, nm, nrad, ntac);  // group 0's block of this wave
  build(wave, nm, nrad, ntac, ops[0][wave]);
  if (W + wave < ntb) load(W + wave, nm, nrad, ntac);  // group 1's
  __syncthreads();
  for (int g = 0; g < ngroups; ++g) {
    const int cur = g & 1;
    if (g + 1 < ngroups) {
      // stage cur ^ 1 was last read before the previous barrier
      const int tbn = (g + 1) * W + wave;
      build(tbn, nm, nrad, ntac, ops[cur ^ 1][wave]);
      if (tbn + W < ntb) load(tbn + W, nm, nrad, ntac);  // in flight
    }
    if (active) {
#pragma unroll 1
      for (int i = 0; i < W; ++i) {
        const int tb = g * W + i;  // uniform
        if (tb >= ntb) break;
        const AdjTileOps& o = ops[cur][i];
        const h8 brh = o.b[0][lane], brl = o.b[1][lane], bih = o.b[2][lane],
                 bil = o.b[3][lane];
        const float2 w1 = o.w1[j], w16 = o.w16[j];
        const float up = o.up[j];
        float2 ph = phasor_cycles((float)centered_frac(k0 * o.step[j]), 1.0f);
        ph = make_float2(ph.x * up, ph.y * up);
#pragma unroll
        for (int kb = 0; kb < kAdjKB; ++kb) {
          const f32x4v zero = {0.0f, 0.0f, 0.0f, 0.0f};
          const f32x4v qr = mma3(ah[kb], al[kb], brh, brl, zero);
          const f32x4v qi = mma3(ah[kb], al[kb], bih, bil, zero);
          float2 p = ph;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            part[kb][r].x = fmaf(p.x, qr[r], fmaf(-p.y, qi[r], part[kb][r].x));
            part[kb][r].y = fmaf(p.x, qi[r], fmaf(p.y, qr[r], part[kb][r].y));
            if (r < 3) p = cmul(p, w1);
          }
          if (kb + 1 < kAdjKB) ph = cmul(ph, w16);
        }
        if ((tb + 1) % kAdjFlush == 0 || tb + 1 == ntb) {
#pragma unroll
          for (int kb = 0; kb < kAdjKB; ++kb)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              accr[kb][r] += (double)part[kb][r].x;
              acci[kb][r] += (double)part[kb][r].y;
              part[kb][r] = make_float2(0.0f, 0.0f);
            }
        }
      }
    }
    __syncthreads();  // stage cur is read; stage cur ^ 1 is built
  }
  if (!active) return;  // after the last barrier
  // sum the row's 16 lanes (tiles j); lane j of row q then writes
  // k = k_w + 16 (j >> 2) + 4 q + (j & 3)
  double outr = 0.0, outi = 0.0;
#pragma unroll
  for (int kb = 0; kb < kAdjKB; ++kb)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      double sr = accr[kb][r], si = acci[kb][r];
#pragma unroll
      for (int m = 1; m < 16; m <<= 1) {
        sr += __shfl_xor(sr, m, kWave);
        si += __shfl_xor(si, m, kWave);
      }
      if (j == 4 * kb + r) outr = sr, outi = si;
    }
  const int k = kw + 16 * (j >> 2) + 4 * q + (j & 3);
  if (k < K) {
    grad_re[(size_t)a * K + k] = (float)(outr * inv_norm);
    grad_im[(size_t)a * K + k] = (float)(outi * inv_norm);
  }
}

template <int NT>
void launch_adj_chunk(const float* gre, const float* gim, const float* tx, const float* rx,
                      const float* voxel, const float4* tile_info, const h8* xi_ops, int a0,
                      int na, int K, int N, int ntiles, int tv, const PhaseConsts& pc, double kc,
                      double inv_norm, float2* mom, double* tau_c, float* grad_re,
                      float* grad_im, hipStream_t stream) {
  const double fcen = pc.f0 + kc * pc.dfs;
  // antenna ranges side by side when the tiles alone give fewer than ~32
  // waves per SIMD (one wave per tile)
  int g = 1;
  while (g < 64 && g * 2 <= na && (int64_t)ntiles * g < 32768) g *= 2;
  const int per_group = (na + g - 1) / g;
  const dim3 mgrid((unsigned)((ntiles + 3) / 4), (unsigned)g);
  if (tv == 512)
    hipLaunchKernelGGL((mf_tile_adj_moments_kernel<NT, 8>), mgrid, dim3(256), 0, stream, gre,
                       gim, tx, rx, voxel, tile_info, a0, na, N, ntiles, pc, fcen, mom, tau_c,
                       per_group);
  else
    hipLaunchKernelGGL((mf_tile_adj_moments_kernel<NT, 2>), mgrid, dim3(256), 0, stream, gre,
                       gim, tx, rx, voxel, tile_info, a0, na, N, ntiles, pc, fcen, mom, tau_c,
                       per_group);
  // up to four waves per block, all on one antenna (fewer when K < 256)
  const int nkw = (K + 16 * kAdjKB - 1) / (16 * kAdjKB);
  const int W = std::min(kAdjWaves, nkw);
  const int64_t blocks = (int64_t)na * ((nkw + W - 1) / W);
  hipLaunchKernelGGL((mf_tile_adj_eval_mma_kernel<NT>), dim3((unsigned)blocks),
                     dim3(64 * W), 0, stream, mom, tau_c, tile_info, xi_ops, a0, na, K,
                     ntiles, pc, kc, inv_norm, grad_re, grad_im);
}

// Instantiated term counts; a request is rounded up to the next one.
constexpr int kTermSets[] = {4, 6, 8, 10, 12, 14, 16, 20, 24, 32};

int round_terms(int nterms) {
  for (int t : kTermSets)
    if (nterms <= t) return t;
  return -1;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Antenna groups run side by side when the tiles alone would give fewer than
// ~16 waves per SIMD (4 waves per tile block), so that two waves are ready to
// dual-issue most of the time; each group has its own FP64 accumulator slice.
int antenna_groups(int64_t ntiles) {
  const int64_t want = 4096;  // tile blocks of 4 waves
  int g = 1;
  while (g < 16 && ntiles * g < want) g *= 2;
  return g;
}

struct TiledLayout {
  size_t acc_bytes, xi_bytes, per_antenna, fixed;
  int groups;
};

TiledLayout tiled_layout(int64_t K, int64_t N, int NT, int tv) {
  const int64_t ntiles = (N + tv - 1) / tv;
  TiledLayout l;
  // evaluation blocks: one per 512-voxel tile, one per four 128-voxel tiles
  l.groups = antenna_groups(tv == 512 ? ntiles : (ntiles + 3) / 4);
  l.acc_bytes = align256((size_t)N * sizeof(double)) * l.groups;
  // the xi^n table in matrix-operand order (hi and lo per step and block)
  // (the forward's per 32 frequencies and 16 terms, the adjoint's per 16 frequencies)
  l.xi_bytes = align256((size_t)std::max<int64_t>((K + 31) / 32 * ((NT + 15) / 16), (K + 15) / 16) *
                        2 * kWave * sizeof(h8));
  l.per_antenna = (size_t)ntiles * (NT * sizeof(float2) + sizeof(double));
  l.fixed = 2 * l.acc_bytes + l.xi_bytes + 2 * 256;
  return l;
}

template <int NT>
void launch_chunk(const float* real, const float* imag, const float* tx, const float* rx,
                  const float* voxel, const float4* tile_info, const h8* xi_ops, int a0,
                  int na, int K, int N, int ntiles, int tv, const PhaseConsts& pc, double kc,
                  float2* coef, double* tau_c, double* acc_re, double* acc_im, int groups,
                  size_t acc_stride, hipStream_t stream) {
  const int groups16 = (ntiles + kMomTiles - 1) / kMomTiles;
  const dim3 mgrid((unsigned)((groups16 + kMomentWaves - 1) / kMomentWaves), (unsigned)na);
  hipLaunchKernelGGL((mf_tile_moments_mma_kernel<NT>), mgrid, dim3(64 * kMomentWaves), 0,
                     stream, real, imag, tx, rx, tile_info, xi_ops, a0, na, K, ntiles, pc, kc,
                     coef, tau_c);
  const double fcen = pc.f0 + kc * pc.dfs;
  const int g = std::min(groups, na);
  const int per_group = (na + g - 1) / g;
  if (tv == 512)
    hipLaunchKernelGGL((mf_tile_eval_kernel<NT, 4>), dim3((unsigned)ntiles, (unsigned)g),
                       dim3(64 * kEvalWaves), 0, stream, tx, rx, voxel, tile_info, coef, tau_c,
                       a0, na, N, ntiles, pc, fcen, acc_re, acc_im, per_group, acc_stride);
  else
    hipLaunchKernelGGL((mf_tile_eval_kernel<NT, 1>),
                       dim3((unsigned)((ntiles + kEvalWaves - 1) / kEvalWaves), (unsigned)g),
                       dim3(64 * kEvalWaves), 0, stream, tx, rx, voxel, tile_info, coef, tau_c,
                       a0, na, N, ntiles, pc, fcen, acc_re, acc_im, per_group, acc_stride);
}

}  // namespace
}  // namespace geraf

using namespace geraf;

namespace geraf {
namespace {
// Per-chunk buffers of the tiled kernels inside the caller's workspace.
struct TiledBuffers {
  double* acc_re;
  double* acc_im;
  h8* xi_ops;
  float2* coef;
  double* tau_c;
  int na_max;
};

TiledBuffers carve(void* workspace, int64_t workspace_bytes, const TiledLayout& l, int64_t A,
                   int ntiles, int NT) {
  if (workspace_bytes < 0 || (size_t)workspace_bytes < l.fixed + 2 * 256 + l.per_antenna)
    throw ArgError("workspace too small (see geraf_mf_tiled_workspace_bytes)");
  const size_t base = align256((size_t)workspace) - (size_t)workspace;
  char* ws = static_cast<char*>(workspace) + base;
  TiledBuffers b;
  b.acc_re = reinterpret_cast<double*>(ws);
  b.acc_im = reinterpret_cast<double*>(ws + l.acc_bytes);  // l.acc_bytes covers all groups
  b.xi_ops = reinterpret_cast<h8*>(ws + 2 * l.acc_bytes);
  char* chunk = ws + 2 * l.acc_bytes + l.xi_bytes;
  const size_t avail = (size_t)workspace_bytes - base - 2 * l.acc_bytes - l.xi_bytes - 256;
  b.na_max = (int)std::min<int64_t>(A, (int64_t)(avail / l.per_antenna));
  if (b.na_max < 1) throw ArgError("workspace too small (see geraf_mf_tiled_workspace_bytes)");
  b.coef = reinterpret_cast<float2*>(chunk);
  b.tau_c = reinterpret_cast<double*>(
      chunk + align256((size_t)b.na_max * ntiles * NT * sizeof(float2)));
  if ((char*)(b.tau_c + (size_t)b.na_max * ntiles) >
      static_cast<char*>(workspace) + workspace_bytes)
    throw ArgError("workspace too small (see geraf_mf_tiled_workspace_bytes)");
  return b;
}
}  // namespace
}  // namespace geraf

namespace geraf {
namespace {
// Tile sizes the kernels are instantiated for.
bool tile_size_ok(int32_t tv) { return tv == 128 || tv == kTileVox; }
}  // namespace
}  // namespace geraf

extern "C" int64_t geraf_mf_tiled_workspace_bytes(int64_t A, int64_t K, int64_t N,
                                                  int32_t nterms, int32_t tile_voxels) {
  const int NT = round_terms(nterms);
  if (A < 0 || K < 0 || N < 0 || NT < 0 || !tile_size_ok(tile_voxels)) return -1;
  const TiledLayout l = tiled_layout(K, N, NT, tile_voxels);
  return (int64_t)(l.fixed + 2 * 256 + (size_t)A * l.per_antenna);
}

extern "C" int geraf_mf_forward_tiled_f32(const float* real, const float* imag,
                                          const float* tx_pos, const float* rx_pos,
                                          float* out_real, float* out_imag, const float* voxel,
                                          const float* tile_info, int64_t A, int64_t K,
                                          int64_t N, int32_t nterms, int32_t tile_voxels,
                                          void* workspace, int64_t workspace_bytes,
                                          const geraf_fmcw_params* params,
                                          geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_params(params);
  if (A < 1 || K < 1 || N < 0) throw ArgError("need A >= 1, K >= 1, N >= 0");
  if (!tile_size_ok(tile_voxels)) throw ArgError("tile_voxels must be 128 or 512");
  if (N == 0) return GERAF_RF_OK;
  if (A * K >= (INT64_C(1) << 31) || N >= (INT64_C(1) << 31) - kTileVox)
    throw ArgError("A*K and N must be < 2^31");
  const int NT = round_terms(nterms);
  if (nterms < 1 || NT < 0) throw ArgError("nterms must be in [1, 32]");
  require_ptrs({real, imag, tx_pos, rx_pos, out_real, out_imag, voxel, tile_info, workspace},
               true);
  const int tv = tile_voxels;
  const TiledLayout l = tiled_layout(K, N, NT, tv);
  const int ntiles = (int)((N + tv - 1) / tv);
  const TiledBuffers b = carve(workspace, workspace_bytes, l, A, ntiles, NT);
  double* acc_re = b.acc_re;
  double* acc_im = b.acc_im;
  h8* xi_ops = b.xi_ops;
  float2* coef = b.coef;
  double* tau_c = b.tau_c;
  const int na_max = b.na_max;

  const hipStream_t st = (hipStream_t)stream;
  const PhaseConsts pc = make_phase_consts(*params);
  const double kc = 0.5 * (double)(K - 1);
  const size_t acc_stride = l.acc_bytes / l.groups / sizeof(double);  // doubles per slice
  hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)((2 * l.acc_bytes / 8 + 255) / 256)),
                     dim3(256), 0, st, acc_re, (long long)(2 * l.acc_bytes / 8));
  {
    const int ops = (int)((K + 31) / 32) * ((NT + 15) / 16) * kWave;
    hipLaunchKernelGGL(xi_ops_kernel, dim3((unsigned)((ops + 255) / 256)), dim3(256), 0, st,
                       xi_ops, (int)K, NT, kc);
  }
  const float4* ti = reinterpret_cast<const float4*>(tile_info);
  for (int a0 = 0; a0 < A; a0 += na_max) {
    const int na = (int)std::min<int64_t>(na_max, A - a0);
#define GERAF_TILED_CASE(T)                                                                    \
  case T:                                                                                      \
    launch_chunk<T>(real, imag, tx_pos, rx_pos, voxel, ti, xi_ops, a0, na, (int)K, (int)N,     \
                    ntiles, tv, pc, kc, coef, tau_c, acc_re, acc_im, l.groups, acc_stride,     \
                    st);                                                                       \
    break;
    switch (NT) {
      GERAF_TILED_CASE(4)
      GERAF_TILED_CASE(6)
      GERAF_TILED_CASE(8)
      GERAF_TILED_CASE(10)
      GERAF_TILED_CASE(12)
      GERAF_TILED_CASE(14)
      GERAF_TILED_CASE(16)
      GERAF_TILED_CASE(20)
      GERAF_TILED_CASE(24)
      GERAF_TILED_CASE(32)
      default: throw ArgError("unsupported term count");
    }
#undef GERAF_TILED_CASE
  }
  hipLaunchKernelGGL(mf_tile_finish_kernel, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st,
                     acc_re, acc_im, out_real, out_imag, (int)N, 1.0 / (double)A,
                     l.groups, acc_stride);
  return check_launch("mf_tile_eval_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mf_backward_tiled_f32(const float* grad_out_real, const float* grad_out_imag,
                                           const float* tx_pos, const float* rx_pos,
                                           float* grad_real, float* grad_imag,
                                           const float* voxel, const float* tile_info,
                                           int64_t A, int64_t K, int64_t N, int64_t norm_count,
                                           int32_t nterms, int32_t tile_voxels, void* workspace,
                                           int64_t workspace_bytes,
                                           const geraf_fmcw_params* params,
                                           geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_params(params);
  if (A < 1 || K < 1 || N < 1) throw ArgError("need A >= 1, K >= 1, N >= 1");
  if (!tile_size_ok(tile_voxels)) throw ArgError("tile_voxels must be 128 or 512");
  if (norm_count < 1) throw ArgError("norm_count must be >= 1");
  if (A * K >= (INT64_C(1) << 31) || N >= (INT64_C(1) << 31) - kTileVox)
    throw ArgError("A*K and N must be < 2^31");
  const int NT = round_terms(nterms);
  if (nterms < 1 || NT < 0) throw ArgError("nterms must be in [1, 32]");
  require_ptrs({grad_out_real, grad_out_imag, tx_pos, rx_pos, grad_real, grad_imag, voxel,
                tile_info, workspace},
               true);
  const int tv = tile_voxels;
  const TiledLayout l = tiled_layout(K, N, NT, tv);
  const int ntiles = (int)((N + tv - 1) / tv);
  const TiledBuffers b = carve(workspace, workspace_bytes, l, A, ntiles, NT);
  const hipStream_t st = (hipStream_t)stream;
  const PhaseConsts pc = make_phase_consts(*params);
  const double kc = 0.5 * (double)(K - 1);
  const float4* ti = reinterpret_cast<const float4*>(tile_info);
  const double inv_norm = 1.0 / (double)norm_count;
  {
    const int ops = (int)((K + 15) / 16) * kWave;
    hipLaunchKernelGGL(xi_adj_ops_kernel, dim3((unsigned)((ops + 255) / 256)), dim3(256), 0, st,
                       b.xi_ops, (int)K, NT, kc);
  }
  for (int a0 = 0; a0 < A; a0 += b.na_max) {
    const int na = (int)std::min<int64_t>(b.na_max, A - a0);
#define GERAF_TILED_ADJ(T)                                                                   \
  case T:                                                                                    \
    launch_adj_chunk<T>(grad_out_real, grad_out_imag, tx_pos, rx_pos, voxel, ti, b.xi_ops, a0, \
                        na, (int)K, (int)N, ntiles, tv, pc, kc, inv_norm, b.coef, b.tau_c,     \
                        grad_real, grad_imag, st);                                             \
    break;
    switch (NT) {
      GERAF_TILED_ADJ(4)
      GERAF_TILED_ADJ(6)
      GERAF_TILED_ADJ(8)
      GERAF_TILED_ADJ(10)
      GERAF_TILED_ADJ(12)
      GERAF_TILED_ADJ(14)
      GERAF_TILED_ADJ(16)
      GERAF_TILED_ADJ(20)
      GERAF_TILED_ADJ(24)
      GERAF_TILED_ADJ(32)
      default: throw ArgError("unsupported term count");
    }
#undef GERAF_TILED_ADJ
  }
  return check_launch("mf_tile_adj_eval_mma_kernel");
  GERAF_TRY_END
}
