// Signal-tracer kernels for gfx950 (MI355X): the [A,P] drop-in op and the
// bilinear depth resampler.
//
// Forward (SignalTracer_cuda_kernel.cu:351-450):
//   out[a,k] = (1/P) sum_{p inbounds, not culled} sigma[a,p] * spec_ap / d_ap^2
//                                              * exp(+j*2*pi*(f0 + k*dfs)*tau_ap)
// Backward (SignalTracer_cuda_kernel.cu:452-587), with
//   h_ap = sum_k (g_re[a,k] cos phi + g_im[a,k] sin phi) = Re sum_k g[a,k] exp(-j phi):
//   grad_sigma[a,p] = (1/P) spec/d^2 h_ap
//   grad_normal[p]  = (1/P) sum_a sigma[a,p]/d^2 * dspec/dn * h_ap
//
// The geometry of a (antenna, point) pair is computed once, not once per
// sample as in the reference.  Forward: a block owns (antenna, 32-sample
// chunk); lanes stride over the points, each seeds its phasor from the FP64
// reduced phase and rotates it through the chunk into 32 private complex
// accumulators, which the block reduces in a fixed order.  Backward: a lane
// owns a point, the block's waves split the antennas, and h_ap is a
// two-level Horner sum over k with the gradient row g[a,:] read through the
// scalar cache (the same evaluation as the matched-filter forward).
#include <hip/hip_runtime.h>

#include <math.h>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

constexpr int kChunk = 32;
constexpr int kWaves = 4;     // forward: waves per (antenna, chunk) block
constexpr int kBwdWaves = 8;  // backward: waves per 64 points, splitting the antennas

__global__ __launch_bounds__(64 * kWaves) void st_forward_kernel(
    const float* __restrict__ normals, const float* __restrict__ sigmas,
    const float* __restrict__ points, const float* __restrict__ tx_pos,
    const float* __restrict__ rx_pos, const uint8_t* __restrict__ inbounds,
    float* __restrict__ out_real, float* __restrict__ out_imag, int A, int P, int K,
    double inv_p, PhaseConsts pc) {
  __shared__ float red[kWaves][kChunk][2];
  const int a = blockIdx.x;
  const int k0 = blockIdx.y * kChunk;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const V3 txa = load3(tx_pos + 3 * a);
  const V3 rxa = load3(rx_pos + 3 * a);

  float pr[kChunk], pi[kChunk];
#pragma unroll
  for (int j = 0; j < kChunk; ++j) pr[j] = pi[j] = 0.0f;

  for (int p = threadIdx.x; p < P; p += blockDim.x) {
    if (!inbounds[p]) continue;
    const V3 n = load3(normals + 3 * (size_t)p);
    const V3 pt = load3(points + 3 * (size_t)p);
    const SpecGeom g = spec_geom(txa, rxa, n, pt);
    if (g.culled()) continue;
    const double d = dist3_f64_fast((double)txa.x, (double)txa.y, (double)txa.z, (double)pt.x,
                                    (double)pt.y, (double)pt.z) +
                     dist3_f64_fast((double)rxa.x, (double)rxa.y, (double)rxa.z, (double)pt.x,
                                    (double)pt.y, (double)pt.z);
    const float amp = sigmas[(size_t)a * P + p] * g.spec / (float)(d * d);
    const double tau = (d + pc.bias) * pc.inv_c;
    const double cyc1 = pc.dfs * tau;
    const float2 seed =
        phasor_cycles((float)centered_frac(pc.f0 * tau + (double)k0 * cyc1), 1.0f);
    const float2 w = phasor_cycles((float)centered_frac(cyc1), 1.0f);
    float2 z = make_float2(amp * seed.x, amp * seed.y);
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      pr[j] += z.x;
      pi[j] += z.y;
      z = cmul(z, w);
    }
  }
#pragma unroll
  for (int j = 0; j < kChunk; ++j) {
    const float sr = wave_sum(pr[j]);
    const float si = wave_sum(pi[j]);
    if (lane == 0) {
      red[wave][j][0] = sr;
      red[wave][j][1] = si;
    }
  }
  __syncthreads();
  if (threadIdx.x < kChunk && k0 + (int)threadIdx.x < K) {
    const int j = threadIdx.x;
    double sr = 0.0, si = 0.0;
    for (int w = 0; w < kWaves; ++w) {
      sr += red[w][j][0];
      si += red[w][j][1];
    }
    out_real[(size_t)a * K + k0 + j] = (float)(sr * inv_p);
    out_imag[(size_t)a * K + k0 + j] = (float)(si * inv_p);
  }
}

// Backward: a block owns 64 points (one per lane) and its kBwdWaves waves
// split the antennas (wave w takes a = w, w + kBwdWaves, ...), so that the
// launch has kBwdWaves waves per 64 points instead of one; each wave writes
// its antennas' sigma gradients and keeps a partial normal gradient per
// point, summed over the waves in LDS in a fixed order (deterministic).
__global__ __launch_bounds__(64 * kBwdWaves) void st_backward_kernel(
    const float* __restrict__ g_real, const float* __restrict__ g_imag,
    const float* __restrict__ normals, const float* __restrict__ sigmas,
    const float* __restrict__ points, const float* __restrict__ tx_pos,
    const float* __restrict__ rx_pos, const uint8_t* __restrict__ inbounds,
    float* __restrict__ grad_normals, float* __restrict__ grad_sigmas, int A, int P, int K,
    float inv_p, PhaseConsts pc) {
  __shared__ float red[kBwdWaves][3][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = blockIdx.x * kWave + lane;
  const bool valid = p < P;
  const int pcl = valid ? p : P - 1;
  const bool inb = valid && inbounds[pcl];
  const V3 n = load3(normals + 3 * (size_t)pcl);
  const V3 pt = load3(points + 3 * (size_t)pcl);
  V3 gn = v3(0.0f, 0.0f, 0.0f);
  for (int a = wave; a < A; a += kBwdWaves) {
    const V3 txa = load3(tx_pos + 3 * a);
    const V3 rxa = load3(rx_pos + 3 * a);
    const SpecGeom g = spec_geom(txa, rxa, n, pt);
    float gs = 0.0f;
    if (inb && !g.culled()) {
      const double d = dist3_f64_fast((double)txa.x, (double)txa.y, (double)txa.z, (double)pt.x,
                                      (double)pt.y, (double)pt.z) +
                       dist3_f64_fast((double)rxa.x, (double)rxa.y, (double)rxa.z, (double)pt.x,
                                      (double)pt.y, (double)pt.z);
      const float att = rcp_fast((float)(d * d));
      const float sv = sigmas[(size_t)a * P + p];
      const PathPhasors ph = path_phasors((d + pc.bias) * pc.inv_c, pc, -1.0f);
      const float2 S = horner2(g_real + (size_t)a * K, g_imag + (size_t)a * K, K, ph.w, ph.W);
      const float h = cmul(ph.base, S).x;
      gn = gn + (sv * att * h) * dspec_dnormal(g, n);
      gs = att * g.spec * h * inv_p;
    }
    if (valid) grad_sigmas[(size_t)a * P + p] = gs;
  }
  red[wave][0][lane] = gn.x;
  red[wave][1][lane] = gn.y;
  red[wave][2][lane] = gn.z;
  __syncthreads();
  if (wave == 0 && valid) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float t = red[0][c][lane];
#pragma unroll
      for (int w = 1; w < kBwdWaves; ++w) t += red[w][c][lane];
      grad_normals[3 * (size_t)p + c] = t * inv_p;
    }
  }
}

// ---------------------------------------------------------------- plain FP64 forms

struct D3 {
  double x, y, z;
};
__device__ __forceinline__ D3 d3load(const double* p) { return D3{p[0], p[1], p[2]}; }
__device__ __forceinline__ D3 operator-(D3 a, D3 b) { return D3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ D3 operator*(double s, D3 a) { return D3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ D3 operator+(D3 a, D3 b) { return D3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ double dot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 nrm_eps(D3 v) { return (1.0 / (sqrt(dot(v, v)) + 1.0e-7)) * v; }

struct SpecGeomD {
  D3 inc, ret, refl;
  double refl_norm, idn, spec;
};
__device__ __forceinline__ SpecGeomD spec_geom_d(D3 t, D3 r, D3 n, D3 p) {
  SpecGeomD g;
  g.inc = nrm_eps(p - t);
  g.idn = dot(g.inc, n);
  const D3 rr = g.inc - (2.0 * g.idn) * n;
  g.refl_norm = sqrt(dot(rr, rr));
  g.refl = (1.0 / (g.refl_norm + 1.0e-7)) * rr;
  g.ret = nrm_eps(r - p);
  g.spec = dot(g.ret, g.refl);
  return g;
}
__device__ __forceinline__ double dist_d(D3 a, D3 b) { return sqrt(dot(a - b, a - b)); }

__global__ __launch_bounds__(256) void st_forward_f64_kernel(
    const double* __restrict__ normals, const double* __restrict__ sigmas,
    const double* __restrict__ points, const double* __restrict__ tx_pos,
    const double* __restrict__ rx_pos, const uint8_t* __restrict__ inbounds,
    double* __restrict__ out_real, double* __restrict__ out_imag, int A, int P, int K,
    double inv_p, PhaseConsts pc) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)A * K) return;
  const int a = (int)(idx / K), k = (int)(idx % K);
  const D3 t = d3load(tx_pos + 3 * a), r = d3load(rx_pos + 3 * a);
  double ar = 0.0, ai = 0.0;
  for (int p = 0; p < P; ++p) {
    if (!inbounds[p]) continue;
    const D3 n = d3load(normals + 3 * (size_t)p), pt = d3load(points + 3 * (size_t)p);
    const SpecGeomD g = spec_geom_d(t, r, n, pt);
    if (g.spec < 1.0e-6 || g.idn > 0.0) continue;
    const double d = dist_d(t, pt) + dist_d(r, pt);
    const double amp = sigmas[(size_t)a * P + p] * g.spec / (d * d);
    const double tau = (d + pc.bias) * pc.inv_c;
    double s, c;
    sincospi(2.0 * centered_frac(centered_frac(pc.f0 * tau) + (double)k * (pc.dfs * tau)), &s,
             &c);
    ar += amp * c;
    ai += amp * s;
  }
  out_real[idx] = ar * inv_p;
  out_imag[idx] = ai * inv_p;
}

__global__ __launch_bounds__(256) void st_backward_f64_kernel(
    const double* __restrict__ g_real, const double* __restrict__ g_imag,
    const double* __restrict__ normals, const double* __restrict__ sigmas,
    const double* __restrict__ points, const double* __restrict__ tx_pos,
    const double* __restrict__ rx_pos, const uint8_t* __restrict__ inbounds,
    double* __restrict__ grad_normals, double* __restrict__ grad_sigmas, int A, int P, int K,
    double inv_p, PhaseConsts pc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  const D3 n = d3load(normals + 3 * (size_t)p), pt = d3load(points + 3 * (size_t)p);
  D3 gn{0.0, 0.0, 0.0};
  const bool inb = inbounds[p];
  for (int a = 0; a < A; ++a) {
    double gs = 0.0;
    const D3 t = d3load(tx_pos + 3 * a), r = d3load(rx_pos + 3 * a);
    const SpecGeomD g = spec_geom_d(t, r, n, pt);
    if (inb && !(g.spec < 1.0e-6 || g.idn > 0.0)) {
      const D3 q = g.ret - g.spec * g.refl;
      const D3 dsn = (-2.0 / (g.refl_norm + 1.0e-7)) * (dot(q, n) * g.inc + g.idn * q);
      const double d = dist_d(t, pt) + dist_d(r, pt);
      const double att = 1.0 / (d * d);
      const double tau = (d + pc.bias) * pc.inv_c;
      const double fr0 = centered_frac(pc.f0 * tau), cyc1 = pc.dfs * tau;
      double h = 0.0;
      for (int k = 0; k < K; ++k) {
        double s, c;
        sincospi(2.0 * centered_frac(fr0 + (double)k * cyc1), &s, &c);
        h += g_real[(size_t)a * K + k] * c + g_imag[(size_t)a * K + k] * s;
      }
      gn = gn + (sigmas[(size_t)a * P + p] * att * h) * dsn;
      gs = att * g.spec * h * inv_p;
    }
    grad_sigmas[(size_t)a * P + p] = gs;
  }
  grad_normals[3 * (size_t)p + 0] = gn.x * inv_p;
  grad_normals[3 * (size_t)p + 1] = gn.y * inv_p;
  grad_normals[3 * (size_t)p + 2] = gn.z * inv_p;
}

// ---------------------------------------------------------------- bilinear resampler

// Binary search of SignalTracer_cuda_kernel.cu:234-243: the largest lo with
// z[lo] <= t among [0, S-2] (lo stays 0 when t < z[1]).
template <typename T>
__device__ __forceinline__ int bl_lower(const T* z, int S, T t) {
  int lo = 0, hi = S - 1;
  while (lo < hi - 1) {
    const int mid = (lo + hi) / 2;
    if (z[mid] <= t)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

// Forward, one thread per (antenna, ray, target).  With STAGE a block covers
// kBlTargets targets of one (antenna, ray) and first copies the ray's source
// depths to LDS (one coalesced load), so each target's binary search reads
// LDS instead of making its dependent steps through the caches; rays longer
// than kBlStageMax samples search global memory.
constexpr int kBlTargets = 256;
constexpr int kBlStageMax = 4096;

template <typename T, bool STAGE>
__global__ __launch_bounds__(kBlTargets) void bilinear_forward_kernel(
    const T* __restrict__ normals, const T* __restrict__ sigmas, const T* __restrict__ z,
    const T* __restrict__ tz, T* __restrict__ out_n, T* __restrict__ out_s,
    uint8_t* __restrict__ out_inb, int A, int R, int S, int Tn) {
  extern __shared__ unsigned char bl_smem[];
  int t, r, a;
  if constexpr (STAGE) {
    // block -> (target chunk, ray, antenna)
    const int chunks = (Tn + kBlTargets - 1) / kBlTargets;
    const long long b = blockIdx.x;
    const int tc = (int)(b % chunks);
    const long long ra = b / chunks;
    r = (int)(ra % R);
    a = (int)(ra / R);
    t = tc * kBlTargets + threadIdx.x;
  } else {
    const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (long long)A * R * Tn) return;
    t = (int)(idx % Tn);
    r = (int)((idx / Tn) % R);
    a = (int)(idx / ((long long)Tn * R));
  }
  const T* zr = z + (size_t)r * S;
  if constexpr (STAGE) {
    T* zs = reinterpret_cast<T*>(bl_smem);
    for (int i = threadIdx.x; i < S; i += blockDim.x) zs[i] = zr[i];
    __syncthreads();
    zr = zs;
    if (t >= Tn) return;  // after the block's only barrier
  }
  const T tv = tz[(size_t)r * Tn + t];
  const int lo = bl_lower(zr, S, tv);
  const T z0 = zr[lo], z1 = zr[lo + 1];
  const bool oob = (tv < z0) || (tv > z1);
  const size_t rt = (size_t)r * Tn + t;
  if (a == 0) out_inb[rt] = !oob;
  T* os = out_s + ((size_t)a * R + r) * Tn + t;
  if (oob) {
    *os = T(0);
    if (a == 0) out_n[3 * rt] = out_n[3 * rt + 1] = out_n[3 * rt + 2] = T(0);
    return;
  }
  const T den = max(z1 - z0, static_cast<T>(1.0e-6f));
  const T wu = (tv - z0) / den;
  const T wl = T(1) - wu;
  const T* sr = sigmas + ((size_t)a * R + r) * S;
  *os = wl * sr[lo] + wu * sr[lo + 1];
  if (a == 0) {
    const T* nr = normals + (size_t)r * S * 3;
    for (int c = 0; c < 3; ++c) out_n[3 * rt + c] = wl * nr[3 * lo + c] + wu * nr[3 * (lo + 1) + c];
  }
}

// Byte offset of the backward's staged source depths: after the targets'
// weights and bracket keys, rounded up to 16 bytes.
template <typename T>
__host__ __device__ constexpr size_t bl_zstage_offset(int Tn) {
  return ((size_t)Tn * (2 * sizeof(T) + sizeof(int)) + 15) & ~(size_t)15;
}

// Byte offset of the backward's table of bracket bounds (S + 2 ints): after
// the staged depths when they are staged, else after the bracket keys.
template <typename T>
__host__ __device__ constexpr size_t bl_lb_offset(int Tn, int S, int stage_z) {
  return stage_z ? ((bl_zstage_offset<T>(Tn) + (size_t)S * sizeof(T) + 15) & ~(size_t)15)
                 : bl_zstage_offset<T>(Tn);
}

// First index t in [0, n) with key[t] >= v, for a non-decreasing key (LDS).
__device__ __forceinline__ int lds_lower_bound(const int* key, int n, int v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (key[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// Atomic-free transpose: block (ray r, antenna a).  The targets' brackets and
// weights go to LDS; each thread then gathers, for its source samples, every
// target that lands on them, in ascending target order (deterministic).
// Source s receives the upper weight of the targets bracketed by (s-1, s)
// and the lower weight of those bracketed by (s, s+1).  When the targets'
// bracket keys do not decrease along the ray (sorted target depths, below-range
// targets first and above-range ones last), each of the two sets is a
// contiguous range of targets found by binary search: O(S log T + T) per ray.
// Otherwise every source scans all targets (O(S T)).  Both visit the
// contributing targets in ascending order, so they give the same sums.
template <typename T>
__global__ __launch_bounds__(64) void bilinear_backward_kernel(
    const T* __restrict__ g_n, const T* __restrict__ g_s, const T* __restrict__ z,
    const T* __restrict__ tz, T* __restrict__ grad_n, T* __restrict__ grad_s, int A, int R,
    int S, int Tn, int stage_z, int use_lb) {
  extern __shared__ unsigned char smem[];
  T* wl = reinterpret_cast<T*>(smem);
  T* wu = wl + Tn;
  // key[t]: the bracket (t between source samples key and key + 1, in
  // [0, S - 2]), -1 below the first source sample, S - 1 above the last
  int* key = reinterpret_cast<int*>(wu + Tn);
  const int r = blockIdx.x;
  const int a = blockIdx.y;
  const T* zr = z + (size_t)r * S;
  if (stage_z) {
    // the ray's source depths in LDS for the targets' binary searches (after
    // the target arrays, 16-byte aligned)
    T* zs = reinterpret_cast<T*>(smem + bl_zstage_offset<T>(Tn));
    for (int i = threadIdx.x; i < S; i += blockDim.x) zs[i] = zr[i];
    __syncthreads();
    zr = zs;
  }
  for (int t = threadIdx.x; t < Tn; t += blockDim.x) {
    const T tv = tz[(size_t)r * Tn + t];
    const int l = bl_lower(zr, S, tv);
    const T z0 = zr[l], z1 = zr[l + 1];
    if (tv < z0 || tv > z1) {
      // contributes to no source sample; sorts before (below) or after
      // (above) every bracket
      wl[t] = wu[t] = T(0);
      key[t] = tv < z0 ? -1 : S - 1;
    } else {
      const T den = max(z1 - z0, static_cast<T>(1.0e-6f));
      const T u = (tv - z0) / den;
      key[t] = l;
      wu[t] = u;
      wl[t] = T(1) - u;
    }
  }
  __syncthreads();
  int unsorted = 0;
  for (int t = threadIdx.x; t + 1 < Tn; t += blockDim.x) unsorted |= key[t] > key[t + 1];
  const bool ranges = __syncthreads_or(unsorted) == 0;
  // with sorted keys: lb[k + 1] = the first target whose key is >= k, for
  // k = -1 .. S (one search per boundary, shared by the two sources beside it)
  int* lb = reinterpret_cast<int*>(smem + bl_lb_offset<T>(Tn, S, stage_z));
  if (ranges && use_lb) {
    for (int k = (int)threadIdx.x - 1; k <= S; k += blockDim.x)
      lb[k + 1] = lds_lower_bound(key, Tn, k);
    __syncthreads();
  }
  const T* gs = g_s + ((size_t)a * R + r) * Tn;
  for (int s = threadIdx.x; s < S; s += blockDim.x) {
    T acc = T(0), an0 = T(0), an1 = T(0), an2 = T(0);
    auto add = [&](int t, T w) {
      acc += w * gs[t];
      if (a == 0) {
        const T* gn = g_n + ((size_t)r * Tn + t) * 3;
        an0 += w * gn[0];
        an1 += w * gn[1];
        an2 += w * gn[2];
      }
    };
    if (ranges) {
      // keys s - 1 (upper weight), then keys s (lower weight): ascending t
      const int t0 = use_lb ? lb[s] : lds_lower_bound(key, Tn, s - 1);
      const int t1 = use_lb ? lb[s + 1] : lds_lower_bound(key, Tn, s);
      const int t2 = use_lb ? lb[s + 2] : lds_lower_bound(key, Tn, s + 1);
      if (s >= 1)
        for (int t = t0; t < t1; ++t) add(t, wu[t]);
      if (s <= S - 2)
        for (int t = t1; t < t2; ++t) add(t, wl[t]);
    } else {
      for (int t = 0; t < Tn; ++t) {
        const int l = key[t];
        if (l < 0 || l > S - 2) continue;
        if (l == s)
          add(t, wl[t]);
        else if (l + 1 == s)
          add(t, wu[t]);
      }
    }
    grad_s[((size_t)a * R + r) * S + s] = acc;
    if (a == 0) {
      T* o = grad_n + ((size_t)r * S + s) * 3;
      o[0] = an0;
      o[1] = an1;
      o[2] = an2;
    }
  }
}

void check_st_sizes(int64_t A, int64_t P, int64_t K) {
  if (A < 0 || P < 0 || K < 0) throw ArgError("sizes must be non-negative");
  if (A > INT32_MAX || P > INT32_MAX || K > INT32_MAX || A * P > (INT64_C(1) << 40) ||
      A * K > (INT64_C(1) << 40))
    throw ArgError("sizes too large");
}

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int geraf_st_forward_f32(const float* normals, const float* sigmas, const float* points,
                                    const float* tx_pos, const float* rx_pos,
                                    const uint8_t* inbounds, float* out_real, float* out_imag,
                                    int64_t A, int64_t P, int64_t K,
                                    const geraf_fmcw_params* params, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_st_sizes(A, P, K);
  check_params(params);
  if (A == 0 || K == 0) return GERAF_RF_OK;
  require_ptrs({tx_pos, rx_pos, out_real, out_imag}, true);
  require_ptrs({normals, sigmas, points, inbounds}, P > 0);
  const dim3 grid((unsigned)A, (unsigned)((K + kChunk - 1) / kChunk));
  hipLaunchKernelGGL(st_forward_kernel, grid, dim3(64 * kWaves), 0, (hipStream_t)stream, normals,
                     sigmas, points, tx_pos, rx_pos, inbounds, out_real, out_imag, (int)A,
                     (int)P, (int)K, 1.0 / (double)P, make_phase_consts(*params));
  return check_launch("st_forward_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_st_backward_f32(const float* grad_out_real, const float* grad_out_imag,
                                     const float* normals, const float* sigmas,
                                     const float* points, const float* tx_pos,
                                     const float* rx_pos, const uint8_t* inbounds,
                                     float* grad_normals, float* grad_sigmas, int64_t A,
                                     int64_t P, int64_t K, const geraf_fmcw_params* params,
                                     geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_st_sizes(A, P, K);
  check_params(params);
  if (P == 0) return GERAF_RF_OK;
  require_ptrs({normals, sigmas, points, inbounds, grad_normals}, true);
  require_ptrs({tx_pos, rx_pos, grad_sigmas}, A > 0);
  require_ptrs({grad_out_real, grad_out_imag}, A > 0 && K > 0);
  hipLaunchKernelGGL(st_backward_kernel, dim3((unsigned)((P + kWave - 1) / kWave)),
                     dim3(64 * kBwdWaves), 0,
                     (hipStream_t)stream, grad_out_real, grad_out_imag, normals, sigmas, points,
                     tx_pos, rx_pos, inbounds, grad_normals, grad_sigmas, (int)A, (int)P, (int)K,
                     (float)(1.0 / (double)P), make_phase_consts(*params));
  return check_launch("st_backward_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_st_forward_f64(const double* normals, const double* sigmas,
                                    const double* points, const double* tx_pos,
                                    const double* rx_pos, const uint8_t* inbounds,
                                    double* out_real, double* out_imag, int64_t A, int64_t P,
                                    int64_t K, const geraf_fmcw_params* params,
                                    geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_st_sizes(A, P, K);
  check_params(params);
  if (A == 0 || K == 0) return GERAF_RF_OK;
  require_ptrs({tx_pos, rx_pos, out_real, out_imag}, true);
  require_ptrs({normals, sigmas, points, inbounds}, P > 0);
  const long long total = A * K;
  hipLaunchKernelGGL(st_forward_f64_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, normals, sigmas, points, tx_pos, rx_pos, inbounds,
                     out_real, out_imag, (int)A, (int)P, (int)K, 1.0 / (double)P,
                     make_phase_consts(*params));
  return check_launch("st_forward_f64_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_st_backward_f64(const double* grad_out_real, const double* grad_out_imag,
                                     const double* normals, const double* sigmas,
                                     const double* points, const double* tx_pos,
                                     const double* rx_pos, const uint8_t* inbounds,
                                     double* grad_normals, double* grad_sigmas, int64_t A,
                                     int64_t P, int64_t K, const geraf_fmcw_params* params,
                                     geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_st_sizes(A, P, K);
  check_params(params);
  if (P == 0) return GERAF_RF_OK;
  require_ptrs({normals, sigmas, points, inbounds, grad_normals}, true);
  require_ptrs({tx_pos, rx_pos, grad_sigmas}, A > 0);
  require_ptrs({grad_out_real, grad_out_imag}, A > 0 && K > 0);
  hipLaunchKernelGGL(st_backward_f64_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, grad_out_real, grad_out_imag, normals, sigmas, points,
                     tx_pos, rx_pos, inbounds, grad_normals, grad_sigmas, (int)A, (int)P, (int)K,
                     1.0 / (double)P, make_phase_consts(*params));
  return check_launch("st_backward_f64_kernel");
  GERAF_TRY_END
}

namespace {
void check_bl_sizes(int64_t A, int64_t R, int64_t S, int64_t T) {
  if (A < 0 || R < 0 || T < 0) throw ArgError("sizes must be non-negative");
  if (S < 2) throw ArgError("need at least 2 source samples per ray (Nz >= 2)");
  if (A * R * (S > T ? S : T) > (INT64_C(1) << 40) || R > INT32_MAX || A > 65535)
    throw ArgError("sizes too large");
  if (T > 8192) throw ArgError("Nz_tgt must be <= 8192");
}

template <typename T>
int bilinear_forward(const T* normals, const T* sigmas, const T* z, const T* tz, T* out_n,
                     T* out_s, uint8_t* out_inb, int64_t A, int64_t R, int64_t S, int64_t Tn,
                     geraf_stream_t stream) {
  check_bl_sizes(A, R, S, Tn);
  const long long total = A * R * Tn;
  if (total == 0) return GERAF_RF_OK;
  require_ptrs({normals, sigmas, z, tz, out_n, out_s, out_inb}, true);
  if (S <= kBlStageMax) {
    const long long blocks = A * R * ((Tn + kBlTargets - 1) / kBlTargets);
    if (blocks > INT32_MAX) throw ArgError("sizes too large");
    hipLaunchKernelGGL((bilinear_forward_kernel<T, true>), dim3((unsigned)blocks),
                       dim3(kBlTargets), (size_t)S * sizeof(T), (hipStream_t)stream, normals,
                       sigmas, z, tz, out_n, out_s, out_inb, (int)A, (int)R, (int)S, (int)Tn);
  } else {
    hipLaunchKernelGGL((bilinear_forward_kernel<T, false>),
                       dim3((unsigned)((total + kBlTargets - 1) / kBlTargets)), dim3(kBlTargets),
                       0, (hipStream_t)stream, normals, sigmas, z, tz, out_n, out_s, out_inb,
                       (int)A, (int)R, (int)S, (int)Tn);
  }
  return check_launch("bilinear_forward_kernel");
}

template <typename T>
int bilinear_backward(const T* g_n, const T* g_s, const T* z, const T* tz, T* grad_n, T* grad_s,
                      int64_t A, int64_t R, int64_t S, int64_t Tn, geraf_stream_t stream) {
  check_bl_sizes(A, R, S, Tn);
  if (A == 0 || R == 0) return GERAF_RF_OK;
  require_ptrs({z, grad_n, grad_s}, true);
  require_ptrs({g_n, g_s, tz}, Tn > 0);
  // the ray's source depths go to LDS too when the whole stays within 64 KB;
  // then the table of bracket bounds (S + 2 ints)
  // (each only when the whole stays within 64 KB)
  const size_t lb_bytes = ((size_t)S + 2) * sizeof(int);
  const int stage_z =
      S <= kBlStageMax &&
              bl_lb_offset<T>((int)Tn, (int)S, 1) + lb_bytes <= 65536 ? 1 : 0;
  // the shared table pays for rays of >= 1,024 source samples (measured: 233 ->
  // 151 us at 300 x 5,000 -> 700, and 352 -> 375 us at 32,768 x 256 -> 512,
  // where three searches per source cost less than the extra barrier)
  const int use_lb =
      S >= 1024 && bl_lb_offset<T>((int)Tn, (int)S, stage_z) + lb_bytes <= 65536 ? 1 : 0;
  const size_t lds = bl_lb_offset<T>((int)Tn, (int)S, stage_z) + (use_lb ? lb_bytes : 0);
  hipLaunchKernelGGL(bilinear_backward_kernel<T>, dim3((unsigned)R, (unsigned)A), dim3(64), lds,
                     (hipStream_t)stream, g_n, g_s, z, tz, grad_n, grad_s, (int)A, (int)R,
                     (int)S, (int)Tn, stage_z, use_lb);
  return check_launch("bilinear_backward_kernel");
}
}  // namespace

extern "C" int geraf_bilinear_forward_f32(const float* normals, const float* sigmas,
                                          const float* z_vals, const float* tgt_z_vals,
                                          float* out_normals, float* out_sigmas,
                                          uint8_t* out_inbounds, int64_t A, int64_t R,
                                          int64_t S, int64_t T, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return bilinear_forward(normals, sigmas, z_vals, tgt_z_vals, out_normals, out_sigmas,
                          out_inbounds, A, R, S, T, stream);
  GERAF_TRY_END
}

extern "C" int geraf_bilinear_forward_f64(const double* normals, const double* sigmas,
                                          const double* z_vals, const double* tgt_z_vals,
                                          double* out_normals, double* out_sigmas,
                                          uint8_t* out_inbounds, int64_t A, int64_t R,
                                          int64_t S, int64_t T, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return bilinear_forward(normals, sigmas, z_vals, tgt_z_vals, out_normals, out_sigmas,
                          out_inbounds, A, R, S, T, stream);
  GERAF_TRY_END
}

extern "C" int geraf_bilinear_backward_f32(const float* grad_out_normals,
                                           const float* grad_out_sigmas, const float* z_vals,
                                           const float* tgt_z_vals, float* grad_normals,
                                           float* grad_sigmas, int64_t A, int64_t R, int64_t S,
                                           int64_t T, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return bilinear_backward(grad_out_normals, grad_out_sigmas, z_vals, tgt_z_vals, grad_normals,
                           grad_sigmas, A, R, S, T, stream);
  GERAF_TRY_END
}

extern "C" int geraf_bilinear_backward_f64(const double* grad_out_normals,
                                           const double* grad_out_sigmas, const double* z_vals,
                                           const double* tgt_z_vals, double* grad_normals,
                                           double* grad_sigmas, int64_t A, int64_t R, int64_t S,
                                           int64_t T, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  return bilinear_backward(grad_out_normals, grad_out_sigmas, z_vals, tgt_z_vals, grad_normals,
                           grad_sigmas, A, R, S, T, stream);
  GERAF_TRY_END
}
