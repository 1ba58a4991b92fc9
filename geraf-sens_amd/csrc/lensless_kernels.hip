// Fused lensless RF renderer for gfx950: per-ray NeuS-style compositing and
// FMCW signal synthesis in one kernel, with a hand-written adjoint.
//
// Per ray r (monostatic antenna at its origin o, unit direction d) and sample s:
//   c      = d . g_s                                 (g = raw SDF gradient)
//   ic     = -(relu(0.5 - 0.5 c)(1 - r_a) + relu(-c) r_a)       rf_rendering.py:145-155
//   prev   = sdf - ic*dist/2,  next = sdf + ic*dist/2
//   alpha  = clip((sig(prev*inv_s) - sig(next*inv_s) + 1e-5)
//                 / (sig(prev*inv_s) + 1e-5), 0, 1)             rf_rendering.py:157-169
//   T      = prod_{i<s} (1 - alpha_i + 1e-7)                    rf_rendering.py:46-52
//   Tc     = T + offset_s        (calibration new_cdf - prev_cdf, rf_rendering.py:750)
//   sigma  = max(refl*power*Tc, 1e-6) * alpha * Tc              sdf_network.py:475-481
//   n      = g / max(|g|, 1e-12)                                 F.normalize
//   amp    = sigma * spec(o, o, n, o + z d) / d_path^2,  d_path = 2|z d|
//   out[r,k] = (1/n_valid) sum_s amp_s exp(+j 2 pi (f0 + k dfs) tau_s)
//                                         (SignalTracer_cuda_kernel.cu:351-450, one antenna)
//
// Layout: one wave per ray, four rays per block; samples are streamed in
// 64-sample tiles (one sample per lane; per-ray base addresses live in SGPRs,
// lanes add 32-bit offsets).  The transmittance is an exclusive prefix product
// over the tile (DPP row shifts and broadcasts) times the carry of the earlier
// tiles.  Divisions, square roots and the sigmoid's exponential use the
// hardware approximations (about 1 ulp); the round-trip distance and the
// carrier phase are formed in FP64.  The two cdfs, their difference and their
// slopes come from exp(-|x|) (logistic_fast): the FP32 difference of two cdfs
// near 1 and the slope p (1 - p), as the reference forms them, lose all
// precision in front of the surface, and over 2048 samples that costs the sdf
// gradient 1e-3 relative.  The antenna is monostatic, so the return direction
// is exactly the negated incident direction.
//
// Forward signal: emitting samples (valid, not culled, amp != 0) are appended
// to a per-wave LDS queue as (amplitude, reduced carrier phase, per-frequency
// phase step in FP64 cycles, step phasor).  The queue is drained in batches
// of 64 (and once at the end).  The drain is a complex matrix product on the
// matrix cores (k = 16 a + b, see synth_mfma), for every K: at K = 32, 64 and
// 128 it measured 22, 18 and 25 % faster than a rotation synthesis on the
// vector ALUs (4,096 rays x 128 samples), although most of its 256-frequency
// tile is idle there; for K <= 64 the 16-block 4 x 4 form (synth_mfma4) leaves
// nothing idle and took the 32,768 x 256, K = 64 forward from 86.6 to 75.4 us.  (A form on FP16 hi/lo operands, 32 samples per
// v_mfma_f32_16x16x32_f16 product, measured 4 % slower at 32,768 x 256,
// K = 512: its splits cost more vector instructions than the matrix time it
// saves, and the kernel is bound by vector issue.)
// Backward: pass A walks the tiles forward, keeps the transmittance carry at
// each tile start, and queues (tau, sample) of every emitting sample; full
// 64-sample batches evaluate h_s = Re sum_k g[r,k] exp(-j phi_sk), for
// K > 128 as a matrix product with FP16 hi/lo operands (see h_mma16; the
// FP32-operand form h_mfma stays for GERAF_LL_H16=0), else by two-level
// Horner (one sample per lane, gradient row read through the scalar cache).  Pass B walks the
// tiles backward, recomputes the compositing and geometry, and propagates
// through the transmittance with a reverse affine scan
// W_s = gT_s + t_s W_{s+1}, so dL/dalpha_i = (direct) - T_i W_{i+1} needs no
// division by (1 - alpha + 1e-7).
#include <hip/hip_runtime.h>

#include <math.h>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

constexpr int kRaysPerBlock = 4;
constexpr int kTile = 64;
constexpr int kMaxTiles = 32;  // S <= 2048
constexpr int kQueue = 2 * kTile;

struct LenslessArgs {
  const float* sdf;           // [R,S]
  const float* grad;          // [R,S,3]
  const float* refl;          // [R,S] or null (= 1)
  const float* dists;         // [R,S]
  const float* z;             // [R,S]
  const float* origin;        // [R,3]
  const float* dirs;          // [R,3]
  const float* offset;        // [R,S] or null (= 0)
  const int32_t* n_valid;     // [R] or null (= S)
  const float* inv_s;         // [1]
  const float* power;         // [1]
  int R, S, K;
  float anneal;
};

// Element `i` of a row whose base is wave-uniform (SGPR base + 32-bit offset).
__device__ __forceinline__ float ld_row(const float* row, unsigned i) {
  return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(row) + i * 4u);
}
__device__ __forceinline__ V3 ld_row3(const float* row, unsigned i) {
  const float* p = reinterpret_cast<const float*>(reinterpret_cast<const char*>(row) + i * 12u);
  return v3(p[0], p[1], p[2]);
}
__device__ __forceinline__ void st_row(float* row, unsigned i, float v) {
  *reinterpret_cast<float*>(reinterpret_cast<char*>(row) + i * 4u) = v;
}

// One ray's wave-uniform state.
struct Ray {
  size_t row;        // r * S
  int nv;            // valid samples
  V3 o, dir;
  double ox, oy, oz;
  float inv_s, power;
};

__device__ __forceinline__ Ray load_ray(const LenslessArgs& a, int r) {
  Ray y;
  y.row = (size_t)r * a.S;
  y.nv = a.n_valid ? min(max(a.n_valid[r], 0), a.S) : a.S;
  y.o = load3(a.origin + 3 * (size_t)r);
  y.dir = load3(a.dirs + 3 * (size_t)r);
  y.ox = y.o.x;
  y.oy = y.o.y;
  y.oz = y.o.z;
  y.inv_s = a.inv_s[0];
  y.power = a.power[0];
  return y;
}

// Per-sample inputs and compositing terms.
struct Sample {
  float sdf, dist, refl, off, zz;
  V3 g;
  float c, prev, next, pc, nc, sp, sn, alpha, t;  // sp, sn: slopes of pc, nc
  bool qmask;
};

// The raw per-sample inputs.
struct RawSample {
  float sdf, dist, zz, refl, off;
  V3 g;
};

__device__ __forceinline__ RawSample load_raw(const LenslessArgs& a, const Ray& y, int s) {
  RawSample w;
  const unsigned u = (unsigned)s;
  w.sdf = ld_row(a.sdf + y.row, u);
  w.dist = ld_row(a.dists + y.row, u);
  w.zz = ld_row(a.z + y.row, u);
  w.refl = a.refl ? ld_row(a.refl + y.row, u) : 1.0f;
  w.off = a.offset ? ld_row(a.offset + y.row, u) : 0.0f;
  w.g = ld_row3(a.grad + 3 * y.row, u);
  return w;
}

__device__ __forceinline__ Sample composite(const LenslessArgs& a, const Ray& y,
                                            const RawSample& w) {
  Sample x;
  x.sdf = w.sdf;
  x.dist = w.dist;
  x.zz = w.zz;
  x.refl = w.refl;
  x.off = w.off;
  x.g = w.g;
  x.c = dot(y.dir, x.g);
  const float ic =
      -(fmaxf(-x.c * 0.5f + 0.5f, 0.0f) * (1.0f - a.anneal) + fmaxf(-x.c, 0.0f) * a.anneal);
  const float half = ic * x.dist * 0.5f;
  x.next = x.sdf + half;
  x.prev = x.sdf - half;
  const Logistic lp = logistic_fast(x.prev * y.inv_s);
  const Logistic ln = logistic_fast(x.next * y.inv_s);
  x.pc = lp.p;
  x.nc = ln.p;
  x.sp = lp.d;
  x.sn = ln.d;
  // pc - nc from the accurate pair: the complements once pc >= 1/2 (far in
  // front of the surface both cdfs round towards 1 and their FP32 difference
  // would be noise of order 1e-7).
  const float dcdf = lp.p >= 0.5f ? ln.c - lp.c : lp.p - ln.p;
  const float q = (dcdf + 1.0e-5f) * rcp_fast(x.pc + 1.0e-5f);
  x.qmask = (q >= 0.0f) && (q <= 1.0f);
  x.alpha = fminf(fmaxf(q, 0.0f), 1.0f);
  x.t = 1.0f - x.alpha + 1.0e-7f;
  return x;
}

__device__ __forceinline__ Sample load_composite(const LenslessArgs& a, const Ray& y, int s) {
  return composite(a, y, load_raw(a, y, s));
}

// Specular geometry of the monostatic sample (ComputeSpecularGeometry,
// SignalTracer_cuda_kernel.cu:172-197, with tx = rx = o, so ret = -inc).
// The reference forms v = p - o, inc = v / (|v| + 1e-7), n = g / |g|,
// rr = inc - 2 (inc.n) n, refl = rr / (|rr| + 1e-7) and spec = ret.refl.
// With nu = |inc| = |v| / (|v| + 1e-7) and |n| = 1, |rr|^2 = nu^2 exactly, so
//   spec = (2 (inc.n)^2 - nu^2) / (nu + 1e-7),
// and inc.n = (v.g) / ((|v| + 1e-7) |g|) with v.g = z (d.g) (v = z d up to
// the rounding of p, ~1e-7 relative).  |v| and the round-trip distance come
// from the FP32 point p = o + z d (separately rounded product and sum, as
// the tensor ops round it) in FP64.
struct Geo {
  V3 n;          // normalised normal
  float gl;      // |g|
  float incs;    // inc = incs * d
  float idn;     // inc . n
  float nu;      // |inc|
  float spec;    // ret . refl
  double dpath;  // round-trip distance (no bias)
  bool live;     // valid and not culled
};

__device__ __forceinline__ Geo geometry(const Ray& y, float zz, const V3& g, bool valid) {
  Geo e;
  e.gl = sqrt_fast(dot(g, g));
  const float inv_gl = rcp_fast(fmaxf(e.gl, 1.0e-12f));
  e.n = inv_gl * g;
  const V3 zd = v3(__fmul_rn(zz, y.dir.x), __fmul_rn(zz, y.dir.y), __fmul_rn(zz, y.dir.z));
  const V3 p = v3(__fadd_rn(y.o.x, zd.x), __fadd_rn(y.o.y, zd.y), __fadd_rn(y.o.z, zd.z));
  const double dx = (double)p.x - y.ox, dy = (double)p.y - y.oy, dz = (double)p.z - y.oz;
  const double vl = sqrt_f64_newton(fma(dx, dx, fma(dy, dy, dz * dz)));
  e.dpath = 2.0 * vl;
  const float vlf = (float)vl;
  const float inv_v = rcp_fast(vlf + 1.0e-7f);
  e.incs = zz * inv_v;
  e.nu = vlf * inv_v;
  e.idn = (e.incs * dot(y.dir, g)) * inv_gl;
  e.spec = fmaf(2.0f * e.idn, e.idn, -e.nu * e.nu) * rcp_fast(e.nu + 1.0e-7f);
  e.live = valid && !(e.spec < 1.0e-6f || e.idn > 0.0f);
  return e;
}

// d spec / d n up to its component along n (which the normalisation's
// Jacobian removes): 4 (inc.n) inc / (nu + 1e-7).
__device__ __forceinline__ V3 dspec_dn_mono(const Ray& y, const Geo& e) {
  return (4.0f * e.idn * e.incs * rcp_fast(e.nu + 1.0e-7f)) * y.dir;
}

// Number of set bits of `mask` below this lane.
__device__ __forceinline__ int lane_prefix(unsigned long long mask) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// ------------------------------------------------------------------ forward

// Per-wave queue of emitting samples: amplitude, carrier phase f0 tau reduced
// to cycles, per-frequency step dfs tau in FP64 cycles, and the 256-frequency
// rotation exp(j 2 pi 256 dfs tau) (used for K > 256).
struct FwdQueue {
  float amp[kRaysPerBlock][kQueue];
  float fr0[kRaysPerBlock][kQueue];
  float2 w[kRaysPerBlock][kQueue];
  double cyc1[kRaysPerBlock][kQueue];
};

// Matrix-core synthesis.  With k = 16 a + b (b < 16) the ray's signal is a
// complex product over its emitting samples s,
//   out[16 a + b] = sum_s X[a,s] Y[s,b],
//   X[a,s] = amp_s exp(+j 2 pi (f0 + 16 a dfs) tau_s),  Y[s,b] = exp(+j 2 pi b dfs tau_s),
// taken four samples at a time by v_mfma_f32_16x16x4f32 (exact FP32
// products, FP32 accumulation), three real products per 16-row tile of a:
// P1 = Xr Yr, P2 = Xi Yi, P3 = (Xr + Xi)(Yr + Yi); Re = P1 - P2 and
// Im = P3 - P1 - P2 (the error stays O(eps sum |X||Y|), as with four).
// Each lane forms one Y and one X phasor per step from FP64-reduced phases
// (hardware sine and cosine); the rows of further tiles follow by one complex
// rotation each (the sample's exp(j 2 pi 256 dfs tau), queued with it).
// Operand lanes of that instruction: A[i][k] and B[k][j] come from lane
// i + 16 k (resp. j + 16 k); D[4 q + v][j] is element v of lane j + 16 q.
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int MT>
__device__ __forceinline__ void synth_mfma(const FwdQueue& q, int wave, int lane, int n,
                                           f32x4* dr, f32x4* di, f32x4* ds) {
  const int j = lane & 15;
  for (int s0 = 0; s0 < n; s0 += 4) {
    const int i = s0 + (lane >> 4);
    float am = 0.0f, f0 = 0.0f;
    double c1 = 0.0;
    if (i < n) {
      am = q.amp[wave][i];
      f0 = q.fr0[wave][i];
      c1 = q.cyc1[wave][i];
    }
    const float2 yv = phasor_cycles_hw((float)centered_frac((double)j * c1));
    const float2 xv =
        phasor_cycles_hw((float)centered_frac(fma((double)(16 * j), c1, (double)f0)));
    float2 x = make_float2(am * xv.x, am * xv.y);
    float2 rot = make_float2(1.0f, 0.0f);
    if constexpr (MT > 1) {
      if (i < n) rot = q.w[wave][i];
    }
    const float ys = yv.x + yv.y;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      if (m > 0) x = cmul(x, rot);  // X[a + 16, s] = X[a, s] exp(j 2 pi 256 dfs tau_s)
      dr[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, yv.x, dr[m], 0, 0, 0);
      di[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, yv.y, di[m], 0, 0, 0);
      ds[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x + x.y, ys, ds[m], 0, 0, 0);
    }
  }
}

// The synthesis for K <= 64 on v_mfma_f32_4x4x1f32 (16 blocks of 4 x 4 outer
// products; each lane 4 b + r gives row r of block b's A and column r of its
// B, and element i of its result is D[b][i][r]).  With k = 16 a + b, a < 4:
// block b = 4 q + c takes queued sample q of the 4-sample step and the
// columns 4 c .. 4 c + 3, so lane 16 q + 4 c + r forms X[a = r, s_q] and
// Y[s_q, b = 4 c + r] and accumulates partial sums of out[16 i + 4 c + r]
// (i < 4) over the samples of slot q; the four slots are summed once per ray.
// None of the output is idle (the 16 x 16 form uses 4 of its 16 rows here),
// and each real product takes a quarter of the 16x16x4 instruction's cycles.
// Three real products per complex one, as synth_mfma.
__device__ __forceinline__ void synth_mfma4(const FwdQueue& q, int wave, int lane, int n,
                                            f32x4& dr, f32x4& di, f32x4& ds) {
  const int r = lane & 3, slot = lane >> 4, col = lane & 15;  // col = 4 c + r
  for (int s0 = 0; s0 < n; s0 += 4) {
    const int i = s0 + slot;
    float am = 0.0f, f0 = 0.0f;
    double c1 = 0.0;
    if (i < n) {
      am = q.amp[wave][i];
      f0 = q.fr0[wave][i];
      c1 = q.cyc1[wave][i];
    }
    const float2 yv = phasor_cycles_hw((float)centered_frac((double)col * c1));
    const float2 xv =
        phasor_cycles_hw((float)centered_frac(fma((double)(16 * r), c1, (double)f0)));
    const float2 x = make_float2(am * xv.x, am * xv.y);
    dr = __builtin_amdgcn_mfma_f32_4x4x1f32(x.x, yv.x, dr, 0, 0, 0);
    di = __builtin_amdgcn_mfma_f32_4x4x1f32(x.y, yv.y, di, 0, 0, 0);
    ds = __builtin_amdgcn_mfma_f32_4x4x1f32(x.x + x.y, yv.x + yv.y, ds, 0, 0, 0);
  }
}

// Matrix-core synthesis with MT tiles of 256 frequencies (K <= 256 MT).
template <int MT>
__global__ __launch_bounds__(64 * kRaysPerBlock)
__attribute__((amdgpu_waves_per_eu(MT == 4 ? 6 : 1))) void lensless_forward_kernel(
    LenslessArgs a, float* __restrict__ out_real, float* __restrict__ out_imag,
    float* __restrict__ sigma_out, float* __restrict__ carry_out, PhaseConsts pc) {
  __shared__ FwdQueue q;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = blockIdx.x * kRaysPerBlock + wave;
  if (r >= a.R) return;  // only wave-level synchronisation below
  const Ray y = load_ray(a, r);

  // MT = 0: the K <= 64 form (synth_mfma4), one set of accumulators
  constexpr int NA = MT > 0 ? MT : 1;
  f32x4 dr[NA], di[NA], ds[NA];
#pragma unroll
  for (int m = 0; m < NA; ++m) dr[m] = di[m] = ds[m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
  auto drain = [&](int n) {
    if constexpr (MT == 0)
      synth_mfma4(q, wave, lane, n, dr[0], di[0], ds[0]);
    else
      synth_mfma<MT>(q, wave, lane, n, dr, di, ds);
  };

  float carry = 1.0f;
  int queued = 0;
  const int ntmax = (a.S + kTile - 1) / kTile;
  for (int base = 0; base < y.nv; base += kTile) {
    const int s = base + lane;
    const bool valid = s < y.nv;
    // the transmittance at the tile start, for the backward's single pass
    if (carry_out != nullptr && lane == 0) carry_out[(size_t)r * ntmax + base / kTile] = carry;
    const Sample x = load_composite(a, y, valid ? s : y.nv - 1);
    const float incl = wave_incl_prod_dpp(valid ? x.t : 1.0f);
    const float T = carry * dpp_mov<dpp::kWaveShr1>(1.0f, incl);
    carry *= readlane(incl, kWave - 1);
    const float Tc = T + x.off;
    const float sigma = fmaxf(x.refl * y.power * Tc, 1.0e-6f) * (x.alpha * Tc);
    if (sigma_out && valid) st_row(sigma_out + y.row, (unsigned)s, sigma);
    const Geo e = geometry(y, x.zz, x.g, valid);
    float amp = 0.0f;
    if (e.live) amp = sigma * e.spec * rcp_fast((float)(e.dpath * e.dpath));
    const bool emit = e.live && amp != 0.0f;
    const unsigned long long mask = __ballot(emit);
    if (emit) {
      const int slot = queued + lane_prefix(mask);
      const double tau = (e.dpath + pc.bias) * pc.inv_c;
      const double c1 = pc.dfs * tau;
      q.amp[wave][slot] = amp;
      q.fr0[wave][slot] = (float)centered_frac(pc.f0 * tau);
      q.cyc1[wave][slot] = c1;
      if constexpr (MT > 1)
        q.w[wave][slot] = phasor_cycles_hw((float)centered_frac(256.0 * c1));
    }
    queued += __popcll(mask);
    wave_sync();
    const int ndrain = queued >= kTile ? kTile : 0;
    if (ndrain > 0) {
      drain(ndrain);
      const int rest = queued - ndrain;
      float am = 0.0f, f0 = 0.0f;
      float2 w = make_float2(0.0f, 0.0f);
      double c1 = 0.0;
      if (lane < rest) {
        am = q.amp[wave][ndrain + lane];
        f0 = q.fr0[wave][ndrain + lane];
        if constexpr (MT > 1) w = q.w[wave][ndrain + lane];
        c1 = q.cyc1[wave][ndrain + lane];
      }
      wave_sync();
      if (lane < rest) {
        q.amp[wave][lane] = am;
        q.fr0[wave][lane] = f0;
        if constexpr (MT > 1) q.w[wave][lane] = w;
        q.cyc1[wave][lane] = c1;
      }
      queued = rest;
      wave_sync();
    }
  }
  if (queued > 0) drain(queued);

  const float inv_n = y.nv > 0 ? 1.0f / (float)y.nv : 0.0f;
  float* orow = out_real + (size_t)r * a.K;
  float* irow = out_imag + (size_t)r * a.K;
  // (Xr Yr - Xi Yi, (Xr + Xi)(Yr + Yi) - Xr Yr - Xi Yi)
#pragma unroll
  for (int m = 0; m < NA; ++m) {
    const f32x4 p1 = dr[m], p2 = di[m];
    dr[m] = p1 - p2;
    di[m] = ds[m] - p1 - p2;
  }
  if constexpr (MT == 0) {
    // the four sample slots (lanes l, l + 16, l + 32, l + 48) summed; lane
    // l < 16 holds out[16 i + l] in element i
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      dr[0][v] += __shfl_xor(dr[0][v], 16, kWave);
      di[0][v] += __shfl_xor(di[0][v], 16, kWave);
      dr[0][v] += __shfl_xor(dr[0][v], 32, kWave);
      di[0][v] += __shfl_xor(di[0][v], 32, kWave);
    }
    if (lane < 16) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int k = 16 * v + lane;
        if (k < a.K) {
          st_row(orow, (unsigned)k, dr[0][v] * inv_n);
          st_row(irow, (unsigned)k, di[0][v] * inv_n);
        }
      }
    }
    if (sigma_out)
      for (int s = y.nv + lane; s < a.S; s += kWave) st_row(sigma_out + y.row, (unsigned)s, 0.0f);
    return;
  }
  const int kl = 64 * (lane >> 4) + (lane & 15);
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int k = 256 * m + 16 * v + kl;
      if (k < a.K) {
        st_row(orow, (unsigned)k, dr[m][v] * inv_n);
        st_row(irow, (unsigned)k, di[m][v] * inv_n);
      }
    }
  if (sigma_out)
    for (int s = y.nv + lane; s < a.S; s += kWave) st_row(sigma_out + y.row, (unsigned)s, 0.0f);
}

// ------------------------------------------------------------------ backward

// Per-wave LDS layout of the backward kernel.
struct BwdLds {
  double* btau;  // [kQueue] tau of queued emitting samples
  int* bsid;     // [kQueue] their sample indices
  float* carry;  // [kMaxTiles] transmittance at each tile start
  float* h;      // [S] h_s (0 for samples that do not emit)
};
__host__ __device__ constexpr size_t bwd_lds_bytes_per_wave(int S) {
  return kQueue * sizeof(double) + kQueue * sizeof(int) + kMaxTiles * sizeof(float) +
         (size_t)S * sizeof(float);
}

// Matrix-core form of h_s (K > 128).  With k = 16 a + b,
//   sum_k g[k] exp(-j phi_sk) = sum_a conj(Xt[a,s]) T[a,s],
//   T[a,s] = sum_b G[a,b] conj(Y[s,b]),  G[a,b] = g[16 a + b],
//   Xt[a,s] = exp(+j 2 pi (f0 + 16 a dfs) tau_s),  Y[s,b] = exp(+j 2 pi b dfs tau_s):
// T is a product of the gradient row (the A operand, held in registers for
// the whole ray: tile m, depth step t, lane i + 16 k holds g[256 m + 16 i +
// 4 t + k]) and 16 queued samples' phasors (the B operand), four depth steps
// of v_mfma_f32_16x16x4f32 per 16-row tile; lane j + 16 q then holds
// T[16 m + 4 q + v, s_j] and contracts it with Xt (one phasor, rotations by
// exp(j 2 pi 16 dfs tau) and exp(j 2 pi 256 dfs tau)); the four lane quarters
// are summed with cross-lane shuffles.
template <int MT>
__device__ __forceinline__ void h_mfma(const double* __restrict__ btau,
                                       const int* __restrict__ bsid, float* __restrict__ hrow,
                                       int cnt, int lane, const float (*gr)[4],
                                       const float (*gi)[4], const PhaseConsts& pc,
                                       float inv_n) {
  // three real products per complex one for up to two tiles (four beyond:
  // the third accumulator set then costs more occupancy than it saves)
  constexpr bool kThree = MT <= 2;
  const int j = lane & 15, qd = lane >> 4;
  for (int s0 = 0; s0 < cnt; s0 += 16) {
    const int idx = s0 + j;
    const bool ok = idx < cnt;
    const double tau = ok ? btau[idx] : 0.0;
    const double c1 = pc.dfs * tau;
    f32x4 tr[MT], ti[MT], tq[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) tr[m] = ti[m] = tq[m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const float2 y = phasor_cycles_hw((float)centered_frac((double)(4 * t + qd) * c1));
      const float yi_neg = -y.y;  // conj(Y) = (y.x, -y.y)
      if constexpr (kThree) {
        // P1 = G_r Y_r, P2 = G_i (-Y_i), P3 = (G_r + G_i)(Y_r - Y_i)
        const float yd = y.x - y.y;
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          tr[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(gr[m][t], y.x, tr[m], 0, 0, 0);
          ti[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(gi[m][t], yi_neg, ti[m], 0, 0, 0);
          tq[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(gr[m][t] + gi[m][t], yd, tq[m], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int m = 0; m < MT; ++m) {
          tr[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(gr[m][t], y.x, tr[m], 0, 0, 0);
          ti[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(gr[m][t], yi_neg, ti[m], 0, 0, 0);
          tr[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(gi[m][t], y.y, tr[m], 0, 0, 0);
          ti[m] = __builtin_amdgcn_mfma_f32_16x16x4f32(gi[m][t], y.x, ti[m], 0, 0, 0);
        }
      }
    }
    if constexpr (kThree) {
#pragma unroll
      for (int m = 0; m < MT; ++m) {  // T = (P1 - P2, P3 - P1 - P2)
        const f32x4 p1 = tr[m], p2 = ti[m];
        tr[m] = p1 - p2;
        ti[m] = tq[m] - p1 - p2;
      }
    }
    float2 x = phasor_cycles_hw((float)centered_frac(fma((double)(64 * qd), c1, pc.f0 * tau)));
    const float2 st16 = phasor_cycles_hw((float)centered_frac(16.0 * c1));
    float2 st256 = make_float2(1.0f, 0.0f);
    if constexpr (MT > 1) st256 = phasor_cycles_hw((float)centered_frac(256.0 * c1));
    float h = 0.0f;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      float2 xm = x;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        h = fmaf(xm.x, tr[m][v], fmaf(xm.y, ti[m][v], h));
        if (v < 3) xm = cmul(xm, st16);
      }
      if (m + 1 < MT) x = cmul(x, st256);
    }
    h += __shfl_xor(h, 16, kWave);
    h += __shfl_xor(h, 32, kWave);
    if (ok && qd == 0) hrow[bsid[idx]] = h * inv_n;
  }
}

// The same h_s with FP16 hi/lo operands on v_mfma_f32_16x16x32_f16, for
// 16 queued samples at a time.  With k = 32 a + b (b < 32),
//   sum_k g[k] exp(-j phi_sk) = sum_a Xc[a,s] T[a,s],  T[a,s] = sum_b G[a,b] Yc[s,b],
//   Xc = exp(-j 2 pi (f0 + 32 a dfs) tau_s),  Yc = exp(-j 2 pi b dfs tau_s):
// the gradient row is the A operand (tile m of 16 rows of a = 512
// frequencies; lane i + 16 g holds G[16 m + i, 8 g .. 8 g + 7], split, scaled
// by a power of two for the ray), Yc the B operand (lane j + 16 g: sample j,
// b = 8 g .. 8 g + 7, one hardware phasor and a rotation tree), four real
// products of three instructions per tile; lane j + 16 q then contracts its
// four rows of T with Xc by Horner in exp(-j 2 pi 32 dfs tau) and the four
// lane quarters are summed with cross-lane shuffles.
template <int MT16>
__device__ __forceinline__ void h_mma16(const double* __restrict__ btau,
                                        const int* __restrict__ bsid, float* __restrict__ hrow,
                                        int cnt, int lane, const h8 (*gop)[4], float up,
                                        const PhaseConsts& pc, float inv_n) {
  const int j = lane & 15, g = lane >> 4;
  for (int s0 = 0; s0 < cnt; s0 += 16) {
    const int idx = s0 + j;
    const bool ok = idx < cnt;
    const double tau = ok ? btau[idx] : 0.0;
    const SplitCycles c = split_cycles13(pc.dfs * tau);
    const float r0 = (float)centered_frac(pc.f0 * tau);
    const float2 w1 = cis_neg(frac_multiple(1.0f, c));
    const float2 w2 = cmul(w1, w1), w4 = cmul(w2, w2);
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
    const float2 st = cis_neg(frac_multiple(32.0f, c));
    float h = 0.0f;
#pragma unroll
    for (int m = 0; m < MT16; ++m) {
      const h8& grh = gop[m][0];
      const h8& grl = gop[m][1];
      const h8& gih = gop[m][2];
      const h8& gil = gop[m][3];
      f32x4v tr = {0.0f, 0.0f, 0.0f, 0.0f}, ti = {0.0f, 0.0f, 0.0f, 0.0f};
      tr = mma3(grh, grl, yrh, yrl, tr);
      tr = mma3(-gih, -gil, yih, yil, tr);
      ti = mma3(grh, grl, yih, yil, ti);
      ti = mma3(gih, gil, yrh, yrl, ti);
      float2 H = make_float2(tr[3], ti[3]);
#pragma unroll
      for (int r = 2; r >= 0; --r) {
        const float2 t = cmul(H, st);
        H = make_float2(t.x + tr[r], t.y + ti[r]);
      }
      const float2 x = cis_neg(r0 + frac_multiple((float)(32 * (16 * m + 4 * g)), c));
      h = fmaf(x.x, H.x, fmaf(-x.y, H.y, h));
    }
    h += __shfl_xor(h, 16, kWave);
    h += __shfl_xor(h, 32, kWave);
    if (ok && g == 0) hrow[bsid[idx]] = h * up * inv_n;
  }
}

#ifndef GERAF_LL_H16
#define GERAF_LL_H16 1
#endif

// MT = 0: h_s by two-level Horner (one sample per lane); MT > 0: matrix-core
// form with MT tiles of 256 frequencies (K <= 256 MT).
template <int MT>
__global__ __launch_bounds__(64 * kRaysPerBlock)
__attribute__((amdgpu_waves_per_eu(MT <= 2 ? 5 : 1))) void lensless_backward_kernel(
    LenslessArgs a, const float* __restrict__ g_real, const float* __restrict__ g_imag,
    float* __restrict__ grad_sdf, float* __restrict__ grad_g, float* __restrict__ grad_refl,
    float* __restrict__ grad_inv_s, float* __restrict__ grad_power,
    const float* __restrict__ carry_in, PhaseConsts pc) {
  extern __shared__ double lds_raw[];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = blockIdx.x * kRaysPerBlock + wave;
  if (r >= a.R) return;  // only wave-level synchronisation below
  unsigned char* base_ptr =
      reinterpret_cast<unsigned char*>(lds_raw) + (size_t)wave * bwd_lds_bytes_per_wave(a.S);
  BwdLds L;
  L.btau = reinterpret_cast<double*>(base_ptr);
  L.bsid = reinterpret_cast<int*>(L.btau + kQueue);
  L.carry = reinterpret_cast<float*>(L.bsid + kQueue);
  L.h = L.carry + kMaxTiles;

  const Ray y = load_ray(a, r);
  const float inv_n = y.nv > 0 ? 1.0f / (float)y.nv : 0.0f;
  const int ntiles = (y.nv + kTile - 1) / kTile;
  const float* gr_row = g_real + (size_t)r * a.K;
  const float* gi_row = g_imag + (size_t)r * a.K;

  // h_s = (1/n) Re sum_k g[r,k] exp(-j phi_sk) for the first `cnt` queued
  // samples, one per lane (the gradient row is wave-uniform: scalar loads).
  constexpr bool kH16 = GERAF_LL_H16 && MT > 0;
  constexpr int kT = (MT && !kH16) ? MT : 1;
  float ga_r[kT][4], ga_i[kT][4];
  if constexpr (MT > 0 && !kH16) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const int k = 256 * m + 16 * (lane & 15) + 4 * t + (lane >> 4);
        ga_r[m][t] = k < a.K ? ld_row(gr_row, (unsigned)k) : 0.0f;
        ga_i[m][t] = k < a.K ? ld_row(gi_row, (unsigned)k) : 0.0f;
      }
  }
  // FP16 operands of the gradient row: 512-frequency tiles, lane i + 16 g
  // holds g[512 m + 32 i + 8 g .. + 7], scaled by 2^-e (2^e > max |g|)
  constexpr int kT16 = kH16 ? (MT + 1) / 2 : 1;
  h8 gop[kT16][4];
  float g_up = 1.0f;
  if constexpr (kH16) {
    float2 gv[kT16][8];
    float mx = 0.0f;
    // a lane's eight elements are contiguous: two 16-byte loads per row when
    // the rows are 16-byte aligned and K is a multiple of 8 (then the eight
    // are all in range or all out), else one element at a time
    const bool vec = (a.K % 8 == 0) &&
                     ((reinterpret_cast<uintptr_t>(gr_row) | reinterpret_cast<uintptr_t>(gi_row)) &
                      15u) == 0;
#pragma unroll
    for (int m = 0; m < kT16; ++m) {
      const int k0 = 512 * m + 32 * (lane & 15) + 8 * (lane >> 4);
      if (vec) {
        float4 r0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), r1 = r0, i0 = r0, i1 = r0;
        if (k0 < a.K) {
          const float4* pr = reinterpret_cast<const float4*>(gr_row + k0);
          const float4* pi = reinterpret_cast<const float4*>(gi_row + k0);
          r0 = pr[0];
          r1 = pr[1];
          i0 = pi[0];
          i1 = pi[1];
        }
        gv[m][0] = make_float2(r0.x, i0.x);
        gv[m][1] = make_float2(r0.y, i0.y);
        gv[m][2] = make_float2(r0.z, i0.z);
        gv[m][3] = make_float2(r0.w, i0.w);
        gv[m][4] = make_float2(r1.x, i1.x);
        gv[m][5] = make_float2(r1.y, i1.y);
        gv[m][6] = make_float2(r1.z, i1.z);
        gv[m][7] = make_float2(r1.w, i1.w);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + e;
          gv[m][e] = k < a.K ? make_float2(ld_row(gr_row, (unsigned)k),
                                           ld_row(gi_row, (unsigned)k))
                             : make_float2(0.0f, 0.0f);
        }
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) mx = fmaxf(mx, fmaxf(fabsf(gv[m][e].x), fabsf(gv[m][e].y)));
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, kWave));
    int ex = 0;
    if (mx > 0.0f && isfinite(mx)) (void)frexpf(mx, &ex);
    const float down = ldexpf(1.0f, -ex);
    g_up = ldexpf(1.0f, ex);
#pragma unroll
    for (int m = 0; m < kT16; ++m) {
#pragma unroll
      for (int e = 0; e < 8; ++e) gv[m][e] = make_float2(gv[m][e].x * down, gv[m][e].y * down);
      split8(gv[m], gop[m][0], gop[m][1], gop[m][2], gop[m][3]);
    }
  }
  auto horner_batch = [&](int cnt) {
    if constexpr (MT == 0) {
      if (lane < cnt) {
        const PathPhasors ph = path_phasors(L.btau[lane], pc, -1.0f);
        const float2 S = horner2(gr_row, gi_row, a.K, ph.w, ph.W);
        L.h[L.bsid[lane]] = cmul(ph.base, S).x * inv_n;
      }
    } else if constexpr (kH16) {
      h_mma16<kT16>(L.btau, L.bsid, L.h, cnt, lane, gop, g_up, pc, inv_n);
    } else {
      h_mfma<MT>(L.btau, L.bsid, L.h, cnt, lane, ga_r, ga_i, pc, inv_n);
    }
  };

  // pass A (forward over tiles): transmittance carries, and h_s of every
  // emitting sample, evaluated in full 64-lane batches of such samples.
  // With the forward's carries (carry_in) there is no pass A: pass B then
  // evaluates h_s tile by tile (each sample's h_s depends on its own delay
  // alone, so the results are the same bit for bit).
  const bool one_pass = carry_in != nullptr;
  const int ntmax = (a.S + kTile - 1) / kTile;
  if (!one_pass) {
  for (int s = lane; s < y.nv; s += kWave) L.h[s] = 0.0f;
  wave_sync();
  float carry = 1.0f;
  int queued = 0;
  for (int tile = 0; tile < ntiles; ++tile) {
    if (lane == 0) L.carry[tile] = carry;
    const int s = tile * kTile + lane;
    const bool valid = s < y.nv;
    const Sample x = load_composite(a, y, valid ? s : y.nv - 1);
    const float incl = wave_incl_prod_dpp(valid ? x.t : 1.0f);
    carry *= readlane(incl, kWave - 1);
    const Geo e = geometry(y, x.zz, x.g, valid);
    const unsigned long long mask = __ballot(e.live);
    if (e.live) {
      const int slot = queued + lane_prefix(mask);
      L.btau[slot] = (e.dpath + pc.bias) * pc.inv_c;
      L.bsid[slot] = s;
    }
    queued += __popcll(mask);
    wave_sync();
    const int ndrain = queued >= kWave ? kWave : 0;
    if (ndrain > 0) {
      horner_batch(ndrain);
      const int rest = queued - ndrain;
      double t_keep = 0.0;
      int s_keep = 0;
      if (lane < rest) {
        t_keep = L.btau[ndrain + lane];
        s_keep = L.bsid[ndrain + lane];
      }
      wave_sync();
      if (lane < rest) {
        L.btau[lane] = t_keep;
        L.bsid[lane] = s_keep;
      }
      queued = rest;
      wave_sync();
    }
  }
  if (queued > 0) horner_batch(queued);
  wave_sync();
  }

  // pass B: reverse over tiles
  float w_carry = 0.0f;  // W at the first sample of the next tile
  float gis = 0.0f, gpow = 0.0f;
  float* gs_row = grad_sdf + y.row;
  float* gg_row = grad_g + 3 * y.row;
  float* gr_out = grad_refl ? grad_refl + y.row : nullptr;
  for (int tile = ntiles - 1; tile >= 0; --tile) {
    const int s = tile * kTile + lane;
    const bool valid = s < y.nv;
    // the tile's inputs and geometry, once: they serve both the h_s batch
    // and the gradients
    const RawSample w = load_raw(a, y, valid ? s : y.nv - 1);
    const Geo e = geometry(y, w.zz, w.g, valid);
    if (one_pass) {
      // h_s of this tile's emitting samples (which are known from z and the
      // SDF gradient alone), before the compositing terms are formed
      const unsigned long long mask = __ballot(e.live);
      if (e.live) {
        const int slot = lane_prefix(mask);
        L.btau[slot] = (e.dpath + pc.bias) * pc.inv_c;
        L.bsid[slot] = s;
      }
      wave_sync();
      const int cnt = __popcll(mask);
      if (cnt > 0) horner_batch(cnt);
      wave_sync();
    }
    const Sample x = composite(a, y, w);
    const float tt = valid ? x.t : 1.0f;
    const float incl = wave_incl_prod_dpp(tt);
    const float tile_T = one_pass ? carry_in[(size_t)r * ntmax + tile] : L.carry[tile];
    const float T = tile_T * dpp_mov<dpp::kWaveShr1>(1.0f, incl);
    const float Tc = T + x.off;
    const float rp = x.refl * y.power;
    const float m = rp * Tc;
    const bool unclamped = m >= 1.0e-6f;
    const float aTc = x.alpha * Tc;
    const float sigma = fmaxf(m, 1.0e-6f) * aTc;

    float h = 0.0f, att = 0.0f;
    if (e.live) {
      att = rcp_fast((float)(e.dpath * e.dpath));
      h = L.h[s];
    }
    const float gsig = h * e.spec * att;
    const float gspec = h * sigma * att;
    const V3 gnhat = e.live ? gspec * dspec_dn_mono(y, e) : v3(0.0f, 0.0f, 0.0f);
    float grefl, gTc, galpha;
    if (unclamped) {
      grefl = gsig * y.power * Tc * aTc;
      gpow += gsig * x.refl * Tc * aTc;
      gTc = gsig * 2.0f * m * x.alpha;
      galpha = gsig * m * Tc;
    } else {
      grefl = 0.0f;
      gTc = gsig * 1.0e-6f * x.alpha;
      galpha = gsig * 1.0e-6f * Tc;
    }
    if (!valid) gTc = 0.0f;
    // reverse affine scan: W_l = gT_l + t_l * W_{l+1}
    float w_next, w_head;
    wave_affine_suffix(tt, gTc, w_carry, lane, &w_next, &w_head);
    w_carry = w_head;

    galpha -= T * w_next;
    const float gq = (valid && x.qmask) ? galpha : 0.0f;
    const float rden = rcp_fast(x.pc + 1.0e-5f);
    const float gpc = gq * x.nc * (rden * rden);
    const float gnc = -gq * rden;
    const float gprev = gpc * y.inv_s * x.sp;
    const float gnext = gnc * y.inv_s * x.sn;
    gis += gpc * x.prev * x.sp + gnc * x.next * x.sn;
    const float gsdf = gprev + gnext;
    const float gic = (gnext - gprev) * x.dist * 0.5f;
    const float gc = gic * ((x.c < 1.0f ? 0.5f * (1.0f - a.anneal) : 0.0f) +
                            (x.c < 0.0f ? a.anneal : 0.0f));
    V3 gg;
    if (e.gl > 1.0e-12f)
      gg = rcp_fast(e.gl) * (gnhat - dot(e.n, gnhat) * e.n);
    else
      gg = 1.0e12f * gnhat;
    gg = gg + gc * y.dir;
    if (valid) {
      const unsigned u = (unsigned)s;
      st_row(gs_row, u, gsdf);
      st_row(gg_row, 3u * u, gg.x);
      st_row(gg_row, 3u * u + 1u, gg.y);
      st_row(gg_row, 3u * u + 2u, gg.z);
      if (gr_out) st_row(gr_out, u, grefl);
    }
  }
  for (int s = y.nv + lane; s < a.S; s += kWave) {
    const unsigned u = (unsigned)s;
    st_row(gs_row, u, 0.0f);
    st_row(gg_row, 3u * u, 0.0f);
    st_row(gg_row, 3u * u + 1u, 0.0f);
    st_row(gg_row, 3u * u + 2u, 0.0f);
    if (gr_out) st_row(gr_out, u, 0.0f);
  }
  gis = wave_sum(gis);
  gpow = wave_sum(gpow);
  if (lane == 0) {
    grad_inv_s[r] = gis;
    grad_power[r] = gpow;
  }
}

LenslessArgs make_args(const float* sdf, const float* grad, const float* refl,
                       const float* dists, const float* z, const float* origin,
                       const float* dirs, const float* offset, const int32_t* n_valid,
                       const float* inv_s, const float* power, int64_t R, int64_t S, int64_t K,
                       float anneal) {
  if (R < 0 || S < 0 || K < 0) throw ArgError("sizes must be non-negative");
  if (R > INT32_MAX / 2 || R * S > (INT64_C(1) << 36)) throw ArgError("too many samples");
  // S <= 2048 keeps the backward's LDS (4 waves x (1.8 KB + 4 S B)) under 64 KB
  if (S > 2048) throw ArgError("at most 2048 samples per ray");
  if (!(anneal >= 0.0f && anneal <= 1.0f)) throw ArgError("cos_anneal_ratio must be in [0,1]");
  require_ptrs({origin, dirs, inv_s, power}, R > 0);
  require_ptrs({sdf, grad, dists, z}, R > 0 && S > 0);
  LenslessArgs a;
  a.sdf = sdf;
  a.grad = grad;
  a.refl = refl;
  a.dists = dists;
  a.z = z;
  a.origin = origin;
  a.dirs = dirs;
  a.offset = offset;
  a.n_valid = n_valid;
  a.inv_s = inv_s;
  a.power = power;
  a.R = (int)R;
  a.S = (int)S;
  a.K = (int)K;
  a.anneal = anneal;
  return a;
}

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int geraf_lensless_forward_f32(
    const float* sdf, const float* grad, const float* refl, const float* dists, const float* z,
    const float* origin, const float* dirs, const float* trans_offset, const int32_t* n_valid,
    const float* inv_s, const float* power, float* out_real, float* out_imag, float* sigma_out,
    float* tile_carry, int64_t R, int64_t S, int64_t K, float cos_anneal_ratio,
    const geraf_fmcw_params* params, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_params(params);
  const LenslessArgs a = make_args(sdf, grad, refl, dists, z, origin, dirs, trans_offset,
                                   n_valid, inv_s, power, R, S, K, cos_anneal_ratio);
  if (K > 2048) throw ArgError("at most 2048 frequencies per ray");
  if (R == 0 || K == 0) return GERAF_RF_OK;
  require_ptrs({out_real, out_imag}, true);
  if (S == 0) throw ArgError("need at least one sample per ray");
  const unsigned blocks = (unsigned)((R + kRaysPerBlock - 1) / kRaysPerBlock);
  const PhaseConsts pc = make_phase_consts(*params);
  const dim3 grid(blocks), block(64 * kRaysPerBlock);
  hipStream_t st = (hipStream_t)stream;
  if (K <= 64)
    hipLaunchKernelGGL(lensless_forward_kernel<0>, grid, block, 0, st, a, out_real, out_imag,
                       sigma_out, tile_carry, pc);
  else if (K <= 256)
    hipLaunchKernelGGL(lensless_forward_kernel<1>, grid, block, 0, st, a, out_real, out_imag,
                       sigma_out, tile_carry, pc);
  else if (K <= 512)
    hipLaunchKernelGGL(lensless_forward_kernel<2>, grid, block, 0, st, a, out_real, out_imag,
                       sigma_out, tile_carry, pc);
  else if (K <= 1024)
    hipLaunchKernelGGL(lensless_forward_kernel<4>, grid, block, 0, st, a, out_real, out_imag,
                       sigma_out, tile_carry, pc);
  else
    hipLaunchKernelGGL(lensless_forward_kernel<8>, grid, block, 0, st, a, out_real, out_imag,
                       sigma_out, tile_carry, pc);
  return check_launch("lensless_forward_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_lensless_backward_f32(
    const float* grad_out_real, const float* grad_out_imag, const float* sdf, const float* grad,
    const float* refl, const float* dists, const float* z, const float* origin,
    const float* dirs, const float* trans_offset, const int32_t* n_valid, const float* inv_s,
    const float* power, float* grad_sdf, float* grad_grad, float* grad_refl,
    float* grad_inv_s_per_ray, float* grad_power_per_ray, const float* tile_carry, int64_t R,
    int64_t S, int64_t K, float cos_anneal_ratio, const geraf_fmcw_params* params,
    geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  check_params(params);
  const LenslessArgs a = make_args(sdf, grad, refl, dists, z, origin, dirs, trans_offset,
                                   n_valid, inv_s, power, R, S, K, cos_anneal_ratio);
  if (K > 2048) throw ArgError("at most 2048 frequencies per ray");
  if (R == 0) return GERAF_RF_OK;
  require_ptrs({grad_inv_s_per_ray, grad_power_per_ray}, true);
  require_ptrs({grad_sdf, grad_grad}, S > 0);
  require_ptrs({grad_out_real, grad_out_imag}, K > 0);
  if (S == 0) throw ArgError("need at least one sample per ray");
  const unsigned blocks = (unsigned)((R + kRaysPerBlock - 1) / kRaysPerBlock);
  const size_t lds = kRaysPerBlock * bwd_lds_bytes_per_wave((int)S);
  const PhaseConsts pc = make_phase_consts(*params);
  const dim3 grid(blocks), block(64 * kRaysPerBlock);
  hipStream_t st = (hipStream_t)stream;
#define GERAF_LL_BWD(MT_)                                                                      \
  hipLaunchKernelGGL(lensless_backward_kernel<MT_>, grid, block, lds, st, a, grad_out_real,   \
                     grad_out_imag, grad_sdf, grad_grad, grad_refl, grad_inv_s_per_ray,        \
                     grad_power_per_ray, tile_carry, pc)
  // h_s: two-level Horner up to K = 128, where it measured 5-26 % faster
  // than the FP16 matrix form
  if (K <= 128)
    GERAF_LL_BWD(0);
  else if (K <= 256)
    GERAF_LL_BWD(1);
  else if (K <= 512)
    GERAF_LL_BWD(2);
  else if (K <= 1024)
    GERAF_LL_BWD(4);
  else
    GERAF_LL_BWD(8);
#undef GERAF_LL_BWD
  return check_launch("lensless_backward_kernel");
  GERAF_TRY_END
}
