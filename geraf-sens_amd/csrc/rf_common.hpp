// Shared device helpers for the GeRaF RF kernels (gfx950).
//
// Phase model shared by the matched filter and the signal tracer
// (MatchedFilter_cuda_kernel.cu:28-63, SignalTracer_cuda_kernel.cu:433-443):
//   tau      = (|tx - p| + |rx - p| + path_bias) / c          [1e-8 s]
//   phi_k    = 2*pi * (f0 + k * slope / fs) * tau              [rad]
// with f0 in 1e8 Hz, slope in 1e8 Hz/s, fs in Hz and c in 1e8 m/s.  All phase
// work here is done in cycles: cycles_k = f0*tau + k*(slope/fs)*tau.  The
// carrier term f0*tau is a few hundred cycles, so it is formed and reduced
// modulo 1 in FP64; everything after the reduction runs in FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "geraf_rf.h"

namespace geraf {

constexpr int kWave = 64;

// Host-side derived constants handed to every kernel by value.
struct PhaseConsts {
  double f0;     // chirp start frequency [1e8 Hz]
  double dfs;    // frequency step per ADC sample, slope / fs [1e8 Hz]
  double inv_c;  // 1 / c  [1e-8 s / m]
  double bias;   // path-length bias added before the phase [m]
};

inline PhaseConsts make_phase_consts(const geraf_fmcw_params& p) {
  PhaseConsts c;
  c.f0 = p.f0;
  c.dfs = p.slope / p.fs;
  c.inv_c = 1.0 / p.c;
  c.bias = p.path_bias;
  return c;
}

__device__ __forceinline__ double dist3_f64(double ax, double ay, double az, double bx,
                                            double by, double bz) {
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  return sqrt(fma(dx, dx, fma(dy, dy, dz * dz)));
}

// sqrt(D) for D >= 0 to ~1e-14 relative: an FP32 reciprocal-square-root seed
// (about 1 ulp, 6e-8) refined by one FP64 Newton step (error ~1.5 * 6e-8^2).
// Far cheaper than a correctly rounded FP64 sqrt.  The seed's argument is
// held at or above FLT_MIN, so D == 0 gives a finite seed and the result 0
// without a select; D below FLT_MIN (distances under 1e-19) is not resolved.
__device__ __forceinline__ double sqrt_f64_newton(double D) {
  const double y = (double)__builtin_amdgcn_rsqf(fmaxf((float)D, 1.17549435e-38f));
  const double h = 0.5 * D * (y * y);
  return D * fma(y, 0.5 - h, y);  // D * y * (1.5 - h)
}

__device__ __forceinline__ double dist3_f64_fast(double ax, double ay, double az, double bx,
                                                 double by, double bz) {
  const double dx = ax - bx, dy = ay - by, dz = az - bz;
  return sqrt_f64_newton(fma(dx, dx, fma(dy, dy, dz * dz)));
}

// Whether antenna a transmits and receives at one position (the three
// coordinates equal bit for bit): then both legs of the round trip are the
// same distance, computed once.  The loads are wave-uniform (scalar), and so
// is the result, which lets the caller branch instead of selecting.
__device__ __forceinline__ bool monostatic(const float* __restrict__ tx_pos,
                                           const float* __restrict__ rx_pos, int a) {
  const unsigned* t = reinterpret_cast<const unsigned*>(tx_pos) + 3 * a;
  const unsigned* r = reinterpret_cast<const unsigned*>(rx_pos) + 3 * a;
  const int same = (t[0] == r[0]) & (t[1] == r[1]) & (t[2] == r[2]);
  return __builtin_amdgcn_readfirstlane(same) != 0;
}

// |t - p| + |r - p| (FP64, dist3_f64_fast).  With Mono the second leg is the
// first one again: d + d, bit for bit what the two-leg sum gives when t == r.
template <bool Mono>
__device__ __forceinline__ double round_trip_f64(double tx, double ty, double tz, double rx,
                                                 double ry, double rz, double px, double py,
                                                 double pz) {
  const double dt = dist3_f64_fast(tx, ty, tz, px, py, pz);
  if constexpr (Mono) return dt + dt;
  return dt + dist3_f64_fast(rx, ry, rz, px, py, pz);
}

// Fractional part of x in [-0.5, 0.5].
__device__ __forceinline__ double centered_frac(double x) { return x - rint(x); }

// (cos(2*pi*x), sign*sin(2*pi*x)) for finite |x| <= 0.5 (x in cycles, already
// range-reduced in FP64 by the caller).  Quadrant reduction q = rint(4x),
// r = x - q/4 (exact), theta = 2*pi*r in [-pi/4, pi/4]; Taylor polynomials of
// degree 9 (sin) and 10 (cos), truncation < 3e-9; then the quadrant rotation.
// About 1 ulp, with none of the general sincospif's range reduction and
// special-case handling.
__device__ __forceinline__ float2 phasor_cycles(float x, float sign) {
  const float qf = rintf(4.0f * x);
  const float r = fmaf(qf, -0.25f, x);
  const float t = 6.28318530717958647692f * r;
  const float t2 = t * t;
  float ps = fmaf(t2, 2.75573192e-6f, -1.98412698e-4f);
  ps = fmaf(t2, ps, 8.33333333e-3f);
  ps = fmaf(t2, ps, -1.66666667e-1f);
  const float sn = fmaf(t * t2, ps, t);
  float pc = fmaf(t2, -2.75573192e-7f, 2.48015873e-5f);
  pc = fmaf(t2, pc, -1.38888889e-3f);
  pc = fmaf(t2, pc, 4.16666667e-2f);
  pc = fmaf(t2, pc, -0.5f);
  const float cs = fmaf(t2, pc, 1.0f);
  const int q = (int)qf;
  // rotate (cs, sn) by q quarter turns
  const bool odd = q & 1;
  float s2 = odd ? cs : sn;
  float c2 = odd ? sn : cs;
  if (q & 2) s2 = -s2;
  if ((q + 1) & 2) c2 = -c2;
  return make_float2(c2, sign * s2);
}

// (cos(2*pi*x), sin(2*pi*x)) from the hardware sine and cosine, whose
// argument is in revolutions, for |x| <= 0.5 (FP64-reduced by the caller).
// Measured on gfx950 over 2^22 points of [-0.5, 0.5]: max abs error 1.25e-7,
// rms 3.5e-8 (two quarter-rate instructions instead of ~20 VALU).
__device__ __forceinline__ float2 phasor_cycles_hw(float x) {
  return make_float2(__builtin_amdgcn_cosf(x), __builtin_amdgcn_sinf(x));
}

// A phase step in cycles split for exact small multiples in FP32 (Cody-Waite):
// hi keeps 16 significant bits, so m * hi is exact for integers 0 <= m < 256;
// lo = c - hi (|lo| <= 2^-16 |c|).
struct SplitCycles {
  float hi, lo;
};
__device__ __forceinline__ SplitCycles split_cycles(double c) {
  SplitCycles s;
  s.hi = __uint_as_float(__float_as_uint((float)c) & 0xFFFFFF00u);
  s.lo = (float)(c - (double)s.hi);
  return s;
}
// A phase step split for exact multiples up to 2^11 in FP32: hi keeps 13
// significant bits (m * hi has at most 24), lo = c - hi.
__device__ __forceinline__ SplitCycles split_cycles13(double c) {
  SplitCycles s;
  s.hi = __uint_as_float(__float_as_uint((float)c) & 0xFFFFF800u);
  s.lo = (float)(c - (double)s.hi);
  return s;
}
// Centered fraction of m * c for an integer 0 <= m < 256: m * hi, its rint
// and their difference are exact; only m * lo and the final add round (about
// 3e-8 cycles, as a cast of the FP64 result to FP32 would).  |result| <= 0.5 + tiny.
__device__ __forceinline__ float frac_multiple(float m, const SplitCycles& c) {
  const float p = m * c.hi;
  return fmaf(m, c.lo, p - rintf(p));
}

// ---------------------------------------------- FP16 hi/lo matrix operands
// Operands of v_mfma_f32_16x16x32_f16: eight FP16 values per lane.  An FP32
// value x is carried as hi = fp16(x) and lo = fp16(x - hi); a real product
// is hi*hi + hi*lo + lo*hi into one FP32 accumulator (the dropped lo*lo is
// ~2^-22 relative), i.e. the products keep FP32 operand precision.
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f32x4v __attribute__((ext_vector_type(4)));

// hi/lo of the real and imaginary parts of eight complex values: hi two at
// a time by one packed conversion, lo = x - hi as one mixed-precision FMA
// (FP16 hi operand, FP32 x) per value, two results per packed conversion.
// The multiplier -1 is hidden from the compiler (an empty asm that claims to
// modify it), which would otherwise fold the FMA into a subtraction that
// first converts hi back to FP32.
__device__ __forceinline__ void split8(const float2* y, h8& rh, h8& rl, h8& ih, h8& il) {
  float negone = -1.0f;
  asm("" : "+v"(negone));
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    const h2 r = __builtin_convertvector((f2v){y[e].x, y[e + 1].x}, h2);
    const h2 i = __builtin_convertvector((f2v){y[e].y, y[e + 1].y}, h2);
    rh[e] = r[0];
    rh[e + 1] = r[1];
    ih[e] = i[0];
    ih[e + 1] = i[1];
    rl[e] = (_Float16)fmaf((float)r[0], negone, y[e].x);
    rl[e + 1] = (_Float16)fmaf((float)r[1], negone, y[e + 1].x);
    il[e] = (_Float16)fmaf((float)i[0], negone, y[e].y);
    il[e + 1] = (_Float16)fmaf((float)i[1], negone, y[e + 1].y);
  }
}

// hi/lo of four complex values as interleaved (Re, Im) pairs: element 2u is
// Re y[u], element 2u + 1 is Im y[u] (same conversions as split8).
__device__ __forceinline__ void split4c(const float2* y, h8& h, h8& l) {
  float negone = -1.0f;
  asm("" : "+v"(negone));
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const h2 c = __builtin_convertvector((f2v){y[u].x, y[u].y}, h2);
    h[2 * u] = c[0];
    h[2 * u + 1] = c[1];
    l[2 * u] = (_Float16)fmaf((float)c[0], negone, y[u].x);
    l[2 * u + 1] = (_Float16)fmaf((float)c[1], negone, y[u].y);
  }
}

// c += A B for split operands (three instructions).
__device__ __forceinline__ f32x4v mma3(const h8& ah, const h8& al, const h8& bh, const h8& bl,
                                       f32x4v c) {
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, c, 0, 0, 0);
  return c;
}

// exp(-j 2 pi x) from the hardware sine and cosine, x reduced to [-0.5, 0.5].
__device__ __forceinline__ float2 cis_neg(float x) {
  const float r = x - rintf(x);
  return make_float2(__builtin_amdgcn_cosf(r), -__builtin_amdgcn_sinf(r));
}

// Orders one wave's LDS writes before its later LDS reads (a wave-level
// barrier with release/acquire fences; no block synchronisation).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// exp(-j 2 pi x) for an x already reduced to about [-0.5, 0.5] (e.g. by
// frac_multiple): the hardware sine and cosine without a second reduction.
__device__ __forceinline__ float2 cis_neg_reduced(float x) {
  return make_float2(__builtin_amdgcn_cosf(x), -__builtin_amdgcn_sinf(x));
}

// Complex helpers on float2 = (re, im).
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

// a^2 (three instructions).
__device__ __forceinline__ float2 csq(float2 a) {
  return make_float2(fmaf(a.x, a.x, -a.y * a.y), 2.0f * a.x * a.y);
}

// a * b + c as two FMA chains (four instructions).
__device__ __forceinline__ float2 cmul_add(float2 a, float2 b, float2 c) {
  return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, c.y)));
}

// Compensated (Kahan) accumulation of a complex value.
struct KahanC {
  float sr, si, cr, ci;
  __device__ __forceinline__ void init() { sr = si = cr = ci = 0.0f; }
  __device__ __forceinline__ void add(float2 v) {
    const float yr = v.x - cr;
    const float tr = sr + yr;
    cr = (tr - sr) - yr;
    sr = tr;
    const float yi = v.y - ci;
    const float ti = si + yi;
    ci = (ti - si) - yi;
    si = ti;
  }
  // Best FP64 estimate of the compensated sum.
  __device__ __forceinline__ double re() const { return (double)sr - (double)cr; }
  __device__ __forceinline__ double im() const { return (double)si - (double)ci; }
};

}  // namespace geraf

namespace geraf {

// ------------------------------------------------------------------ 3-vectors

struct V3 {
  float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 load3(const float* p) { return v3(p[0], p[1], p[2]); }

// v / (|v| + 1e-7), the normalisation of SignalTracer_cuda_kernel.cu:40-47.
__device__ __forceinline__ V3 normalize_eps(V3 v) {
  const float inv = 1.0f / (sqrtf(dot(v, v)) + 1.0e-7f);
  return inv * v;
}

// Specular geometry of one (tx, rx, point, normal) tuple, as
// ComputeSpecularGeometry (SignalTracer_cuda_kernel.cu:172-197): the normal is
// used as given (not re-normalised).
struct SpecGeom {
  V3 inc;            // normalised incident direction (point - tx)
  V3 ret;            // normalised return direction (rx - point)
  V3 refl;           // normalised reflected direction
  float refl_norm;   // |inc - 2 (inc.n) n|
  float idn;         // inc . n
  float spec;        // ret . refl
  __device__ __forceinline__ bool culled() const { return spec < 1.0e-6f || idn > 0.0f; }
};

__device__ __forceinline__ SpecGeom spec_geom(V3 tx, V3 rx, V3 n, V3 p) {
  SpecGeom g;
  g.inc = normalize_eps(p - tx);
  g.idn = dot(g.inc, n);
  const V3 rr = g.inc - (2.0f * g.idn) * n;
  g.refl_norm = sqrtf(dot(rr, rr));
  g.refl = (1.0f / (g.refl_norm + 1.0e-7f)) * rr;
  g.ret = normalize_eps(rx - p);
  g.spec = dot(g.ret, g.refl);
  return g;
}

// d spec / d n with the return and incident directions held fixed: the product
// of the three Jacobians of SignalTracer_cuda_kernel.cu:91-170,
//   J = -2 n inc^T - 2 (inc.n) I,  NJ = (I - refl refl^T) J / (|refl_raw| + eps),
//   grad = NJ^T ret,
// folded to grad = -2 [ (q.n) inc + (inc.n) q ] / (|refl_raw| + eps) with
// q = ret - spec * refl.
__device__ __forceinline__ V3 dspec_dnormal(const SpecGeom& g, V3 n) {
  const V3 q = g.ret - g.spec * g.refl;
  const float s = -2.0f / (g.refl_norm + 1.0e-7f);
  return s * (dot(q, n) * g.inc + g.idn * q);
}

__device__ __forceinline__ double dist3_f64(V3 a, V3 b) {
  return dist3_f64((double)a.x, (double)a.y, (double)a.z, (double)b.x, (double)b.y, (double)b.z);
}

// ------------------------------------------------------------ cross-lane (DPP)

// v_mov_b32_dpp: each enabled lane reads `src` from the lane the DPP control
// names; lanes whose source lies outside the row/wave, or which the row/bank
// masks disable, keep `old` (bound_ctrl off).
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ float dpp_mov(float old, float src) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, old),
                                                               __builtin_bit_cast(int, src), CTRL,
                                                               ROW_MASK, BANK_MASK, false));
}
namespace dpp {
constexpr int kQuadXor1 = 0xB1;    // quad_perm [1,0,3,2]
constexpr int kQuadXor2 = 0x4E;    // quad_perm [2,3,0,1]
constexpr int kRowShl = 0x100;     // + n: lane l reads l+n within its row of 16
constexpr int kRowShr = 0x110;     // + n: lane l reads l-n within its row of 16
constexpr int kWaveShl1 = 0x130;   // lane l reads l+1 (lane 63 keeps old)
constexpr int kWaveShr1 = 0x138;   // lane l reads l-1 (lane 0 keeps old)
constexpr int kRowMirror = 0x140;  // lane l reads 15-l within its row
constexpr int kRowHalfMirror = 0x141;  // lane l reads 7-l within its half-row
constexpr int kRowBcast15 = 0x142;     // lane 15 of each row feeds the next row
constexpr int kRowBcast31 = 0x143;     // lane 31 feeds rows 2 and 3
}  // namespace dpp

__device__ __forceinline__ float readlane(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// Inclusive prefix product over the 64 lanes (row shifts, then the row
// broadcasts 15 and 31).
__device__ __forceinline__ float wave_incl_prod_dpp(float v) {
  float p = v * dpp_mov<dpp::kRowShr + 1>(1.0f, v);
  p *= dpp_mov<dpp::kRowShr + 2>(1.0f, v);
  p *= dpp_mov<dpp::kRowShr + 3>(1.0f, v);
  p *= dpp_mov<dpp::kRowShr + 4, 0xf, 0xe>(1.0f, p);
  p *= dpp_mov<dpp::kRowShr + 8, 0xf, 0xc>(1.0f, p);
  p *= dpp_mov<dpp::kRowBcast15, 0xa, 0xf>(1.0f, p);
  p *= dpp_mov<dpp::kRowBcast31, 0xc, 0xf>(1.0f, p);
  return p;
}

// Sum over the 64 lanes, returned wave-uniform: DPP sums within each row of
// 16 lanes, then the four row totals through readlane (fixed order).
__device__ __forceinline__ float wave_total_dpp(float v) {
  v += dpp_mov<dpp::kQuadXor1>(0.0f, v);
  v += dpp_mov<dpp::kQuadXor2>(0.0f, v);
  v += dpp_mov<dpp::kRowHalfMirror>(0.0f, v);
  v += dpp_mov<dpp::kRowMirror>(0.0f, v);
  return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}

// Maximum over the 64 lanes, returned wave-uniform (as wave_total_dpp).
__device__ __forceinline__ float wave_max_dpp(float v) {
  v = fmaxf(v, dpp_mov<dpp::kQuadXor1>(v, v));
  v = fmaxf(v, dpp_mov<dpp::kQuadXor2>(v, v));
  v = fmaxf(v, dpp_mov<dpp::kRowHalfMirror>(v, v));
  v = fmaxf(v, dpp_mov<dpp::kRowMirror>(v, v));
  return fmaxf(fmaxf(readlane(v, 0), readlane(v, 16)), fmaxf(readlane(v, 32), readlane(v, 48)));
}

// Suffix composition of the affine maps W_l = b_l + m_l * W_{l+1} over the
// wave, with W_64 = tail.  Returns W_l; *w_next = W_{l+1} and *head = W_0.
// Rows of 16 are scanned with row_shl steps; the four row results are joined
// through readlane.
__device__ __forceinline__ float wave_affine_suffix(float m, float b, float tail, int lane,
                                                    float* w_next, float* head) {
#define GERAF_AFFINE_STEP(N)                                  \
  {                                                           \
    const float m2 = dpp_mov<dpp::kRowShl + N>(1.0f, m);      \
    const float b2 = dpp_mov<dpp::kRowShl + N>(0.0f, b);      \
    b = fmaf(m, b2, b);                                       \
    m *= m2;                                                  \
  }
  GERAF_AFFINE_STEP(1)
  GERAF_AFFINE_STEP(2)
  GERAF_AFFINE_STEP(4)
  GERAF_AFFINE_STEP(8)
#undef GERAF_AFFINE_STEP
  const float w48 = fmaf(readlane(m, 48), tail, readlane(b, 48));
  const float w32 = fmaf(readlane(m, 32), w48, readlane(b, 32));
  const float w16 = fmaf(readlane(m, 16), w32, readlane(b, 16));
  *head = fmaf(readlane(m, 0), w16, readlane(b, 0));
  const int row = lane >> 4;
  const float after = row == 3 ? tail : row == 2 ? w48 : row == 1 ? w32 : w16;
  const float w = fmaf(m, after, b);
  *w_next = dpp_mov<dpp::kWaveShl1>(tail, w);
  return w;
}

// ------------------------------------------------------------ fast FP32 math
// Hardware approximations (about 1 ulp each) used where the reference formula
// is a plain division, square root or exponential.
__device__ __forceinline__ float rcp_fast(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float sqrt_fast(float x) { return __builtin_amdgcn_sqrtf(x); }
// The logistic function p = 1 / (1 + exp(-x)), its complement 1 - p and its
// slope p (1 - p), each to a few ulp of its own size: with e = exp(-|x|) <= 1
// and r = 1 / (1 + e), the pair {p, 1 - p} is {r, e r} (ordered by the sign
// of x) and the slope is e r^2.  Forming 1 - p or p (1 - p) from a rounded p
// instead loses all relative precision once p rounds to 1 (x > ~17).
struct Logistic {
  float p, c, d;  // p, 1 - p, dp/dx
};
__device__ __forceinline__ Logistic logistic_fast(float x) {
  const float e = __builtin_amdgcn_exp2f(-1.44269504088896341f * fabsf(x));
  const float r = rcp_fast(1.0f + e);
  const float er = e * r;
  return x >= 0.0f ? Logistic{r, er, er * r} : Logistic{er, r, er * r};
}

// Sum over the 64 lanes of a wave (butterfly, every lane gets the total).
__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, kWave);
  return x;
}

// Two-level Horner evaluation of sum_k s_k w^k for wave-uniform coefficient
// rows (read through the scalar cache): 32-term inner Horner in w, outer
// Horner in W = w^32.  Returns the complex sum.
__device__ __forceinline__ float2 horner2(const float* __restrict__ er,
                                          const float* __restrict__ ei, int K, float2 w,
                                          float2 W) {
  constexpr int B = 32;
  const int nb = K / B;
  const int rem = K - nb * B;
  float Sr = 0.0f, Si = 0.0f;
  if (rem) {
    const int k0 = nb * B;
    float ir = er[K - 1], ii = ei[K - 1];
    for (int k = K - 2; k >= k0; --k) {
      const float u1 = fmaf(ir, w.x, er[k]);
      const float u2 = fmaf(ii, w.x, ei[k]);
      const float nr = fmaf(-ii, w.y, u1);
      ii = fmaf(ir, w.y, u2);
      ir = nr;
    }
    Sr = ir;
    Si = ii;
  }
  for (int b = nb - 1; b >= 0; --b) {
    const float* __restrict__ xr = er + b * B;
    const float* __restrict__ xi = ei + b * B;
    float ir = xr[B - 1], ii = xi[B - 1];
#pragma unroll
    for (int j = B - 2; j >= 0; --j) {
      const float u1 = fmaf(ir, w.x, xr[j]);
      const float u2 = fmaf(ii, w.x, xi[j]);
      const float nr = fmaf(-ii, w.y, u1);
      ii = fmaf(ir, w.y, u2);
      ir = nr;
    }
    const float u1 = fmaf(Sr, W.x, ir);
    const float u2 = fmaf(Si, W.x, ii);
    const float nr = fmaf(-Si, W.y, u1);
    Si = fmaf(Sr, W.y, u2);
    Sr = nr;
  }
  return make_float2(Sr, Si);
}

// Phasors of one path for the k-sum: carrier exp(sign*j*2*pi*f0*tau), step
// w = exp(sign*j*2*pi*dfs*tau) and W = w^32, from FP64-reduced cycles.
struct PathPhasors {
  float2 base, w, W;
};
__device__ __forceinline__ PathPhasors path_phasors(double tau, const PhaseConsts& pc,
                                                    float sign) {
  const double cyc1 = pc.dfs * tau;
  PathPhasors r;
  r.base = phasor_cycles((float)centered_frac(pc.f0 * tau), sign);
  r.w = phasor_cycles((float)centered_frac(cyc1), sign);
  r.W = phasor_cycles((float)centered_frac(32.0 * cyc1), sign);
  return r;
}

}  // namespace geraf
