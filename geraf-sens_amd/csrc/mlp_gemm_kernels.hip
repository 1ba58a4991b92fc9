// BF16 matrix-core kernels of the SDF network (sdf_network.py:78-199) on
// gfx950: the layer products of the forward, of the analytic SDF gradient
// (the normals) and of the hand-derived backward, with the element-wise
// work of each step fused into the product's prologue or epilogue, and the
// weight gradients.
//
// Layer products y[M,N] = x[M,K] . W^T (x: points x input features, BF16,
// row-major; W: [N,K]).  N <= 272 and K <= 288 (the network's widths: 256
// hidden, 257 outputs, 39 encoded inputs).  The weights are small and the
// batch of points is tall, so the kernel is weight-stationary: a workgroup
// (8 waves, one per CU) copies the whole packed W into LDS once and then
// streams tiles of 32 points from HBM straight into matrix operands.  On
// v_mfma_f32_16x16x32_bf16 the weights are the A operand (16 output
// features x 32 inputs) and the points the B operand (32 inputs x 16
// points), so a lane's B fragment is 8 consecutive input features of one
// point (one 16-byte load) and its accumulator holds 4 consecutive output
// features of one point (one 8-byte store).  Operand layout (the same as
// the FP16 form used by the matched filter, checked on the GPU): A[i][d] is
// element d % 8 of lane i + 16 (d / 8), B[d][j] element d % 8 of lane
// j + 16 (d / 8), D[4 q + r][j] element r of lane j + 16 q.
//
// Prologues (on x): none, or x * s(x2) with s the softplus slope of a stored
// pre-activation x2 (the gradient pass and its adjoint).  Epilogues (on
// z = acc + bias): z; softplus(z); softplus(z) with z kept (training);
// relu(z); z and z * e1 * beta * curv(e0) (the adjoint of the gradient pass,
// its second-order term); z * s(e0) + e1 (the adjoint of a hidden layer).
// Softplus follows torch.nn.Softplus(beta, threshold): y = z if beta z >
// threshold else log1p(exp(beta z)) / beta, slope s = sigmoid(beta z) (1
// above the threshold), curvature s (1 - s) = e / (1 + e)^2 (0 above).
//
// Weight gradients dW[N,K] = G^T X (G: [M,N], X: [M,K], both tall) with the
// same optional prologue on either side: split over the points, each
// workgroup stages 32-point tiles of G and X transposed in LDS and keeps a
// 256 x 256 block of dW in registers; the partial blocks go to a workspace
// and a second kernel sums them in a fixed order (deterministic).  The bias
// gradient (column sums of G) comes from one more product against ones.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <string>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));

constexpr int kGemmWaves = 8;
constexpr int kMaxNB = 17;   // output feature blocks of 16 (N <= 272)
constexpr int kPts = 32;     // points per wave tile (two blocks of 16)

enum Pro { PRO_NONE = 0, PRO_MUL_SLOPE = 1 };
enum Epi {
  EPI_BIAS = 0,
  EPI_SOFTPLUS = 1,
  EPI_SOFTPLUS_TRAIN = 2,
  EPI_RELU = 3,
  EPI_GRAD_ADJ = 4,
  EPI_LAYER_ADJ = 5
};

__device__ __forceinline__ float bf(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ uint16_t to_bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

struct SpTerms {
  float y, s, c;  // softplus, slope, curvature / beta
};

__device__ __forceinline__ SpTerms sp_terms(float z, float beta, float threshold) {
  const float t = z * beta;
  if (t > threshold) return SpTerms{z, 1.0f, 0.0f};
  const float e = __expf(t);
  const float r = 1.0f / (1.0f + e);
  const float u = 1.0f + e;
  const float l1p = (u == 1.0f) ? e : __logf(u) * (e / (u - 1.0f));
  return SpTerms{l1p / beta, e * r, e * r * r};
}

__device__ __forceinline__ float sp_slope(float z, float beta, float threshold) {
  const float t = z * beta;
  if (t > threshold) return 1.0f;
  const float e = __expf(t);
  return e / (1.0f + e);
}

struct GemmArgs {
  const uint16_t* x;
  const uint16_t* x2;
  const bf8* wpack;  // [nbN][ksK][64]
  const float* bias;
  uint16_t* y;
  uint16_t* y2;
  const uint16_t* e0;
  const uint16_t* e1;
  int64_t ldx, ldx2, ldy, ldy2, lde0, lde1;
  int M, N, K, nbN, ksK;
  float beta, threshold;
};

// The B fragment of points p (this lane's point) for input features
// k0 .. k0 + 7: a 16-byte load, zeros past K or M.
template <int PRO>
__device__ __forceinline__ bf8 load_frag(const GemmArgs& a, int p, int k0) {
  u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
  if (p < a.M && k0 < a.K) {
    v = *reinterpret_cast<const u16x8*>(a.x + (size_t)p * a.ldx + k0);
    if (PRO == PRO_MUL_SLOPE) {
      const u16x8 w = *reinterpret_cast<const u16x8*>(a.x2 + (size_t)p * a.ldx2 + k0);
#pragma unroll
      for (int e = 0; e < 8; ++e) v[e] = to_bf(bf(v[e]) * sp_slope(bf(w[e]), a.beta, a.threshold));
    }
    if (k0 + 8 > a.K) {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (k0 + e >= a.K) v[e] = 0;
    }
  }
  return __builtin_bit_cast(bf8, v);
}

template <int KS, int PRO, int EPI>
__global__ __launch_bounds__(64 * kGemmWaves) void mlp_gemm_kernel(GemmArgs a) {
  extern __shared__ bf8 wl[];
  const int nw = a.nbN * a.ksK * kWave;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) wl[i] = a.wpack[i];
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int tiles = (a.M + kPts - 1) / kPts;
  for (int t = blockIdx.x * kGemmWaves + wave; t < tiles; t += gridDim.x * kGemmWaves) {
    const int p0 = t * kPts;
    bf8 xf[2][KS];
#pragma unroll
    for (int pb = 0; pb < 2; ++pb)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) xf[pb][ks] = load_frag<PRO>(a, p0 + 16 * pb + j, 32 * ks + 8 * g);
    f32x4v acc[2][kMaxNB];
#pragma unroll
    for (int nb = 0; nb < kMaxNB; ++nb) {
      acc[0][nb] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
      acc[1][nb] = acc[0][nb];
    }
#pragma unroll
    for (int nb = 0; nb < kMaxNB; ++nb) {
      if (nb < a.nbN) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const bf8 w = wl[(nb * a.ksK + ks) * kWave + lane];
          acc[0][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xf[0][ks], acc[0][nb], 0, 0, 0);
          acc[1][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xf[1][ks], acc[1][nb], 0, 0, 0);
        }
      }
    }
    // epilogue: lane (j, g) holds features 16 nb + 4 g + r of point p0 + 16 pb + j
#pragma unroll
    for (int pb = 0; pb < 2; ++pb) {
      const int p = p0 + 16 * pb + j;
      if (p >= a.M) continue;
#pragma unroll
      for (int nb = 0; nb < kMaxNB; ++nb) {
        if (nb >= a.nbN) continue;
        const int f0 = 16 * nb + 4 * g;
        if (f0 >= a.N) continue;
        const bool full = f0 + 4 <= a.N;
        float z[4];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          z[r] = acc[pb][nb][r] + ((a.bias && f0 + r < a.N) ? a.bias[f0 + r] : 0.0f);
        u16x4 o = {0, 0, 0, 0}, o2 = {0, 0, 0, 0};
        u16x4 ea = {0, 0, 0, 0}, eb = {0, 0, 0, 0};
        if (EPI == EPI_GRAD_ADJ || EPI == EPI_LAYER_ADJ) {
          const uint16_t* pa = a.e0 + (size_t)p * a.lde0 + f0;
          const uint16_t* pbp = a.e1 + (size_t)p * a.lde1 + f0;
          if (full) {
            ea = *reinterpret_cast<const u16x4*>(pa);
            eb = *reinterpret_cast<const u16x4*>(pbp);
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (f0 + r < a.N) {
                ea[r] = pa[r];
                eb[r] = pbp[r];
              }
          }
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          if (EPI == EPI_BIAS) {
            o[r] = to_bf(z[r]);
          } else if (EPI == EPI_SOFTPLUS) {
            o[r] = to_bf(sp_terms(bf(to_bf(z[r])), a.beta, a.threshold).y);
          } else if (EPI == EPI_SOFTPLUS_TRAIN) {
            o2[r] = to_bf(z[r]);
            // the activation of the stored (rounded) pre-activation, as the
            // backward sees it
            o[r] = to_bf(sp_terms(bf(o2[r]), a.beta, a.threshold).y);
          } else if (EPI == EPI_RELU) {
            o[r] = to_bf(fmaxf(z[r], 0.0f));
          } else if (EPI == EPI_GRAD_ADJ) {
            // y = vbar (the product), y2 = vbar * u * beta * curv(z_l)
            o[r] = to_bf(z[r]);
            const SpTerms st = sp_terms(bf(ea[r]), a.beta, a.threshold);
            o2[r] = to_bf(z[r] * bf(eb[r]) * a.beta * st.c);
          } else {  // EPI_LAYER_ADJ: zbar = hbar * s(z) + t
            const float s = sp_slope(bf(ea[r]), a.beta, a.threshold);
            o[r] = to_bf(fmaf(z[r], s, bf(eb[r])));
          }
        }
        uint16_t* py = a.y + (size_t)p * a.ldy + f0;
        if (full) {
          *reinterpret_cast<u16x4*>(py) = o;
        } else {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (f0 + r < a.N) py[r] = o[r];
        }
        if (EPI == EPI_SOFTPLUS_TRAIN || EPI == EPI_GRAD_ADJ) {
          uint16_t* py2 = a.y2 + (size_t)p * a.ldy2 + f0;
          if (full) {
            *reinterpret_cast<u16x4*>(py2) = o2;
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (f0 + r < a.N) py2[r] = o2[r];
          }
        }
      }
    }
  }
}

// W [N,K] (FP32, w[n * ldw + k], or w[k * ldw + n] with transpose) times
// `scale` -> the A operands in LDS order: [nb][ks][lane] of 8 BF16, zero
// padded to 16 nb x 32 ks.
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ w, int64_t ldw,
                                                       int transpose, float scale, int N, int K,
                                                       int ksK, u16x8* __restrict__ out,
                                                       int total) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int lane = idx & 63;
  const int ks = (idx >> 6) % ksK;
  const int nb = (idx >> 6) / ksK;
  const int n = 16 * nb + (lane & 15);
  const int k0 = 32 * ks + 8 * (lane >> 4);
  u16x8 v;
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const int k = k0 + e;
    float x = 0.0f;
    if (n < N && k < K) x = transpose ? w[(size_t)k * ldw + n] : w[(size_t)n * ldw + k];
    v[e] = to_bf(x * scale);
  }
  out[idx] = v;
}

// ------------------------------------------------------------------ weight gradients

constexpr int kWgWaves = 8;
constexpr int kWgTile = 256;        // dW block per workgroup: 256 x 256
constexpr int kWgStride = 40;       // LDS row stride of the transposed tiles (32 points + pad)

struct WgradArgs {
  const uint16_t* g;
  const uint16_t* g2;
  const uint16_t* x;
  const uint16_t* x2;
  float* partial;  // [gridDim.x][N_t][K_t + 1] for this (n, k) super-tile
  int64_t ldg, ldg2, ldx, ldx2;
  int M, N, K;
  int with_bias;
  int pts_per_block;
  float beta, threshold;
};

// A 32-point tile of a [M, C] tensor (columns c0 .. c0 + 255), transposed into
// LDS: t[c][p] (row stride kWgStride), with the optional slope prologue.
template <int PRO>
__device__ __forceinline__ void stage_tile(const uint16_t* src, int64_t ld, const uint16_t* src2,
                                           int64_t ld2, int M, int C, int c0, int p0,
                                           uint16_t* t, float beta, float threshold) {
  // 32 points x 32 column groups of 8 = 1024 loads of 16 bytes, two per thread
  for (int q = threadIdx.x; q < kPts * (kWgTile / 8); q += blockDim.x) {
    const int p = q / (kWgTile / 8);
    const int cg = q % (kWgTile / 8);
    const int c = c0 + 8 * cg;
    u16x8 v = {0, 0, 0, 0, 0, 0, 0, 0};
    if (p0 + p < M && c < C) {
      v = *reinterpret_cast<const u16x8*>(src + (size_t)(p0 + p) * ld + c);
      if (PRO == PRO_MUL_SLOPE) {
        const u16x8 w = *reinterpret_cast<const u16x8*>(src2 + (size_t)(p0 + p) * ld2 + c);
#pragma unroll
        for (int e = 0; e < 8; ++e) v[e] = to_bf(bf(v[e]) * sp_slope(bf(w[e]), beta, threshold));
      }
      if (c + 8 > C) {
#pragma unroll
        for (int e = 0; e < 8; ++e)
          if (c + e >= C) v[e] = 0;
      }
    }
#pragma unroll
    for (int e = 0; e < 8; ++e) t[(8 * cg + e) * kWgStride + p] = v[e];
  }
}

// Workgroup: 8 waves in 4 (n) x 2 (k); wave tile 64 (n) x 128 (k) of a
// 256 x 256 block of dW (the host offsets g and x to the block's columns).
template <int PRO_G, int PRO_X>
__global__ __launch_bounds__(64 * kWgWaves) void mlp_wgrad_kernel(WgradArgs a) {
  __shared__ uint16_t gt[kWgTile * kWgStride];
  __shared__ uint16_t xt[kWgTile * kWgStride];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int wn = wave & 3, wk = wave >> 2;
  const int pbeg = blockIdx.x * a.pts_per_block;
  const int pend = min(a.M, pbeg + a.pts_per_block);
  const bool bias_col = a.with_bias && wk == 0;
  f32x4v acc[4][8];
  f32x4v bacc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    bacc[i] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) acc[i][kk] = bacc[i];
  }
  u16x8 ones_u;
#pragma unroll
  for (int e = 0; e < 8; ++e) ones_u[e] = 0x3f80;  // 1.0 in BF16
  const bf8 ones = __builtin_bit_cast(bf8, ones_u);
  for (int p0 = pbeg; p0 < pend; p0 += kPts) {
    __syncthreads();
    stage_tile<PRO_G>(a.g, a.ldg, a.g2, a.ldg2, pend, a.N, 0, p0, gt, a.beta, a.threshold);
    stage_tile<PRO_X>(a.x, a.ldx, a.x2, a.ldx2, pend, a.K, 0, p0, xt, a.beta, a.threshold);
    __syncthreads();
    // A = G^T rows (n), depth = the 32 points; B = X (points x k)
    bf8 af[4], bfr[8];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      af[i] = *reinterpret_cast<const bf8*>(gt + (64 * wn + 16 * i + j) * kWgStride + 8 * g);
#pragma unroll
    for (int kk = 0; kk < 8; ++kk)
      bfr[kk] = *reinterpret_cast<const bf8*>(xt + (128 * wk + 16 * kk + j) * kWgStride + 8 * g);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int kk = 0; kk < 8; ++kk)
        acc[i][kk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[kk], acc[i][kk], 0, 0, 0);
      if (bias_col)
        bacc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], ones, bacc[i], 0, 0, 0);
    }
  }
  // partial block: [N_t = 256][K_t + 1 = 257] (the last column: the bias gradient)
  float* out = a.partial + (size_t)blockIdx.x * kWgTile * (kWgTile + 1);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 64 * wn + 16 * i + 4 * g + r;
#pragma unroll
      for (int kk = 0; kk < 8; ++kk) out[(size_t)n * (kWgTile + 1) + 128 * wk + 16 * kk + j] =
          acc[i][kk][r];
      if (bias_col && j == 0) out[(size_t)n * (kWgTile + 1) + kWgTile] = bacc[i][r];
    }
  }
}

// dW[n][k] (+)= sum over the partial blocks in order; db[n] likewise.
__global__ __launch_bounds__(256) void mlp_wgrad_reduce_kernel(const float* __restrict__ partial,
                                                               int blocks, int n0, int k0, int N,
                                                               int K, float* __restrict__ dw,
                                                               int64_t lddw, float* __restrict__ db,
                                                               int accumulate) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int cols = kWgTile + 1;
  if (idx >= kWgTile * cols) return;
  const int n = idx / cols, c = idx % cols;
  const bool is_bias = c == kWgTile;
  if (n0 + n >= N) return;
  if (is_bias ? (db == nullptr || k0 != 0) : (k0 + c >= K)) return;
  double s = 0.0;
  for (int b = 0; b < blocks; ++b) s += (double)partial[(size_t)b * kWgTile * cols + idx];
  if (is_bias) {
    db[n0 + n] = (float)(accumulate ? (double)db[n0 + n] + s : s);
  } else {
    float* o = dw + (size_t)(n0 + n) * lddw + k0 + c;
    *o = (float)(accumulate ? (double)*o + s : s);
  }
}

// ------------------------------------------------------------------ positional encoding

// gamma(x) = [x, sin(2^l x), cos(2^l x) for l < L] (x scaled), BF16, into
// out[p * ldo + c0 + i] for i < 3 + 6 L (the band order of sdf_network.py's
// Embedder: per band the 3 sines, then the 3 cosines); columns up to
// pad_to are zero-filled.
__global__ __launch_bounds__(256) void pe_kernel(const float* __restrict__ x, int M, int L,
                                                 float scale, uint16_t* __restrict__ out,
                                                 int64_t ldo, int c0, int pad_to) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  float v[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) v[d] = x[3 * (size_t)p + d] * scale;
  uint16_t* o = out + (size_t)p * ldo + c0;
  for (int d = 0; d < 3; ++d) o[d] = to_bf(v[d]);
  for (int l = 0; l < L; ++l) {
    const float f = (float)(1 << l);
    for (int d = 0; d < 3; ++d) {
      const float t = v[d] * f;
      o[3 + 6 * l + d] = to_bf(sinf(t));
      o[3 + 6 * l + 3 + d] = to_bf(cosf(t));
    }
  }
  for (int c = 3 + 6 * L; c < pad_to; ++c) o[c] = 0;
}

// n = d sdf / d x from d sdf / d gamma (the sum of two BF16 parts, [M, 3+6L]
// each at their own stride): n_d = scale * (g[d] + sum_l 2^l (cos(2^l x_d)
// g[3+6l+d] - sin(2^l x_d) g[3+6l+3+d])), FP32.
__global__ __launch_bounds__(256) void pe_grad_kernel(const float* __restrict__ x, int M, int L,
                                                      float scale, const uint16_t* __restrict__ ga,
                                                      int64_t lda, const uint16_t* __restrict__ gb,
                                                      int64_t ldb, float* __restrict__ n) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  const uint16_t* a = ga + (size_t)p * lda;
  const uint16_t* b = gb + (size_t)p * ldb;
  for (int d = 0; d < 3; ++d) {
    const float v = x[3 * (size_t)p + d] * scale;
    float s = bf(a[d]) + bf(b[d]);
    for (int l = 0; l < L; ++l) {
      const float f = (float)(1 << l);
      const float t = v * f;
      const float gs = bf(a[3 + 6 * l + d]) + bf(b[3 + 6 * l + d]);
      const float gc = bf(a[3 + 6 * l + 3 + d]) + bf(b[3 + 6 * l + 3 + d]);
      s += f * (cosf(t) * gs - sinf(t) * gc);
    }
    n[3 * (size_t)p + d] = s * scale;
  }
}

// The adjoint of pe_grad for d loss / d n: d loss / d gamma = J dn with J the
// encoding's Jacobian (scaled), BF16, into out[p * ldo + c0 + i], zero-padded
// up to pad_to columns.
__global__ __launch_bounds__(256) void pe_grad_adj_kernel(const float* __restrict__ x, int M,
                                                          int L, float scale,
                                                          const float* __restrict__ dn,
                                                          uint16_t* __restrict__ out,
                                                          int64_t ldo, int c0, int pad_to) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= M) return;
  uint16_t* o = out + (size_t)p * ldo + c0;
  for (int d = 0; d < 3; ++d) {
    const float v = x[3 * (size_t)p + d] * scale;
    const float gd = dn[3 * (size_t)p + d] * scale;
    o[d] = to_bf(gd);
    for (int l = 0; l < L; ++l) {
      const float f = (float)(1 << l);
      const float t = v * f;
      o[3 + 6 * l + d] = to_bf(f * cosf(t) * gd);
      o[3 + 6 * l + 3 + d] = to_bf(-f * sinf(t) * gd);
    }
  }
  for (int c = 3 + 6 * L; c < pad_to; ++c) o[c] = 0;
}

// ld == 0 broadcasts one row to every point; otherwise rows are 16-byte
// aligned and hold the columns rounded up to 8 (the fragments' width).
void check_ld(int64_t ld, int64_t cols, const char* what) {
  if (ld == 0) return;
  if (ld < 0 || ld % 8 != 0 || ld < ((cols + 7) / 8) * 8)
    throw ArgError(std::string(what) + ": leading dimension must be 0 (broadcast) or a "
                   "multiple of 8 and >= the columns rounded up to 8");
}

// Dynamic LDS above the default launch limit needs the kernel attribute.
template <typename F>
void allow_lds(F* kernel, size_t bytes) {
  if (bytes > 65536) {
    const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)bytes);
    if (e != hipSuccess) throw HipError(std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  }
}

int ks_of(int64_t K) { return (int)((K + 31) / 32); }
int nb_of(int64_t N) { return (int)((N + 15) / 16); }

int gemm_grid(int64_t M) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  const int64_t tiles = (M + kPts - 1) / kPts;
  const int64_t want = (tiles + kGemmWaves - 1) / kGemmWaves;
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, cus));
}

template <int KS, int PRO>
int launch_gemm_epi(int epi, const GemmArgs& a, size_t lds, hipStream_t st) {
  const dim3 grid(gemm_grid(a.M)), block(64 * kGemmWaves);
  switch (epi) {
#define GERAF_EPI(E)                                                               \
  case E:                                                                          \
    allow_lds(mlp_gemm_kernel<KS, PRO, E>, lds);                                   \
    hipLaunchKernelGGL((mlp_gemm_kernel<KS, PRO, E>), grid, block, lds, st, a);    \
    break;
    GERAF_EPI(EPI_BIAS)
    GERAF_EPI(EPI_SOFTPLUS)
    GERAF_EPI(EPI_SOFTPLUS_TRAIN)
    GERAF_EPI(EPI_RELU)
    GERAF_EPI(EPI_GRAD_ADJ)
    GERAF_EPI(EPI_LAYER_ADJ)
#undef GERAF_EPI
    default:
      throw ArgError("unknown epilogue");
  }
  return check_launch("mlp_gemm_kernel");
}

template <int PRO>
int launch_gemm(int epi, const GemmArgs& a, size_t lds, hipStream_t st) {
  if (a.ksK <= 2) return launch_gemm_epi<2, PRO>(epi, a, lds, st);
  if (a.ksK <= 8) return launch_gemm_epi<8, PRO>(epi, a, lds, st);
  return launch_gemm_epi<9, PRO>(epi, a, lds, st);
}

}  // namespace
}  // namespace geraf

using namespace geraf;

extern "C" int64_t geraf_mlp_pack_bytes(int64_t N, int64_t K) {
  if (N <= 0 || K <= 0 || N > 16 * kMaxNB || K > 32 * 9) return -1;
  return (int64_t)nb_of(N) * ks_of(K) * kWave * 16;
}

extern "C" int geraf_mlp_pack_weights(const float* w, int64_t ldw, int32_t transpose, float scale,
                                      int64_t N, int64_t K, void* packed, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (geraf_mlp_pack_bytes(N, K) < 0) throw ArgError("N must be in [1, 272] and K in [1, 288]");
  require_ptrs({w, packed}, true);
  if ((reinterpret_cast<uintptr_t>(packed) & 15u) != 0) throw ArgError("packed must be 16-byte aligned");
  const int total = nb_of(N) * ks_of(K) * kWave;
  hipLaunchKernelGGL(mlp_pack_kernel, dim3((total + 255) / 256), dim3(256), 0,
                     (hipStream_t)stream, w, ldw, (int)transpose, scale, (int)N, (int)K,
                     ks_of(K), reinterpret_cast<u16x8*>(packed), total);
  return check_launch("mlp_pack_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mlp_gemm_bf16(const uint16_t* x, int64_t ldx, const uint16_t* x2,
                                   int64_t ldx2, const void* wpacked, const float* bias,
                                   uint16_t* y, int64_t ldy, uint16_t* y2, int64_t ldy2,
                                   const uint16_t* e0, int64_t lde0, const uint16_t* e1,
                                   int64_t lde1, int64_t M, int64_t N, int64_t K,
                                   int32_t prologue, int32_t epilogue, float beta,
                                   float threshold, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX) throw ArgError("M out of range");
  if (geraf_mlp_pack_bytes(N, K) < 0) throw ArgError("N must be in [1, 272] and K in [1, 288]");
  if (M == 0) return GERAF_RF_OK;
  if (prologue != PRO_NONE && prologue != PRO_MUL_SLOPE) throw ArgError("unknown prologue");
  require_ptrs({x, wpacked, y}, true);
  check_ld(ldx, K, "x");
  check_ld(ldy, N, "y");
  if (ldy % 4 != 0) throw ArgError("y: leading dimension must be a multiple of 4");
  if (prologue == PRO_MUL_SLOPE) {
    require_ptrs({x2}, true);
    check_ld(ldx2, K, "x2");
  }
  if (epilogue == EPI_SOFTPLUS_TRAIN || epilogue == EPI_GRAD_ADJ) {
    require_ptrs({y2}, true);
    check_ld(ldy2, N, "y2");
    if (ldy2 % 4 != 0) throw ArgError("y2: leading dimension must be a multiple of 4");
  }
  if (epilogue == EPI_GRAD_ADJ || epilogue == EPI_LAYER_ADJ) {
    require_ptrs({e0, e1}, true);
    check_ld(lde0, N, "e0");
    check_ld(lde1, N, "e1");
    if (lde0 % 4 != 0 || lde1 % 4 != 0)
      throw ArgError("e0/e1: leading dimensions must be multiples of 4");
  }
  const uintptr_t align = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(x2) |
                          reinterpret_cast<uintptr_t>(wpacked);
  const uintptr_t align8 = reinterpret_cast<uintptr_t>(y) | reinterpret_cast<uintptr_t>(y2) |
                           reinterpret_cast<uintptr_t>(e0) | reinterpret_cast<uintptr_t>(e1);
  if ((align & 15u) != 0 || (align8 & 7u) != 0)
    throw ArgError("x, x2, wpacked must be 16-byte and y, y2, e0, e1 8-byte aligned");
  GemmArgs a;
  a.x = x;
  a.x2 = x2;
  a.wpack = reinterpret_cast<const bf8*>(wpacked);
  a.bias = bias;
  a.y = y;
  a.y2 = y2;
  a.e0 = e0;
  a.e1 = e1;
  a.ldx = ldx;
  a.ldx2 = ldx2;
  a.ldy = ldy;
  a.ldy2 = ldy2;
  a.lde0 = lde0;
  a.lde1 = lde1;
  a.M = (int)M;
  a.N = (int)N;
  a.K = (int)K;
  a.nbN = nb_of(N);
  a.ksK = ks_of(K);
  a.beta = beta;
  a.threshold = threshold;
  const size_t lds = (size_t)a.nbN * a.ksK * kWave * sizeof(bf8);
  const hipStream_t st = (hipStream_t)stream;
  if (prologue == PRO_MUL_SLOPE) return launch_gemm<PRO_MUL_SLOPE>(epilogue, a, lds, st);
  return launch_gemm<PRO_NONE>(epilogue, a, lds, st);
  GERAF_TRY_END
}

extern "C" int64_t geraf_mlp_wgrad_workspace_bytes(int64_t M, int64_t N, int64_t K) {
  if (M < 0 || N <= 0 || K <= 0) return -1;
  (void)M;
  (void)N;
  (void)K;
  return (int64_t)256 * kWgTile * (kWgTile + 1) * (int64_t)sizeof(float);
}

extern "C" int geraf_mlp_wgrad_bf16(const uint16_t* g, int64_t ldg, const uint16_t* g2,
                                    int64_t ldg2, int32_t prologue_g, const uint16_t* x,
                                    int64_t ldx, const uint16_t* x2, int64_t ldx2,
                                    int32_t prologue_x, float* dw, int64_t lddw, float* db,
                                    int32_t accumulate, int64_t M, int64_t N, int64_t K,
                                    float beta, float threshold, void* workspace,
                                    int64_t workspace_bytes, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX || N <= 0 || K <= 0 || N > 512 || K > 512)
    throw ArgError("sizes out of range (N, K in [1, 512])");
  require_ptrs({dw, workspace}, true);
  if (workspace_bytes < geraf_mlp_wgrad_workspace_bytes(M, N, K))
    throw ArgError("workspace too small (see geraf_mlp_wgrad_workspace_bytes)");
  if (lddw < K) throw ArgError("dw: leading dimension must be >= K");
  const hipStream_t st = (hipStream_t)stream;
  if (M > 0) {
    require_ptrs({g, x}, true);
    check_ld(ldg, N, "g");
    check_ld(ldx, K, "x");
    if (prologue_g == PRO_MUL_SLOPE) {
      require_ptrs({g2}, true);
      check_ld(ldg2, N, "g2");
    }
    if (prologue_x == PRO_MUL_SLOPE) {
      require_ptrs({x2}, true);
      check_ld(ldx2, K, "x2");
    }
  }
  const uintptr_t align = reinterpret_cast<uintptr_t>(g) | reinterpret_cast<uintptr_t>(g2) |
                          reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(x2);
  if ((align & 15u) != 0) throw ArgError("g, g2, x, x2 must be 16-byte aligned");
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  const int nsup = (int)((N + kWgTile - 1) / kWgTile), ksup = (int)((K + kWgTile - 1) / kWgTile);
  // point chunks: a multiple of 32 points each, at most 256 chunks (the workspace)
  int64_t per = (M + cus - 1) / std::max(1, cus);
  per = std::max<int64_t>(kPts, ((per + kPts - 1) / kPts) * kPts);
  const int blocks = M == 0 ? 1 : (int)std::min<int64_t>(256, (M + per - 1) / per);
  if ((int64_t)blocks * per < M) per = ((M + blocks - 1) / blocks + kPts - 1) / kPts * kPts;
  float* partial = reinterpret_cast<float*>(workspace);
  for (int ns = 0; ns < nsup; ++ns) {
    for (int kz = 0; kz < ksup; ++kz) {
      WgradArgs a;
      a.g = g;
      a.g2 = g2;
      a.x = x;
      a.x2 = x2;
      a.partial = partial;
      a.ldg = ldg;
      a.ldg2 = ldg2;
      a.ldx = ldx;
      a.ldx2 = ldx2;
      a.M = (int)M;
      a.N = (int)N;
      a.K = (int)K;
      a.with_bias = db != nullptr;
      a.pts_per_block = (int)per;
      a.beta = beta;
      a.threshold = threshold;
      // one super-tile at a time: blockIdx.y / z select it
      const dim3 grid((unsigned)blocks, 1, 1);
      WgradArgs b = a;
      // shift the super-tile origin into the pointers' column offsets
      b.g = g ? g + (size_t)ns * kWgTile : nullptr;
      b.g2 = g2 ? g2 + (size_t)ns * kWgTile : nullptr;
      b.x = x ? x + (size_t)kz * kWgTile : nullptr;
      b.x2 = x2 ? x2 + (size_t)kz * kWgTile : nullptr;
      b.N = (int)std::min<int64_t>(kWgTile, N - (int64_t)ns * kWgTile);
      b.K = (int)std::min<int64_t>(kWgTile, K - (int64_t)kz * kWgTile);
      b.with_bias = (db != nullptr && kz == 0);
      if (M > 0) {
#define GERAF_WG(PG, PX)                                                                     \
  hipLaunchKernelGGL((mlp_wgrad_kernel<PG, PX>), grid, dim3(64 * kWgWaves), 0, st, b)
        if (prologue_g == PRO_MUL_SLOPE && prologue_x == PRO_MUL_SLOPE)
          GERAF_WG(PRO_MUL_SLOPE, PRO_MUL_SLOPE);
        else if (prologue_g == PRO_MUL_SLOPE)
          GERAF_WG(PRO_MUL_SLOPE, PRO_NONE);
        else if (prologue_x == PRO_MUL_SLOPE)
          GERAF_WG(PRO_NONE, PRO_MUL_SLOPE);
        else
          GERAF_WG(PRO_NONE, PRO_NONE);
#undef GERAF_WG
        const int rc = check_launch("mlp_wgrad_kernel");
        if (rc != GERAF_RF_OK) return rc;
      }
      const int cells = kWgTile * (kWgTile + 1);
      hipLaunchKernelGGL(mlp_wgrad_reduce_kernel, dim3((cells + 255) / 256), dim3(256), 0, st,
                         partial, M == 0 ? 0 : blocks, ns * kWgTile, kz * kWgTile, (int)N,
                         (int)K, dw, lddw, (kz == 0) ? db : nullptr, (int)accumulate);
      const int rc = check_launch("mlp_wgrad_reduce_kernel");
      if (rc != GERAF_RF_OK) return rc;
    }
  }
  return GERAF_RF_OK;
  GERAF_TRY_END
}

extern "C" int geraf_mlp_positional_encoding(const float* x, int64_t M, int32_t L, float scale,
                                             uint16_t* out, int64_t ldo, int32_t col0,
                                             int32_t pad_to, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX || L < 0 || L > 16) throw ArgError("sizes out of range");
  if (col0 < 0 || pad_to < 3 + 6 * L || col0 + pad_to > ldo)
    throw ArgError("columns out of range");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, out}, true);
  hipLaunchKernelGGL(pe_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, x, (int)M, (int)L, scale, out, ldo, (int)col0,
                     (int)pad_to);
  return check_launch("pe_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mlp_pe_gradient(const float* x, int64_t M, int32_t L, float scale,
                                     const uint16_t* ga, int64_t lda, const uint16_t* gb,
                                     int64_t ldb, float* n, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX || L < 0 || L > 16) throw ArgError("sizes out of range");
  if (lda < 3 + 6 * L || ldb < 3 + 6 * L) throw ArgError("leading dimensions too small");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, ga, gb, n}, true);
  hipLaunchKernelGGL(pe_grad_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, x, (int)M, (int)L, scale, ga, lda, gb, ldb, n);
  return check_launch("pe_grad_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mlp_pe_gradient_adjoint(const float* x, int64_t M, int32_t L, float scale,
                                             const float* dn, uint16_t* out, int64_t ldo,
                                             int32_t col0, int32_t pad_to,
                                             geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX || L < 0 || L > 16) throw ArgError("sizes out of range");
  if (col0 < 0 || pad_to < 3 + 6 * L || col0 + pad_to > ldo)
    throw ArgError("columns out of range");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, dn, out}, true);
  hipLaunchKernelGGL(pe_grad_adj_kernel, dim3((unsigned)((M + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, x, (int)M, (int)L, scale, dn, out, ldo, (int)col0,
                     (int)pad_to);
  return check_launch("pe_grad_adj_kernel");
  GERAF_TRY_END
}
