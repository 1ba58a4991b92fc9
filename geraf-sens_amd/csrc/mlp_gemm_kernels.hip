// BF16 matrix-core kernels of the SDF network (sdf_network.py:78-199) on
// gfx950: the layer products of the forward, of the analytic SDF gradient
// (the normals) and of the hand-derived backward, with the element-wise
// work of each step fused into the product's prologue or epilogue, and the
// weight gradients.
//
// Layer products y[M,N] = x[M,K] . W^T (x: points x input features, BF16,
// row-major; W: [N,K]); the outputs are written in groups of 8 columns, so
// columns N .. 8 ceil(N / 8) - 1 of y (and y2) receive unspecified values.  N <= 272 and K <= 288 (the network's widths: 256
// hidden, 257 outputs, 39 encoded inputs).  The weights are small and the
// batch of points is tall, so the kernel is weight-stationary: a workgroup
// (8 waves, one per CU) copies the whole packed W into LDS once and then
// streams tiles of 16 or 32 points from HBM straight into matrix operands,
// the next tile's loads in flight while the current one is multiplied.  On
// v_mfma_f32_16x16x32_bf16 the weights are the A operand (16 output
// features x 32 inputs) and the points the B operand (32 inputs x 16
// points), so a lane's B fragment is 8 consecutive input features of one
// point (one 16-byte load).  The pack permutes W's rows so that the
// accumulators of output blocks 2m and 2m + 1 hold, in lane (j, q), the 8
// consecutive features 32 m + 8 q .. + 7 of point j (one 16-byte store; a
// wave's store writes 64 contiguous bytes of each of 16 rows).  Operand layout (the same as
// the FP16 form used by the matched filter, checked on the GPU): A[i][d] is
// element d % 8 of lane i + 16 (d / 8), B[d][j] element d % 8 of lane
// j + 16 (d / 8), D[4 q + r][j] element r of lane j + 16 q.
//
// Softplus slope codes.  Every use of a stored pre-activation z in the
// network's backward is through the softplus slope s = sigmoid(beta z) or
// its curvature s (1 - s), so the forward stores instead the code
// q = copysign(m, t), m = sigmoid(-|t|) in (0, 1/2], t = beta z, computed
// from the FP32 accumulator and kept as FP16 (11 significant bits; m's
// range stays clear of FP16's, subnormals included, wherever s or s (1 - s)
// is not negligible).  Decoding is exact arithmetic on q: s = 1 - m for
// q >= +0, s = m for q <= -0 (the sign bit; -0 is t -> -inf), s (1 - s) =
// m (1 - m).  Above torch's threshold (t > 20) m < 2.1e-9 rounds to +0:
// slope 1, curvature 0, as torch.nn.Softplus's linear branch has.
//
// Prologues (on x): none, or x * s(q) (the gradient pass and its adjoint).
// Epilogues (on z = acc + bias): z; softplus(z); softplus(z) with the code
// of z (training); relu(z); z and z * e1 * beta * s(1-s) (the adjoint of
// the gradient pass, its second-order term); z * s(e0) + e1 (the adjoint of
// a hidden layer); z * s(e0) (the gradient pass, storing u s); z * s(e0)
// and z * e1 * beta * (1 - s(e0)) with e1 = u s, or z * e1 * beta *
// s(1 - s) with e1 = u (its adjoint, storing vbar s, so that the weight
// gradients and the next product need no slope prologue).  Softplus
// follows torch.nn.Softplus(beta, threshold): y = z if beta z > threshold
// else log1p(exp(beta z)) / beta, evaluated as (max(t, 0) +
// log1p(exp(-|t|))) / beta with the hardware exp2 / log2 / reciprocal
// (log1p by its series below 2^-6).
//
// Weight gradients dW[N,K] = G^T X (G: [M,N], X: [M,K], both tall) with the
// same optional prologue on either side: split over the points, each
// workgroup stages steps of 32 or 64 points of G and X transposed in LDS
// (4-point register transposes, 8-byte LDS stores), keeps a 256 x 256 block
// of dW (plus one 16-row block for N up to 272) in registers and has the
// next step's loads in flight while it multiplies; the partial blocks go to
// a workspace and a second kernel sums them in a fixed order
// (deterministic).  The bias gradient (column sums of G) comes from one
// more product against ones.
#include <hip/hip_runtime.h>

#include <stdint.h>

#include <algorithm>
#include <string>
#include <type_traits>

#include "abi_util.hpp"
#include "rf_common.hpp"

namespace geraf {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef uint16_t u16x4 __attribute__((ext_vector_type(4)));

constexpr int kGemmWaves = 8;
constexpr int kMaxN = 272;   // output features
constexpr int kMaxK = 288;   // input features
constexpr int kNbChunk = 8;  // output blocks of 16 per accumulator pass

enum Pro { PRO_NONE = 0, PRO_MUL_SLOPE = 1 };
enum Epi {
  EPI_BIAS = 0,
  EPI_SOFTPLUS = 1,
  EPI_SOFTPLUS_TRAIN = 2,
  EPI_RELU = 3,
  EPI_GRAD_ADJ = 4,
  EPI_LAYER_ADJ = 5,
  EPI_MUL_SLOPE = 6,
  EPI_GRAD_ADJ_S = 7,
  EPI_GRAD_ADJ_SU = 8
};

// ------------------------------------------------------------------ element-wise math

__device__ __forceinline__ float bf(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (v_cvt_pk_bf16_f32; NaN stays NaN)
__device__ __forceinline__ uint16_t to_bf(float f) {
  return __builtin_bit_cast(uint16_t, (__bf16)f);
}

__device__ __forceinline__ float hf(uint16_t h) {
  return (float)__builtin_bit_cast(_Float16, h);
}

__device__ __forceinline__ uint16_t to_hf(float f) {
  return __builtin_bit_cast(uint16_t, (_Float16)f);
}

constexpr float kLog2e = 1.44269504088896341f;
constexpr float kLn2 = 0.693147180559945309f;

// s from the code q: 1/2 + copysign(1/2 - |q|, q)
__device__ __forceinline__ float code_slope(uint16_t q) {
  const float f = hf(q);
  return 0.5f + copysignf(0.5f - fabsf(f), f);
}

// 1 - s from the code q: m for q >= +0, 1 - m for q <= -0
__device__ __forceinline__ float code_one_minus_slope(uint16_t q) {
  const float f = hf(q);
  return 0.5f - copysignf(0.5f - fabsf(f), f);
}

// s (1 - s) = m (1 - m) from the code q
__device__ __forceinline__ float code_curv(uint16_t q) {
  const float m = fabsf(hf(q));
  return fmaf(-m, m, m);
}

struct Softplus {
  float y, q;
};

// Above torch's threshold (t > 20) the log1p term is below 2.1e-9, so the
// linear branch (y = z) and this form agree to within an FP32 ulp of z; the
// select is left out (it would cost a branch per element).
template <bool CODE>
__device__ __forceinline__ Softplus softplus_of(float z, float beta, float inv_beta) {
  const float t = z * beta;
  const float e = __builtin_amdgcn_exp2f(-fabsf(t) * kLog2e);  // exp(-|t|) in (0, 1]
  const float u = 1.0f + e;
  // log1p(e): e itself below 2^-12 (relative error < 2^-13), else log(1 + e)
  // (the rounding of 1 + e costs at most 2^-24 / e relative)
  const float l = e < 0.000244140625f ? e : __builtin_amdgcn_logf(u) * kLn2;
  Softplus r;
  r.y = (fmaxf(t, 0.0f) + l) * inv_beta;
  r.q = CODE ? copysignf(e * __builtin_amdgcn_rcpf(u), t) : 0.0f;
  return r;
}

// ------------------------------------------------------------------ layer products

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
  float beta, inv_beta;
};

__device__ __forceinline__ u16x8 zero8() { return u16x8{0, 0, 0, 0, 0, 0, 0, 0}; }

// Raw inputs of the B fragments of one point (its row pointers, already at
// this lane's 8-feature group), step ks of 32 input features: 16-byte loads
// of x (and of the code for the prologue).  Steps below ksK - 1 lie inside
// K; the last step loads only where it starts inside K (rows hold K rounded
// up to 8) and zeros its features past K; steps past ksK are zero (their W
// fragments are zero in LDS).
template <int PRO>
__device__ __forceinline__ void load_step(const GemmArgs& a, const uint16_t* xp,
                                          const uint16_t* qp, int ks, int g, u16x8& v, u16x8& q) {
  v = zero8();
  if (PRO == PRO_MUL_SLOPE) q = zero8();
  if (ks < a.ksK - 1) {
    v = *reinterpret_cast<const u16x8*>(xp + 32 * ks);
    if (PRO == PRO_MUL_SLOPE) q = *reinterpret_cast<const u16x8*>(qp + 32 * ks);
  } else if (ks == a.ksK - 1) {
    if (32 * ks + 8 * g < a.K) {
      v = *reinterpret_cast<const u16x8*>(xp + 32 * ks);
      if (PRO == PRO_MUL_SLOPE) q = *reinterpret_cast<const u16x8*>(qp + 32 * ks);
    }
  }
}

template <int PRO>
__device__ __forceinline__ bf8 make_frag(const GemmArgs& a, int ks, int g, u16x8 v, u16x8 q) {
  if (PRO == PRO_MUL_SLOPE) {
#pragma unroll
    for (int e = 0; e < 8; ++e) v[e] = to_bf(bf(v[e]) * code_slope(q[e]));
  }
  if (ks == a.ksK - 1 && 32 * a.ksK != a.K) {
    const int k0 = 32 * ks + 8 * g;
#pragma unroll
    for (int e = 0; e < 8; ++e)
      if (k0 + e >= a.K) v[e] = 0;
  }
  return __builtin_bit_cast(bf8, v);
}

// epilogues that read e0 (and e1) per output element
constexpr bool adj_epi(int epi) { return epi >= EPI_GRAD_ADJ; }
// epilogues that write y2
constexpr bool two_out_epi(int epi) {
  return epi == EPI_SOFTPLUS_TRAIN || epi == EPI_GRAD_ADJ || epi == EPI_GRAD_ADJ_S ||
         epi == EPI_GRAD_ADJ_SU;
}
// epilogues that read e1
constexpr bool e1_epi(int epi) { return adj_epi(epi) && epi != EPI_MUL_SLOPE; }

// 8 consecutive 16-bit values at column f0 of row p (f0 a multiple of 8 and
// below N; rows hold N rounded up to 8, so the access stays in the row)
__device__ __forceinline__ u16x8 load8(const uint16_t* base, int64_t ld, int p, int f0) {
  return *reinterpret_cast<const u16x8*>(base + (size_t)p * ld + f0);
}

__device__ __forceinline__ void store8(uint16_t* base, int64_t ld, int p, int f0, u16x8 v) {
  *reinterpret_cast<u16x8*>(base + (size_t)p * ld + f0) = v;
}

// Output feature of row i of MFMA block nb in the packed W: blocks 2m and
// 2m + 1 share features 32 m .. 32 m + 31, row 4 q + r of block 2m + h
// holding feature 32 m + 8 q + 4 h + r.
__host__ __device__ constexpr int packed_feature(int nb, int i) {
  return 32 * (nb >> 1) + 8 * (i >> 2) + 4 * (nb & 1) + (i & 3);
}

// Wave tile: 16 PB points; the workgroup's waves stride over the tiles.
template <int KS, int PB, int PRO, int EPI>
__global__ __launch_bounds__(64 * kGemmWaves) void mlp_gemm_kernel(GemmArgs a) {
  constexpr int kTile = 16 * PB;
  constexpr bool kAdj = adj_epi(EPI);
  // W in LDS as [nbN][KS][64] fragments, zero past ksK (so every product
  // runs KS steps without a guard), then the bias
  extern __shared__ bf8 wl[];
  float* bias_l = reinterpret_cast<float*>(wl + a.nbN * KS * kWave);
  const int nw = a.nbN * KS * kWave;
  for (int i = threadIdx.x; i < nw; i += blockDim.x) {
    const int l = i & (kWave - 1), ks = (i / kWave) % KS, nb = i / (kWave * KS);
    wl[i] = ks < a.ksK ? a.wpack[(nb * a.ksK + ks) * kWave + l] : __builtin_bit_cast(bf8, zero8());
  }
  for (int i = threadIdx.x; i < 16 * (a.nbN + 1); i += blockDim.x)
    bias_l[i] = (a.bias != nullptr && i < a.N) ? a.bias[i] : 0.0f;
  __syncthreads();
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int tiles = (a.M + kTile - 1) / kTile;
  const int stride = gridDim.x * kGemmWaves;
  int t = blockIdx.x * kGemmWaves + wave;
  u16x8 xr[PB][KS], qr[PB][KS];
  // rows past M load row M - 1 (their results are not stored)
  auto load_tile = [&](int tt) {
#pragma unroll
    for (int pb = 0; pb < PB; ++pb) {
      const int p = min(tt * kTile + 16 * pb + j, a.M - 1);
      const uint16_t* xp = a.x + (size_t)p * a.ldx + 8 * g;
      const uint16_t* qp = PRO == PRO_MUL_SLOPE ? a.x2 + (size_t)p * a.ldx2 + 8 * g : nullptr;
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) load_step<PRO>(a, xp, qp, ks, g, xr[pb][ks], qr[pb][ks]);
    }
  };
  // The next tile's inputs are loaded while this one is multiplied, except
  // with an adjoint epilogue: its operands take those registers (they are
  // loaded per chunk of output blocks, ahead of the chunk's products).
  constexpr bool kPrefetch = !kAdj;
  if (kPrefetch && t < tiles) load_tile(t);
  for (; t < tiles; t += stride) {
    const int p0 = t * kTile;
    const bool full_tile = p0 + kTile <= a.M;
    if constexpr (!kPrefetch) load_tile(t);
    bf8 xf[PB][KS];
#pragma unroll
    for (int pb = 0; pb < PB; ++pb)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) xf[pb][ks] = make_frag<PRO>(a, ks, g, xr[pb][ks], qr[pb][ks]);
    if (kPrefetch && t + stride < tiles) load_tile(t + stride);
    // chunks of output blocks: a rolled loop (one chunk's code in the
    // instruction cache); arrays inside are indexed by unrolled inner loops
#pragma unroll 1
    for (int ci = 0; ci * kNbChunk < a.nbN; ++ci) {
      // every block of the chunk inside N and every point inside M: no guards
      const bool fast = full_tile && 16 * (ci + 1) * kNbChunk <= a.N;
      // the chunk's epilogue operands (8 features per lane and block pair),
      // in flight during its products
      constexpr int kPairs = kNbChunk / 2;
      u16x8 ea[PB][kAdj ? kPairs : 1], eb[PB][kAdj ? kPairs : 1];
      if constexpr (kAdj) {
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) {
          const int p = min(p0 + 16 * pb + j, a.M - 1);
#pragma unroll
          for (int pm = 0; pm < kPairs; ++pm) {
            const int f0 = 16 * ci * kNbChunk + 32 * pm + 8 * g;
            if (fast || f0 < a.N) {
              ea[pb][pm] = load8(a.e0, a.lde0, p, f0);
              eb[pb][pm] = e1_epi(EPI) ? load8(a.e1, a.lde1, p, f0) : zero8();
            } else {
              ea[pb][pm] = zero8();
              eb[pb][pm] = zero8();
            }
          }
        }
      }
      f32x4v acc[PB][kNbChunk];
#pragma unroll
      for (int c = 0; c < kNbChunk; ++c)
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) acc[pb][c] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
      // blocks past nbN (a partial last chunk) multiply the last block's
      // weights again; the epilogue stores none of them
#pragma unroll
      for (int c = 0; c < kNbChunk; ++c) {
        const bf8* wrow = wl + min(ci * kNbChunk + c, a.nbN - 1) * (KS * kWave) + lane;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const bf8 w = wrow[ks * kWave];
#pragma unroll
          for (int pb = 0; pb < PB; ++pb)
            acc[pb][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, xf[pb][ks], acc[pb][c], 0, 0, 0);
        }
      }
      // lane (j, g) holds features f0 .. f0 + 7 of point p0 + 16 pb + j in
      // blocks 2 pm (the first 4) and 2 pm + 1 of the chunk
      auto emit = [&](int pb, int pm) {
        const int p = p0 + 16 * pb + j;
        const int f0 = 16 * ci * kNbChunk + 32 * pm + 8 * g;
        const f32x4v b0 = *reinterpret_cast<const f32x4v*>(bias_l + f0);
        const f32x4v b1 = *reinterpret_cast<const f32x4v*>(bias_l + f0 + 4);
        u16x8 o, o2;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const float z = (e < 4 ? acc[pb][2 * pm][e] + b0[e] : acc[pb][2 * pm + 1][e - 4] + b1[e - 4]);
          if constexpr (EPI == EPI_BIAS) {
            o[e] = to_bf(z);
          } else if constexpr (EPI == EPI_SOFTPLUS) {
            o[e] = to_bf(softplus_of<false>(z, a.beta, a.inv_beta).y);
          } else if constexpr (EPI == EPI_SOFTPLUS_TRAIN) {
            const Softplus sp = softplus_of<true>(z, a.beta, a.inv_beta);
            o[e] = to_bf(sp.y);
            o2[e] = to_hf(sp.q);
          } else if constexpr (EPI == EPI_RELU) {
            o[e] = to_bf(fmaxf(z, 0.0f));
          } else if constexpr (EPI == EPI_GRAD_ADJ) {
            // y = vbar (the product), y2 = vbar * u * beta * s (1 - s)
            o[e] = to_bf(z);
            o2[e] = to_bf(z * bf(eb[pb][pm][e]) * (a.beta * code_curv(ea[pb][pm][e])));
          } else if constexpr (EPI == EPI_LAYER_ADJ) {  // zbar = hbar * s + t
            o[e] = to_bf(fmaf(z, code_slope(ea[pb][pm][e]), bf(eb[pb][pm][e])));
          } else if constexpr (EPI == EPI_MUL_SLOPE) {
            o[e] = to_bf(z * code_slope(ea[pb][pm][e]));
          } else if constexpr (EPI == EPI_GRAD_ADJ_S) {
            // y = vbar * s, y2 = vbar * (u s) * beta * (1 - s)
            o[e] = to_bf(z * code_slope(ea[pb][pm][e]));
            o2[e] = to_bf(z * bf(eb[pb][pm][e]) * (a.beta * code_one_minus_slope(ea[pb][pm][e])));
          } else {  // EPI_GRAD_ADJ_SU: y = vbar * s, y2 = vbar * u * beta * s (1 - s)
            o[e] = to_bf(z * code_slope(ea[pb][pm][e]));
            o2[e] = to_bf(z * bf(eb[pb][pm][e]) * (a.beta * code_curv(ea[pb][pm][e])));
          }
        }
        store8(a.y, a.ldy, p, f0, o);
        if constexpr (two_out_epi(EPI)) store8(a.y2, a.ldy2, p, f0, o2);
      };
      if (fast) {
#pragma unroll
        for (int pb = 0; pb < PB; ++pb)
#pragma unroll
          for (int pm = 0; pm < kPairs; ++pm) emit(pb, pm);
      } else {
#pragma unroll
        for (int pb = 0; pb < PB; ++pb) {
          if (p0 + 16 * pb + j >= a.M) continue;
#pragma unroll
          for (int pm = 0; pm < kPairs; ++pm)
            if (16 * ci * kNbChunk + 32 * pm + 8 * g < a.N) emit(pb, pm);
        }
      }
    }
  }
}

// W [N,K] (FP32, w[n * ldw + k], or w[k * ldw + n] with transpose) times
// `scale` -> the A operands in LDS order: [nb][ks][lane] of 8 BF16, rows in
// packed_feature order, zero padded to 16 nb x 32 ks.
__global__ __launch_bounds__(256) void mlp_pack_kernel(const float* __restrict__ w, int64_t ldw,
                                                       int transpose, float scale, int N, int K,
                                                       int ksK, u16x8* __restrict__ out,
                                                       int total) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int lane = idx & 63;
  const int ks = (idx >> 6) % ksK;
  const int nb = (idx >> 6) / ksK;
  const int n = packed_feature(nb, lane & 15);
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

// ------------------------------------------------------------------ chained forward
//
// The SDF network's eight hidden layers in one kernel, for the forward
// alone (the sdf of each point, no intermediate in HBM) or for training
// (h_l and the slope codes q_l stored, as the layer products store them).
// A wave carries 32 points through every layer with their activations in
// registers: the packed weights permute each layer's output rows so that
// the BF16 outputs of blocks 2m and 2m + 1 in lane (j, g) are exactly the
// B fragment of step m of the next layer (8 consecutive features of point
// j), so a layer's epilogue writes the next layer's operands directly.
// The weights stream through LDS in chunks of 4 output blocks (32 KB,
// double-buffered): while a chunk is multiplied, the next one is loaded
// into registers and written to the other buffer, one workgroup barrier
// per chunk.  A workgroup (8 waves, 256 points per tile) also keeps its
// tile's encoded inputs gamma in LDS, for layer 0 and the skip at layer 4.
// Every product, bias and softplus is the layer product's own (same
// instruction, step order and rounding), so h_l and q_l are bit for bit
// those of the layer-by-layer form; the sdf is W_8[0] . h_7 + b_8[0] summed
// in FP32 (the layer form rounds it to BF16).

constexpr int kChainWaves = 8;
constexpr int kChainPB = 2;                               // 16-point blocks per wave
constexpr int kChainPts = 16 * kChainPB * kChainWaves;    // points per workgroup tile
constexpr int kChainBlk = 4;                              // output blocks per weight chunk
constexpr int kChainFrags = kChainBlk * 8 * kWave;        // bf8 per chunk buffer (32 KB)
constexpr int kChainLoads = kChainFrags / (kWave * kChainWaves);  // per thread
constexpr int kGamStride = 72;                            // halfwords per gamma row in LDS

struct ChainArgs {
  const float* x;      // [M,3]
  const bf8* w;        // packed W_0 .. W_7 back to back (layer 0: K = din; layer 3: N = n3)
  const float* bias;   // [8][256]
  const float* w_sdf;  // W_8[0] (rounded to BF16 here, as the pack does), then b_8[0]
  float* sdf;          // [M] or null
  uint16_t* h;         // layer l at h + l * sh, [M, ldh] BF16, or null (training)
  uint16_t* q;         // the FP16 codes, likewise
  int64_t ldh, ldq, sh, sq;
  float scale, beta, inv_beta;
  int M, L, din, n3, nb3;
};

// fragments (bf8) before layer l in the packed buffer: layer 0 has 2 steps
// of 32 inputs, the others 8; layer 3 has nb3 output blocks, the others 16
__device__ __forceinline__ int chain_w_offset(const ChainArgs& a, int l) {
  if (l == 0) return 0;
  const int full = l - 1 - (l > 3 ? 1 : 0);  // layers 1 .. l - 1 other than 3
  return (16 * 2 + 16 * 8 * full + (l > 3 ? a.nb3 * 8 : 0)) * kWave;
}

// gamma feature f (< din) of the point v = x * scale, as pe_kernel forms it
__device__ __forceinline__ float gamma_feature(const float (&v)[3], int f) {
  if (f < 3) return v[f];
  const int b = (f - 3) / 6, r = (f - 3) % 6, d = r % 3;
  float sn, cs;
  sincosf(v[d] * (float)(1 << b), &sn, &cs);
  return r < 3 ? sn : cs;
}

template <bool TRAIN>
__global__ __launch_bounds__(64 * kChainWaves) void mlp_sdf_chain_kernel(ChainArgs a) {
  extern __shared__ bf8 cw[];  // [2][kChainFrags] weight chunks
  float* bias_l = reinterpret_cast<float*>(cw + 2 * kChainFrags);  // [8][256]
  float* wsdf_l = bias_l + 8 * 256;                                 // [257]: w, b
  uint16_t* gam = reinterpret_cast<uint16_t*>(wsdf_l + 260);        // [kChainPts][kGamStride]
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  for (int i = tid; i < 8 * 256; i += blockDim.x) {
    const int l = i >> 8, f = i & 255;
    const int n = l == 3 ? a.n3 : 256;
    bias_l[i] = f < n ? a.bias[i] : 0.0f;
  }
  if (a.sdf != nullptr)
    for (int i = tid; i < 257; i += blockDim.x)
      wsdf_l[i] = i < 256 ? bf(to_bf(a.w_sdf[i])) : a.w_sdf[i];
  const int tiles = (a.M + kChainPts - 1) / kChainPts;

  // chunk ci of layer l: blocks 4 ci .. 4 ci + 3 of the packed W_l, steps
  // of 32 inputs each (2 for layer 0), zero past the layer's blocks
  bf8 stage[kChainLoads];
  auto load_chunk = [&](int l, int ci) __attribute__((always_inline)) {
    const int ks = l == 0 ? 2 : 8;
    const int nb = l == 3 ? a.nb3 : 16;
    const int valid = max(0, min(kChainBlk, nb - kChainBlk * ci)) * ks * kWave;
    const bf8* src = a.w + chain_w_offset(a, l) + kChainBlk * ci * ks * kWave;
#pragma unroll
    for (int i = 0; i < kChainLoads; ++i) {
      const int e = tid + i * kWave * kChainWaves;
      stage[i] = e < valid ? src[e] : __builtin_bit_cast(bf8, zero8());
    }
  };
  auto store_chunk = [&](int buf, int l) __attribute__((always_inline)) {
    const int n = (l == 0 ? 2 : 8) * kChainBlk * kWave;
#pragma unroll
    for (int i = 0; i < kChainLoads; ++i) {
      const int e = tid + i * kWave * kChainWaves;
      if (e < n) cw[buf * kChainFrags + e] = stage[i];
    }
  };
  int t = blockIdx.x;
  if (t < tiles) {
    load_chunk(0, 0);
    store_chunk(0, 0);
  }

  bf8 B[kChainPB][8], Bn[kChainPB][8];
  for (; t < tiles; t += gridDim.x) {
    const bool more = t + (int)gridDim.x < tiles;
    // this tile's gamma rows: two threads per point (x and bands < 3, bands >= 3)
    {
      const int pt = tid >> 1, half = tid & 1;
      const int p = min(t * kChainPts + pt, a.M - 1);
      float v[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) v[d] = a.x[3 * (size_t)p + d] * a.scale;
      uint16_t* row = gam + pt * kGamStride;
      const int fa = half == 0 ? 0 : 3 + 6 * ((a.L + 1) / 2);
      const int fb = half == 0 ? 3 + 6 * ((a.L + 1) / 2) : 64;
      for (int f = fa; f < fb; ++f) row[f] = f < a.din ? to_bf(gamma_feature(v, f)) : (uint16_t)0;
    }
    const int pw = t * kChainPts + 32 * wave;  // the wave's first point
    // a lane's fragment of gamma for inputs 32 ks + 8 g .. + 7 of its point
    auto gamma_frag = [&](int pb, int ks) __attribute__((always_inline)) {
      const uint16_t* row = gam + (32 * wave + 16 * pb + j) * kGamStride + 32 * ks + 8 * g;
      return *reinterpret_cast<const bf8*>(row);
    };

    // one chunk of layer l: products, epilogue into Bn (and the training
    // stores); the next chunk's weights loaded meanwhile
    auto chunk = [&](auto ks_tag, int l, int ci) __attribute__((always_inline)) {
      constexpr int KS = decltype(ks_tag)::value;
      __syncthreads();  // chunk ci is in its buffer; the other buffer is free
      const int nl = ci < 3 ? l : l + 1;
      const int nci = ci < 3 ? ci + 1 : 0;
      const bool has_next = nl < 8 || more;
      if (has_next) load_chunk(nl & 7, nci);
      const bf8* wb = cw + (ci & 1) * kChainFrags + lane;
      f32x4v acc[kChainPB][kChainBlk];
#pragma unroll
      for (int b = 0; b < kChainBlk; ++b)
#pragma unroll
        for (int pb = 0; pb < kChainPB; ++pb) acc[pb][b] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int b = 0; b < kChainBlk; ++b) {
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          const bf8 w = wb[(b * KS + ks) * kWave];
#pragma unroll
          for (int pb = 0; pb < kChainPB; ++pb)
            acc[pb][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, B[pb][ks], acc[pb][b], 0, 0, 0);
        }
      }
      const float* bl = bias_l + 256 * l;
#pragma unroll
      for (int pb = 0; pb < kChainPB; ++pb) {
        const int p = pw + 16 * pb + j;
#pragma unroll
        for (int pm = 0; pm < 2; ++pm) {
          const int f0 = 64 * ci + 32 * pm + 8 * g;
          const f32x4v b0 = *reinterpret_cast<const f32x4v*>(bl + f0);
          const f32x4v b1 = *reinterpret_cast<const f32x4v*>(bl + f0 + 4);
          u16x8 o, o2;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const float z = e < 4 ? acc[pb][2 * pm][e] + b0[e] : acc[pb][2 * pm + 1][e - 4] + b1[e - 4];
            const Softplus sp = softplus_of<TRAIN>(z, a.beta, a.inv_beta);
            o[e] = to_bf(sp.y);
            if (TRAIN) o2[e] = to_hf(sp.q);
          }
          if (l == 3 && f0 + 8 > a.n3) {
            // the skip input [h_3, gamma]: gamma past layer 3's outputs, slope 1
            const uint16_t* row = gam + (32 * wave + 16 * pb + j) * kGamStride;
#pragma unroll
            for (int e = 0; e < 8; ++e)
              if (f0 + e >= a.n3) {
                o[e] = row[f0 + e - a.n3];
                o2[e] = 0;
              }
          }
          Bn[pb][2 * ci + pm] = __builtin_bit_cast(bf8, o);
          if (TRAIN && p < a.M) {
            store8(a.h + l * a.sh, a.ldh, p, f0, o);
            store8(a.q + l * a.sq, a.ldq, p, f0, o2);
          }
        }
      }
      if (has_next) store_chunk(ci & 1 ? 0 : 1, nl & 7);
    };
    using Ks2 = std::integral_constant<int, 2>;
    using Ks8 = std::integral_constant<int, 8>;

    // layer 0 reads gamma (visible after the first chunk's barrier)
    __syncthreads();
#pragma unroll
    for (int pb = 0; pb < kChainPB; ++pb) {
      B[pb][0] = gamma_frag(pb, 0);
      B[pb][1] = gamma_frag(pb, 1);
    }
    chunk(Ks2{}, 0, 0);
    chunk(Ks2{}, 0, 1);
    chunk(Ks2{}, 0, 2);
    chunk(Ks2{}, 0, 3);
#pragma unroll 1
    for (int l = 1; l < 8; ++l) {
#pragma unroll
      for (int pb = 0; pb < kChainPB; ++pb)
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) B[pb][ks] = Bn[pb][ks];
      chunk(Ks8{}, l, 0);
      chunk(Ks8{}, l, 1);
      chunk(Ks8{}, l, 2);
      chunk(Ks8{}, l, 3);
    }
    if (a.sdf != nullptr) {
#pragma unroll
      for (int pb = 0; pb < kChainPB; ++pb) {
        float s = 0.0f;
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
          const u16x8 hv = __builtin_bit_cast(u16x8, Bn[pb][ks]);
#pragma unroll
          for (int e = 0; e < 8; ++e) s = fmaf(bf(hv[e]), wsdf_l[32 * ks + 8 * g + e], s);
        }
        s += __shfl_xor(s, 16, kWave);
        s += __shfl_xor(s, 32, kWave);
        const int p = pw + 16 * pb + j;
        if (g == 0 && p < a.M) a.sdf[p] = s + wsdf_l[256];
      }
    }
  }
}

// The reverse sweeps chained the same way: the normal pass (g_{l-1} =
// (g_l W_l) * s_{l-1}, from g_7 = u_7 * s_7) and the primal adjoint (zbar_{l-1}
// = (zbar_l W_l) * s_{l-1} + t_{l-1}) over `steps` layers of 256 x 256 (W_l^T
// packed, 16 blocks x 8 steps each, back to back in the order of the
// sweep).  Each step's output is stored (the weight gradients and the later
// passes read it) and kept in registers as the next step's operand, so a
// step reads the codes (and t) and writes its output: no operand re-read.
// Step s reads e0 + s * se0 (and e1 + s * se1) and writes y + s * sy; y may
// be e1 (in place: each lane reads its elements before it writes them).
// Products, step order and epilogues are the layer product's (epilogues 6
// and 5), so the outputs are bit for bit those of the layer-by-layer form.
// The same sweep serves ReLU networks (RELU: the slopes are the masks
// a > 0 of BF16 activations instead of softplus codes), with the first
// operand optionally scaled per point (PRO 2: x * s(x2) * xs[p]) and stored
// (x_out), and steps from mask_steps on unmasked (y = z [+ e1]).
struct ChainAdjArgs {
  const uint16_t* x;   // the first operand [M, ldx] (ldx 0: one broadcast row)
  const uint16_t* x2;  // its slopes (prologue x * s(x2)) or null
  const float* xs;     // per-point scale of the prologue (PRO 2)
  uint16_t* x_out;     // the first operand as multiplied, or null
  int64_t ldx, ldx2, ldxo;
  const bf8* w;
  const uint16_t* e0;
  const uint16_t* e1;  // null: no addend
  uint16_t* y;
  int64_t lde0, se0, lde1, se1, ldy, sy;
  int M, steps, mask_steps;
};

template <bool RELU>
__device__ __forceinline__ float sweep_slope(uint16_t m) {
  if constexpr (RELU) return bf(m) > 0.0f ? 1.0f : 0.0f;
  return code_slope(m);
}

template <int PRO, bool ADD, bool RELU>
__global__ __launch_bounds__(64 * kChainWaves) void mlp_chain_adj_kernel(ChainAdjArgs a) {
  extern __shared__ bf8 cw[];  // [2][kChainFrags] weight chunks
  const int tid = threadIdx.x;
  const int lane = tid & (kWave - 1);
  const int wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int tiles = (a.M + kChainPts - 1) / kChainPts;
  constexpr int kStepFrags = 16 * 8 * kWave;
  bf8 stage[kChainLoads];
  auto load_chunk = [&](int st, int ci) __attribute__((always_inline)) {
    const bf8* src = a.w + (size_t)st * kStepFrags + kChainBlk * ci * 8 * kWave;
#pragma unroll
    for (int i = 0; i < kChainLoads; ++i) stage[i] = src[tid + i * kWave * kChainWaves];
  };
  auto store_chunk = [&](int buf) __attribute__((always_inline)) {
#pragma unroll
    for (int i = 0; i < kChainLoads; ++i)
      cw[buf * kChainFrags + tid + i * kWave * kChainWaves] = stage[i];
  };
  int t = blockIdx.x;
  if (t < tiles) {
    load_chunk(0, 0);
    store_chunk(0);
  }
  bf8 B[kChainPB][8], Bn[kChainPB][8];
  for (; t < tiles; t += gridDim.x) {
    const bool more = t + (int)gridDim.x < tiles;
    const int pw = t * kChainPts + 32 * wave;
    // the first operand (with the slope prologue), rows past M read row M - 1
#pragma unroll
    for (int pb = 0; pb < kChainPB; ++pb) {
      const int p = min(pw + 16 * pb + j, a.M - 1);
      const float sc = PRO == 2 ? a.xs[p] : 1.0f;
#pragma unroll
      for (int ks = 0; ks < 8; ++ks) {
        u16x8 v = load8(a.x, a.ldx, p, 32 * ks + 8 * g);
        if (PRO > 0) {
          const u16x8 c = load8(a.x2, a.ldx2, p, 32 * ks + 8 * g);
#pragma unroll
          for (int e = 0; e < 8; ++e)
            v[e] = PRO == 2 ? to_bf(bf(v[e]) * sc * sweep_slope<RELU>(c[e]))
                            : to_bf(bf(v[e]) * sweep_slope<RELU>(c[e]));
        }
        if (a.x_out != nullptr && pw + 16 * pb + j < a.M)
          store8(a.x_out, a.ldxo, p, 32 * ks + 8 * g, v);
        B[pb][ks] = __builtin_bit_cast(bf8, v);
      }
    }
    auto chunk = [&](int st, int ci) __attribute__((always_inline)) {
      __syncthreads();  // chunk ci is in its buffer; the other buffer is free
      const int nst = ci < 3 ? st : st + 1;
      const int nci = ci < 3 ? ci + 1 : 0;
      const bool has_next = nst < a.steps || more;
      if (has_next) load_chunk(nst < a.steps ? nst : 0, nci);
      const bool masked = st < a.mask_steps;  // uniform
      const uint16_t* e0 = masked ? a.e0 + st * a.se0 : nullptr;
      const uint16_t* e1 = ADD ? a.e1 + st * a.se1 : nullptr;
      u16x8 ea[kChainPB][2], eb[kChainPB][2];
#pragma unroll
      for (int pb = 0; pb < kChainPB; ++pb) {
        const int p = min(pw + 16 * pb + j, a.M - 1);
#pragma unroll
        for (int pm = 0; pm < 2; ++pm) {
          const int f0 = 64 * ci + 32 * pm + 8 * g;
          ea[pb][pm] = masked ? load8(e0, a.lde0, p, f0) : zero8();
          if (ADD) eb[pb][pm] = load8(e1, a.lde1, p, f0);
        }
      }
      const bf8* wb = cw + (ci & 1) * kChainFrags + lane;
      f32x4v acc[kChainPB][kChainBlk];
#pragma unroll
      for (int b = 0; b < kChainBlk; ++b)
#pragma unroll
        for (int pb = 0; pb < kChainPB; ++pb) acc[pb][b] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int b = 0; b < kChainBlk; ++b) {
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
          const bf8 w = wb[(b * 8 + ks) * kWave];
#pragma unroll
          for (int pb = 0; pb < kChainPB; ++pb)
            acc[pb][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, B[pb][ks], acc[pb][b], 0, 0, 0);
        }
      }
      uint16_t* y = a.y + st * a.sy;
#pragma unroll
      for (int pb = 0; pb < kChainPB; ++pb) {
        const int p = pw + 16 * pb + j;
#pragma unroll
        for (int pm = 0; pm < 2; ++pm) {
          const int f0 = 64 * ci + 32 * pm + 8 * g;
          u16x8 o;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            const float z = e < 4 ? acc[pb][2 * pm][e] : acc[pb][2 * pm + 1][e - 4];
            const float sl = masked ? sweep_slope<RELU>(ea[pb][pm][e]) : 1.0f;
            o[e] = ADD ? to_bf(fmaf(z, sl, bf(eb[pb][pm][e]))) : to_bf(z * sl);
          }
          Bn[pb][2 * ci + pm] = __builtin_bit_cast(bf8, o);
          if (p < a.M) store8(y, a.ldy, p, f0, o);
        }
      }
      if (has_next) store_chunk(ci & 1 ? 0 : 1);
    };
#pragma unroll 1
    for (int st = 0; st < a.steps; ++st) {
      chunk(st, 0);
      chunk(st, 1);
      chunk(st, 2);
      chunk(st, 3);
#pragma unroll
      for (int pb = 0; pb < kChainPB; ++pb)
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) B[pb][ks] = Bn[pb][ks];
    }
  }
}

// ------------------------------------------------------------------ weight gradients

constexpr int kWgWaves = 8;
constexpr int kWgTile = 256;          // dW block per workgroup: 256 (+ 16) x 256
constexpr int kWgRows = kWgTile + 16;  // rows of a partial block (N up to 272)
constexpr int kWgCols = kWgTile + 1;   // columns of a partial block (the last: the bias)
constexpr int kWgMaxBlocks = 256;      // point chunks (partial blocks) per launch

struct WgradArgs {
  const uint16_t* g;
  const uint16_t* g2;
  const uint16_t* x;
  const uint16_t* x2;
  float* partial;  // [gridDim.x][kWgRows][kWgCols] for this (n, k) super-tile
  int64_t ldg, ldg2, ldx, ldx2;
  int M, N, K;
  int with_bias;
  int pts_per_block;
};

// One staging task: 4 consecutive points x 8 consecutive columns of a
// [M, C] tensor (16-byte loads; zeros past pend or C).
template <bool PRO>
__device__ __forceinline__ void load_task(const uint16_t* src, int64_t ld, const uint16_t* src2,
                                          int64_t ld2, int pend, int C, int p, int c, u16x8 (&r)[4],
                                          u16x8 (&q)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    r[i] = zero8();
    if (PRO) q[i] = zero8();
    if (p + i < pend && c < C) {
      r[i] = *reinterpret_cast<const u16x8*>(src + (size_t)(p + i) * ld + c);
      if (PRO) q[i] = *reinterpret_cast<const u16x8*>(src2 + (size_t)(p + i) * ld2 + c);
    }
  }
}

// The task's prologue, then its transpose into t[column][point] (row stride S
// halfwords): one 8-byte LDS store per column.
template <bool PRO, int S>
__device__ __forceinline__ void store_task(uint16_t* t, int C, int c, int col, int pt,
                                           u16x8 (&r)[4], const u16x8 (&q)[4]) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    if (PRO) {
#pragma unroll
      for (int e = 0; e < 8; ++e) r[i][e] = to_bf(bf(r[i][e]) * code_slope(q[i][e]));
    }
    if (c + 8 > C) {
#pragma unroll
      for (int e = 0; e < 8; ++e)
        if (c + e >= C) r[i][e] = 0;
    }
  }
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    const u16x4 v = {r[0][e], r[1][e], r[2][e], r[3][e]};
    *reinterpret_cast<u16x4*>(t + (col + e) * S + pt) = v;
  }
}

// Workgroup: 8 waves in 4 (n) x 2 (k); wave tile 64 (n) x 128 (k) of a
// 256 x 256 block of dW (the host offsets g and x to the block's columns);
// with EXTRA, rows 256 .. 271 as well (each wave 16 x 32 of them).  STEP
// points are staged per synchronisation: 64 for plain operands (each thread
// loads one task of G and one of X), 32 with a prologue or EXTRA (waves 0-3
// load G, waves 4-7 X; with EXTRA 16 threads of wave 4 also the extra rows).
template <int PRO_G, int PRO_X, int STEP, bool EXTRA>
__global__ __launch_bounds__(64 * kWgWaves) void mlp_wgrad_kernel(WgradArgs a) {
  static_assert(STEP == 32 || STEP == 64, "staging step");
  static_assert(STEP == 32 || (PRO_G == PRO_NONE && PRO_X == PRO_NONE && !EXTRA),
                "64-point steps stage plain operands only");
  constexpr int S = STEP + 8;  // LDS row stride (halfwords)
  constexpr int kRowsG = EXTRA ? kWgRows : kWgTile;
  extern __shared__ uint16_t wg_lds[];
  uint16_t* gt = wg_lds;               // [kRowsG][S]: G^T
  uint16_t* xt = wg_lds + kRowsG * S;  // [256][S]: X^T
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int wn = wave & 3, wk = wave >> 2;
  const int pbeg = blockIdx.x * a.pts_per_block;
  const int pend = min(a.M, pbeg + a.pts_per_block);
  const bool bias_col = a.with_bias && wk == 0;
  // staging tasks of this thread
  const bool g_task = STEP == 64 || wave < 4;
  int tp, tc;  // task's first point (in the step) and first column
  if (STEP == 64) {
    tp = 4 * (lane & 15);
    tc = 8 * (4 * wave + (lane >> 4));
  } else {
    tp = 4 * (lane & 7);
    tc = 8 * (8 * (wave & 3) + (lane >> 3));
  }
  const bool x_extra = EXTRA && wave == 4 && lane < 16;  // rows 256 .. 271 of G^T
  const int xp = 4 * (lane & 7), xc = kWgTile + 8 * (lane >> 3);
  u16x8 r0[4], q0[4], r1[4], q1[4];
  auto load_step = [&](int p0) {
    if (STEP == 64) {
      load_task<false>(a.g, a.ldg, nullptr, 0, pend, a.N, p0 + tp, tc, r0, q0);
      load_task<false>(a.x, a.ldx, nullptr, 0, pend, a.K, p0 + tp, tc, r1, q1);
    } else {
      if (g_task)
        load_task<PRO_G == PRO_MUL_SLOPE>(a.g, a.ldg, a.g2, a.ldg2, pend, a.N, p0 + tp, tc, r0, q0);
      else
        load_task<PRO_X == PRO_MUL_SLOPE>(a.x, a.ldx, a.x2, a.ldx2, pend, a.K, p0 + tp, tc, r0, q0);
      if (EXTRA && x_extra) load_task<false>(a.g, a.ldg, nullptr, 0, pend, a.N, p0 + xp, xc, r1, q1);
    }
  };
  f32x4v acc[4][8];
  f32x4v bacc[4];
  f32x4v accx[2], baccx;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    bacc[i] = f32x4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) acc[i][kk] = bacc[i];
  }
  accx[0] = accx[1] = baccx = bacc[0];
  u16x8 ones_u;
#pragma unroll
  for (int e = 0; e < 8; ++e) ones_u[e] = 0x3f80;  // 1.0 in BF16
  const bf8 ones = __builtin_bit_cast(bf8, ones_u);
  if (pbeg < pend) load_step(pbeg);
  for (int p0 = pbeg; p0 < pend; p0 += STEP) {
    __syncthreads();  // the previous step's products are done with the tiles
    if (STEP == 64) {
      store_task<false, S>(gt, a.N, tc, tc, tp, r0, q0);
      store_task<false, S>(xt, a.K, tc, tc, tp, r1, q1);
    } else {
      if (g_task)
        store_task<PRO_G == PRO_MUL_SLOPE, S>(gt, a.N, tc, tc, tp, r0, q0);
      else
        store_task<PRO_X == PRO_MUL_SLOPE, S>(xt, a.K, tc, tc, tp, r0, q0);
      if (EXTRA && x_extra) store_task<false, S>(gt, a.N, xc, xc, xp, r1, q1);
    }
    __syncthreads();
    if (p0 + STEP < pend) load_step(p0 + STEP);
    // A = G^T rows (n), depth = 32 points; B = X (points x k)
#pragma unroll
    for (int ks = 0; ks < STEP / 32; ++ks) {
      bf8 af[4], bfr[8];
#pragma unroll
      for (int i = 0; i < 4; ++i)
        af[i] = *reinterpret_cast<const bf8*>(gt + (64 * wn + 16 * i + j) * S + 32 * ks + 8 * g);
#pragma unroll
      for (int kk = 0; kk < 8; ++kk)
        bfr[kk] =
            *reinterpret_cast<const bf8*>(xt + (128 * wk + 16 * kk + j) * S + 32 * ks + 8 * g);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int kk = 0; kk < 8; ++kk)
          acc[i][kk] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[kk], acc[i][kk], 0, 0, 0);
        if (bias_col)
          bacc[i] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], ones, bacc[i], 0, 0, 0);
      }
      if (EXTRA) {
        const bf8 ax = *reinterpret_cast<const bf8*>(gt + (kWgTile + j) * S + 32 * ks + 8 * g);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const bf8 bx = *reinterpret_cast<const bf8*>(
              xt + (128 * wk + 16 * (2 * wn + h) + j) * S + 32 * ks + 8 * g);
          accx[h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax, bx, accx[h], 0, 0, 0);
        }
        if (a.with_bias && wave == 4)
          baccx = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ax, ones, baccx, 0, 0, 0);
      }
    }
  }
  // partial block [kWgRows][kWgCols] (the last column: the bias gradient)
  float* out = a.partial + (size_t)blockIdx.x * kWgRows * kWgCols;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = 64 * wn + 16 * i + 4 * g + r;
#pragma unroll
      for (int kk = 0; kk < 8; ++kk)
        out[(size_t)n * kWgCols + 128 * wk + 16 * kk + j] = acc[i][kk][r];
      if (bias_col && j == 0) out[(size_t)n * kWgCols + kWgTile] = bacc[i][r];
    }
  }
  if (EXTRA) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int n = kWgTile + 4 * g + r;
#pragma unroll
      for (int h = 0; h < 2; ++h)
        out[(size_t)n * kWgCols + 128 * wk + 16 * (2 * wn + h) + j] = accx[h][r];
      if (a.with_bias && wave == 4 && j == 0) out[(size_t)n * kWgCols + kWgTile] = baccx[r];
    }
  }
}

// dW[n][k] (+)= the sum over the partial blocks; db[n] likewise.  A
// workgroup takes 64 consecutive cells: wave w sums the blocks b = w mod 8
// (coalesced rows, eight waves in flight per cell instead of one serial
// chain), then the eight sums are added in wave order (deterministic).
constexpr int kRedWaves = 8;

__global__ __launch_bounds__(64 * kRedWaves) void mlp_wgrad_reduce_kernel(
    const float* __restrict__ partial, int blocks, int rows, int n0, int k0, int N, int K,
    float* __restrict__ dw, int64_t lddw, float* __restrict__ db, int accumulate) {
  __shared__ double part[kRedWaves][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x >> 6;
  const int idx = blockIdx.x * kWave + lane;
  const bool in = idx < rows * kWgCols;
  double s = 0.0;
  if (in)
    for (int b = wave; b < blocks; b += kRedWaves)
      s += (double)partial[(size_t)b * kWgRows * kWgCols + idx];
  part[wave][lane] = s;
  __syncthreads();
  if (wave != 0 || !in) return;
  s = part[0][lane];
#pragma unroll
  for (int w = 1; w < kRedWaves; ++w) s += part[w][lane];
  const int n = idx / kWgCols, c = idx % kWgCols;
  const bool is_bias = c == kWgTile;
  if (n0 + n >= N) return;
  if (is_bias ? (db == nullptr || k0 != 0) : (k0 + c >= K)) return;
  if (is_bias) {
    db[n0 + n] = (float)(accumulate ? (double)db[n0 + n] + s : s);
  } else {
    float* o = dw + (size_t)(n0 + n) * lddw + k0 + c;
    *o = (float)(accumulate ? (double)*o + s : s);
  }
}

// ------------------------------------------------------------------ positional encoding
//
// A workgroup takes 256 points: each thread forms (or reads) its point's row
// in LDS (BF16, an odd row stride), and the workgroup moves the rows between
// LDS and HBM a row per wave instruction, consecutive lanes on consecutive
// columns (coalesced 2-byte accesses, any column offset and stride).

constexpr int kPePts = 256;
constexpr int kPeMaxCols = 64;

__device__ __forceinline__ int pe_stride(int cols) { return cols | 1; }

// gamma(x) = [x, sin(2^l x), cos(2^l x) for l < L] (x scaled), BF16, into
// out[p * ldo + c0 + i] for i < 3 + 6 L (the band order of sdf_network.py's
// Embedder: per band the 3 sines, then the 3 cosines); columns up to
// pad_to are zero-filled.
__global__ __launch_bounds__(kPePts) void pe_kernel(const float* __restrict__ x, int M, int L,
                                                    float scale, uint16_t* __restrict__ out,
                                                    int64_t ldo, int c0, int pad_to) {
  extern __shared__ uint16_t pe_rows[];
  const int rs = pe_stride(pad_to);
  const int pbase = blockIdx.x * kPePts;
  const int p = pbase + threadIdx.x;
  uint16_t* row = pe_rows + threadIdx.x * rs;
  if (p < M) {
    float v[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = x[3 * (size_t)p + d] * scale;
      row[d] = to_bf(v[d]);
    }
    for (int l = 0; l < L; ++l) {
      const float f = (float)(1 << l);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float sn, cs;
        sincosf(v[d] * f, &sn, &cs);
        row[3 + 6 * l + d] = to_bf(sn);
        row[3 + 6 * l + 3 + d] = to_bf(cs);
      }
    }
    for (int c = 3 + 6 * L; c < pad_to; ++c) row[c] = 0;
  }
  __syncthreads();
  const int np = min(kPePts, M - pbase);
  for (int i = threadIdx.x; i < np * kPeMaxCols; i += kPePts) {
    const int q = i >> 6, c = i & (kPeMaxCols - 1);
    if (c < pad_to) out[(size_t)(pbase + q) * ldo + c0 + c] = pe_rows[q * rs + c];
  }
}

// n = d sdf / d x from d sdf / d gamma (the sum of two BF16 parts, [M, 3+6L]
// each at their own stride): n_d = scale * (g[d] + sum_l 2^l (cos(2^l x_d)
// g[3+6l+d] - sin(2^l x_d) g[3+6l+3+d])), FP32.
__global__ __launch_bounds__(kPePts) void pe_grad_kernel(const float* __restrict__ x, int M, int L,
                                                         float scale,
                                                         const uint16_t* __restrict__ ga,
                                                         int64_t lda,
                                                         const uint16_t* __restrict__ gb,
                                                         int64_t ldb, float* __restrict__ n) {
  extern __shared__ uint16_t pe_rows[];
  const int nc = 3 + 6 * L;
  const int rs = pe_stride(nc);
  uint16_t* ra = pe_rows;
  uint16_t* rb = pe_rows + kPePts * rs;
  const int pbase = blockIdx.x * kPePts;
  const int np = min(kPePts, M - pbase);
  for (int i = threadIdx.x; i < np * kPeMaxCols; i += kPePts) {
    const int q = i >> 6, c = i & (kPeMaxCols - 1);
    if (c < nc) {
      ra[q * rs + c] = ga[(size_t)(pbase + q) * lda + c];
      rb[q * rs + c] = gb[(size_t)(pbase + q) * ldb + c];
    }
  }
  __syncthreads();
  const int p = pbase + threadIdx.x;
  if (p >= M) return;
  const uint16_t* a = ra + threadIdx.x * rs;
  const uint16_t* b = rb + threadIdx.x * rs;
  float v[3], acc[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v[d] = x[3 * (size_t)p + d] * scale;
    acc[d] = bf(a[d]) + bf(b[d]);
  }
  for (int l = 0; l < L; ++l) {
    const float f = (float)(1 << l);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float sn, cs;
      sincosf(v[d] * f, &sn, &cs);
      const float gs = bf(a[3 + 6 * l + d]) + bf(b[3 + 6 * l + d]);
      const float gc = bf(a[3 + 6 * l + 3 + d]) + bf(b[3 + 6 * l + 3 + d]);
      acc[d] += f * (cs * gs - sn * gc);
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) n[3 * (size_t)p + d] = acc[d] * scale;
}

// The adjoint of pe_grad for d loss / d n: d loss / d gamma = J dn with J the
// encoding's Jacobian (scaled), BF16, into out[p * ldo + c0 + i], zero-padded
// up to pad_to columns.
__global__ __launch_bounds__(kPePts) void pe_grad_adj_kernel(const float* __restrict__ x, int M,
                                                             int L, float scale,
                                                             const float* __restrict__ dn,
                                                             uint16_t* __restrict__ out,
                                                             int64_t ldo, int c0, int pad_to) {
  extern __shared__ uint16_t pe_rows[];
  const int rs = pe_stride(pad_to);
  const int pbase = blockIdx.x * kPePts;
  const int p = pbase + threadIdx.x;
  uint16_t* row = pe_rows + threadIdx.x * rs;
  if (p < M) {
    float v[3], gd[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      v[d] = x[3 * (size_t)p + d] * scale;
      gd[d] = dn[3 * (size_t)p + d] * scale;
      row[d] = to_bf(gd[d]);
    }
    for (int l = 0; l < L; ++l) {
      const float f = (float)(1 << l);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        float sn, cs;
        sincosf(v[d] * f, &sn, &cs);
        row[3 + 6 * l + d] = to_bf(f * cs * gd[d]);
        row[3 + 6 * l + 3 + d] = to_bf(-f * sn * gd[d]);
      }
    }
    for (int c = 3 + 6 * L; c < pad_to; ++c) row[c] = 0;
  }
  __syncthreads();
  const int np = min(kPePts, M - pbase);
  for (int i = threadIdx.x; i < np * kPeMaxCols; i += kPePts) {
    const int q = i >> 6, c = i & (kPeMaxCols - 1);
    if (c < pad_to) out[(size_t)(pbase + q) * ldo + c0 + c] = pe_rows[q * rs + c];
  }
}

// ------------------------------------------------------------------ host side

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
// output blocks of 16: pairs of blocks cover 32 features (packed_feature)
int nb_of(int64_t N) { return (int)(2 * ((N + 31) / 32)); }

// LDS of the layer product: W padded to the kernel's KS steps, and the bias
size_t gemm_lds_bytes(int64_t N, int64_t K) {
  const int ksK = ks_of(K);
  const int ks_kernel = ksK <= 2 ? 2 : (ksK <= 8 ? 8 : 9);
  return (size_t)nb_of(N) * ks_kernel * kWave * 16 + (size_t)16 * (nb_of(N) + 1) * sizeof(float);
}
constexpr size_t kLdsBytes = 160 * 1024;

int cu_count() {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        v > 0)
      cus = v;
  }
  return cus;
}

// Points per wave tile: 16 where the prologue's codes or the epilogue's
// operands take the registers, 32 otherwise (twice the reuse of each
// weight fragment read from LDS).
constexpr int gemm_pb(int pro, int epi) { return (pro == PRO_MUL_SLOPE || adj_epi(epi)) ? 1 : 2; }

template <int KS, int PRO>
int launch_gemm_epi(int epi, const GemmArgs& a, size_t lds, hipStream_t st) {
  const dim3 block(64 * kGemmWaves);
  switch (epi) {
#define GERAF_EPI(E)                                                                        \
  case E: {                                                                                 \
    constexpr int PB = gemm_pb(PRO, E);                                                     \
    const int64_t tiles = (a.M + 16 * PB - 1) / (16 * PB);                                  \
    const dim3 grid((unsigned)std::max<int64_t>(                                            \
        1, std::min<int64_t>((tiles + kGemmWaves - 1) / kGemmWaves, cu_count())));          \
    allow_lds(mlp_gemm_kernel<KS, PB, PRO, E>, lds);                                        \
    hipLaunchKernelGGL((mlp_gemm_kernel<KS, PB, PRO, E>), grid, block, lds, st, a);         \
    break;                                                                                  \
  }
    GERAF_EPI(EPI_BIAS)
    GERAF_EPI(EPI_SOFTPLUS)
    GERAF_EPI(EPI_SOFTPLUS_TRAIN)
    GERAF_EPI(EPI_RELU)
    GERAF_EPI(EPI_GRAD_ADJ)
    GERAF_EPI(EPI_LAYER_ADJ)
    GERAF_EPI(EPI_MUL_SLOPE)
    GERAF_EPI(EPI_GRAD_ADJ_S)
    GERAF_EPI(EPI_GRAD_ADJ_SU)
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

template <int PG, int PX, int STEP, bool EXTRA>
int launch_wgrad(const WgradArgs& b, int blocks, hipStream_t st) {
  constexpr int S = STEP + 8;
  const size_t lds = (size_t)((EXTRA ? kWgRows : kWgTile) + kWgTile) * S * sizeof(uint16_t);
  allow_lds(mlp_wgrad_kernel<PG, PX, STEP, EXTRA>, lds);
  hipLaunchKernelGGL((mlp_wgrad_kernel<PG, PX, STEP, EXTRA>), dim3((unsigned)blocks),
                     dim3(64 * kWgWaves), lds, st, b);
  return check_launch("mlp_wgrad_kernel");
}

size_t chain_lds_bytes() {
  return (size_t)2 * kChainFrags * sizeof(bf8) + (size_t)(8 * 256 + 260) * sizeof(float) +
         (size_t)kChainPts * kGamStride * sizeof(uint16_t);
}

}  // namespace
}  // namespace geraf

using namespace geraf;

// layer 0 takes two steps of 32 inputs: 33 <= 3 + 6 L <= 64
extern "C" int64_t geraf_mlp_sdf_chain_pack_bytes(int32_t L) {
  if (L < 5 || L > 10) return -1;
  const int n3 = 256 - (3 + 6 * L);
  return (int64_t)(16 * 2 + 16 * 8 * 6 + nb_of(n3) * 8) * kWave * 16;
}

extern "C" int geraf_mlp_sdf_chain_bf16(const float* x, int64_t M, int32_t L, float scale,
                                        const void* wpacked, const float* bias,
                                        const float* w_sdf, float* sdf,
                                        uint16_t* h, int64_t ldh, int64_t layer_stride_h,
                                        uint16_t* q, int64_t ldq, int64_t layer_stride_q,
                                        float beta, float threshold, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX) throw ArgError("M out of range");
  if (L < 5 || L > 10) throw ArgError("L must be in [5, 10]");
  if (!(beta > 0.0f)) throw ArgError("beta must be positive");
  if (!(threshold >= 15.0f)) throw ArgError("threshold must be >= 15 (torch's default is 20)");
  const bool train = h != nullptr || q != nullptr;
  if (train && (h == nullptr || q == nullptr)) throw ArgError("h and q go together");
  if (!train && sdf == nullptr) throw ArgError("nothing to compute (sdf, h and q are null)");
  if (sdf != nullptr && w_sdf == nullptr) throw ArgError("sdf needs w_sdf");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, wpacked, bias}, true);
  if ((reinterpret_cast<uintptr_t>(wpacked) & 15u) != 0) throw ArgError("wpacked must be 16-byte aligned");
  if (train) {
    require_ptrs({h, q}, true);
    if (((reinterpret_cast<uintptr_t>(h) | reinterpret_cast<uintptr_t>(q)) & 15u) != 0)
      throw ArgError("h and q must be 16-byte aligned");
    if (ldh < 256 || ldh % 8 != 0 || ldq < 256 || ldq % 8 != 0)
      throw ArgError("ldh and ldq must be multiples of 8 and >= 256");
    if (layer_stride_h % 8 != 0 || layer_stride_q % 8 != 0 || layer_stride_h < M * ldh ||
        layer_stride_q < M * ldq)
      throw ArgError("layer strides must be multiples of 8 covering M rows");
  }
  ChainArgs a;
  a.x = x;
  a.w = reinterpret_cast<const bf8*>(wpacked);
  a.bias = bias;
  a.h = h;
  a.q = q;
  a.sh = layer_stride_h;
  a.sq = layer_stride_q;
  a.w_sdf = w_sdf;
  a.sdf = sdf;
  a.ldh = ldh;
  a.ldq = ldq;
  a.scale = scale;
  a.beta = beta;
  a.inv_beta = 1.0f / beta;
  a.M = (int)M;
  a.L = L;
  a.din = 3 + 6 * L;
  a.n3 = 256 - a.din;
  a.nb3 = nb_of(a.n3);
  const size_t lds = chain_lds_bytes();
  const int64_t tiles = (M + kChainPts - 1) / kChainPts;
  const dim3 grid((unsigned)std::min<int64_t>(tiles, cu_count()));
  const hipStream_t st = (hipStream_t)stream;
  if (train) {
    allow_lds(mlp_sdf_chain_kernel<true>, lds);
    hipLaunchKernelGGL(mlp_sdf_chain_kernel<true>, grid, dim3(64 * kChainWaves), lds, st, a);
  } else {
    allow_lds(mlp_sdf_chain_kernel<false>, lds);
    hipLaunchKernelGGL(mlp_sdf_chain_kernel<false>, grid, dim3(64 * kChainWaves), lds, st, a);
  }
  return check_launch("mlp_sdf_chain_kernel");
  GERAF_TRY_END
}

extern "C" int64_t geraf_mlp_pack_bytes(int64_t N, int64_t K) {
  if (N <= 0 || K <= 0 || N > kMaxN || K > kMaxK || gemm_lds_bytes(N, K) > kLdsBytes) return -1;
  return (int64_t)nb_of(N) * ks_of(K) * kWave * 16;
}

extern "C" int geraf_mlp_pack_weights(const float* w, int64_t ldw, int32_t transpose, float scale,
                                      int64_t N, int64_t K, void* packed, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (geraf_mlp_pack_bytes(N, K) < 0)
    throw ArgError("N must be in [1, 272] and K in [1, 288], N > 256 with K <= 256");
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
  if (geraf_mlp_pack_bytes(N, K) < 0)
    throw ArgError("N must be in [1, 272] and K in [1, 288], N > 256 with K <= 256");
  if (M == 0) return GERAF_RF_OK;
  if (prologue != PRO_NONE && prologue != PRO_MUL_SLOPE) throw ArgError("unknown prologue");
  if (epilogue < EPI_BIAS || epilogue > EPI_GRAD_ADJ_SU) throw ArgError("unknown epilogue");
  if (!(beta > 0.0f)) throw ArgError("beta must be positive");
  require_ptrs({x, wpacked, y}, true);
  check_ld(ldx, K, "x");
  check_ld(ldy, N, "y");
  if (ldy == 0) throw ArgError("y: leading dimension must not be 0");
  if (prologue == PRO_MUL_SLOPE) {
    require_ptrs({x2}, true);
    check_ld(ldx2, K, "x2");
  }
  if (two_out_epi(epilogue)) {
    require_ptrs({y2}, true);
    check_ld(ldy2, N, "y2");
    if (ldy2 == 0) throw ArgError("y2: leading dimension must not be 0");
  }
  if (adj_epi(epilogue)) {
    require_ptrs({e0}, true);
    check_ld(lde0, N, "e0");
  }
  if (e1_epi(epilogue)) {
    require_ptrs({e1}, true);
    check_ld(lde1, N, "e1");
  }
  const uintptr_t align = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(x2) |
                          reinterpret_cast<uintptr_t>(wpacked);
  const uintptr_t align2 = reinterpret_cast<uintptr_t>(y) | reinterpret_cast<uintptr_t>(y2) |
                           reinterpret_cast<uintptr_t>(e0) | reinterpret_cast<uintptr_t>(e1);
  if (((align | align2) & 15u) != 0)
    throw ArgError("x, x2, wpacked, y, y2, e0, e1 must be 16-byte aligned");
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
  a.inv_beta = 1.0f / beta;
  // the softplus form has no linear branch: it equals torch's above
  // thresholds where exp(-threshold) is below BF16's and the code's resolution
  if (!(threshold >= 15.0f)) throw ArgError("threshold must be >= 15 (torch's default is 20)");
  const size_t lds = gemm_lds_bytes(N, K);
  const hipStream_t st = (hipStream_t)stream;
  if (prologue == PRO_MUL_SLOPE) return launch_gemm<PRO_MUL_SLOPE>(epilogue, a, lds, st);
  return launch_gemm<PRO_NONE>(epilogue, a, lds, st);
  GERAF_TRY_END
}

extern "C" int geraf_mlp_chain_adjoint_bf16(const uint16_t* x, int64_t ldx, const uint16_t* x2,
                                           int64_t ldx2, const float* x_scale, int32_t mask_kind,
                                           const void* wpacked, int32_t steps,
                                           int32_t mask_steps, const uint16_t* e0, int64_t lde0,
                                           int64_t step_e0, const uint16_t* e1, int64_t lde1,
                                           int64_t step_e1, uint16_t* y, int64_t ldy,
                                           int64_t step_y, uint16_t* x_out, int64_t ldx_out,
                                           int64_t M, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX) throw ArgError("M out of range");
  if (steps < 1 || steps > 64) throw ArgError("steps must be in [1, 64]");
  if (mask_steps < 0 || mask_steps > steps) throw ArgError("mask_steps must be in [0, steps]");
  if (mask_kind != 0 && mask_kind != 1) throw ArgError("mask_kind must be 0 (codes) or 1 (relu)");
  if (x_scale != nullptr && x2 == nullptr) throw ArgError("x_scale needs x2");
  if (x2 != nullptr && e1 != nullptr) throw ArgError("a prologue with an addend is not supported");
  if (mask_kind == 1 && e1 != nullptr) throw ArgError("relu masks with an addend are not supported");
  if (mask_kind == 1 && x2 == nullptr) throw ArgError("relu masks need the prologue (x2)");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, wpacked, y}, true);
  if (mask_steps > 0) {
    require_ptrs({e0}, true);
    check_ld(lde0, 256, "e0");
  }
  check_ld(ldx, 256, "x");
  check_ld(ldy, 256, "y");
  if (ldy == 0) throw ArgError("y: leading dimension must not be 0");
  if (x2 != nullptr) check_ld(ldx2, 256, "x2");
  if (e1 != nullptr) check_ld(lde1, 256, "e1");
  if (x_out != nullptr) {
    check_ld(ldx_out, 256, "x_out");
    if (ldx_out == 0) throw ArgError("x_out: leading dimension must not be 0");
  }
  if (((step_e0 | step_e1 | step_y) & 7) != 0) throw ArgError("step strides must be multiples of 8");
  const uintptr_t align = reinterpret_cast<uintptr_t>(x) | reinterpret_cast<uintptr_t>(x2) |
                          reinterpret_cast<uintptr_t>(wpacked) | reinterpret_cast<uintptr_t>(e0) |
                          reinterpret_cast<uintptr_t>(e1) | reinterpret_cast<uintptr_t>(y) |
                          reinterpret_cast<uintptr_t>(x_out);
  if ((align & 15u) != 0) throw ArgError("x, x2, wpacked, e0, e1, y, x_out must be 16-byte aligned");
  ChainAdjArgs a;
  a.x = x;
  a.x2 = x2;
  a.xs = x_scale;
  a.x_out = x_out;
  a.ldx = ldx;
  a.ldx2 = ldx2;
  a.ldxo = ldx_out;
  a.w = reinterpret_cast<const bf8*>(wpacked);
  a.e0 = e0;
  a.e1 = e1;
  a.y = y;
  a.lde0 = lde0;
  a.se0 = step_e0;
  a.lde1 = lde1;
  a.se1 = step_e1;
  a.ldy = ldy;
  a.sy = step_y;
  a.M = (int)M;
  a.steps = steps;
  a.mask_steps = mask_steps;
  const size_t lds = (size_t)2 * kChainFrags * sizeof(bf8);
  const int64_t tiles = (M + kChainPts - 1) / kChainPts;
  const dim3 grid((unsigned)std::min<int64_t>(tiles, cu_count()));
  const dim3 block(64 * kChainWaves);
  const hipStream_t st = (hipStream_t)stream;
#define GERAF_SWEEP(P, A, R)                                                   \
  {                                                                            \
    allow_lds(mlp_chain_adj_kernel<P, A, R>, lds);                             \
    hipLaunchKernelGGL((mlp_chain_adj_kernel<P, A, R>), grid, block, lds, st, a); \
  }
  if (mask_kind == 1) {
    if (x_scale != nullptr)
      GERAF_SWEEP(2, false, true)
    else
      GERAF_SWEEP(1, false, true)
  } else if (x2 != nullptr) {
    if (x_scale != nullptr)
      GERAF_SWEEP(2, false, false)
    else
      GERAF_SWEEP(1, false, false)
  } else if (e1 != nullptr) {
    GERAF_SWEEP(0, true, false)
  } else {
    GERAF_SWEEP(0, false, false)
  }
#undef GERAF_SWEEP
  return check_launch("mlp_chain_adj_kernel");
  GERAF_TRY_END
}

extern "C" int64_t geraf_mlp_wgrad_workspace_bytes(int64_t M, int64_t N, int64_t K) {
  if (M < 0 || N <= 0 || K <= 0) return -1;
  return (int64_t)kWgMaxBlocks * kWgRows * kWgCols * (int64_t)sizeof(float);
}

extern "C" int geraf_mlp_wgrad_bf16(const uint16_t* g, int64_t ldg, const uint16_t* g2,
                                    int64_t ldg2, int32_t prologue_g, const uint16_t* x,
                                    int64_t ldx, const uint16_t* x2, int64_t ldx2,
                                    int32_t prologue_x, float* dw, int64_t lddw, float* db,
                                    int32_t accumulate, int64_t M, int64_t N, int64_t K,
                                    float beta, float threshold, void* workspace,
                                    int64_t workspace_bytes, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  (void)beta;
  (void)threshold;
  if (M < 0 || M > INT32_MAX || N <= 0 || K <= 0 || N > 512 || K > 512)
    throw ArgError("sizes out of range (N, K in [1, 512])");
  if (prologue_g != PRO_NONE && prologue_g != PRO_MUL_SLOPE) throw ArgError("unknown prologue_g");
  if (prologue_x != PRO_NONE && prologue_x != PRO_MUL_SLOPE) throw ArgError("unknown prologue_x");
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
  // row super-tiles: one of up to 272 rows, else blocks of 256
  const int nsup = N <= kWgRows ? 1 : (int)((N + kWgTile - 1) / kWgTile);
  const int ksup = (int)((K + kWgTile - 1) / kWgTile);
  // point chunks: a multiple of 64 points each, at most kWgMaxBlocks chunks
  const int cus = cu_count();
  int64_t per = (M + cus - 1) / std::max(1, cus);
  per = std::max<int64_t>(64, ((per + 63) / 64) * 64);
  if ((M + per - 1) / per > kWgMaxBlocks) per = ((M + kWgMaxBlocks - 1) / kWgMaxBlocks + 63) / 64 * 64;
  const int blocks = M == 0 ? 1 : (int)((M + per - 1) / per);
  float* partial = reinterpret_cast<float*>(workspace);
  for (int ns = 0; ns < nsup; ++ns) {
    const int64_t n0 = (int64_t)ns * kWgTile;
    const int rows = (int)std::min<int64_t>(nsup == 1 ? N : kWgTile, N - n0);
    for (int kz = 0; kz < ksup; ++kz) {
      WgradArgs b;
      // the super-tile's origin as column offsets of the operands
      b.g = g ? g + (size_t)n0 : nullptr;
      b.g2 = g2 ? g2 + (size_t)n0 : nullptr;
      b.x = x ? x + (size_t)kz * kWgTile : nullptr;
      b.x2 = x2 ? x2 + (size_t)kz * kWgTile : nullptr;
      b.partial = partial;
      b.ldg = ldg;
      b.ldg2 = ldg2;
      b.ldx = ldx;
      b.ldx2 = ldx2;
      b.M = (int)M;
      b.N = rows;
      b.K = (int)std::min<int64_t>(kWgTile, K - (int64_t)kz * kWgTile);
      b.with_bias = (db != nullptr && kz == 0);
      b.pts_per_block = (int)per;
      if (M > 0) {
        const bool pg = prologue_g == PRO_MUL_SLOPE, px = prologue_x == PRO_MUL_SLOPE;
        int rc;
        if (rows > kWgTile) {
          if (pg || px) throw ArgError("a prologue needs N <= 256 (or N > 272)");
          rc = launch_wgrad<PRO_NONE, PRO_NONE, 32, true>(b, blocks, st);
        } else if (pg && px) {
          rc = launch_wgrad<PRO_MUL_SLOPE, PRO_MUL_SLOPE, 32, false>(b, blocks, st);
        } else if (pg) {
          rc = launch_wgrad<PRO_MUL_SLOPE, PRO_NONE, 32, false>(b, blocks, st);
        } else if (px) {
          rc = launch_wgrad<PRO_NONE, PRO_MUL_SLOPE, 32, false>(b, blocks, st);
        } else {
          rc = launch_wgrad<PRO_NONE, PRO_NONE, 64, false>(b, blocks, st);
        }
        if (rc != GERAF_RF_OK) return rc;
      }
      const int cells = rows * kWgCols;
      hipLaunchKernelGGL(mlp_wgrad_reduce_kernel, dim3((cells + kWave - 1) / kWave),
                         dim3(64 * kRedWaves), 0, st,
                         partial, M == 0 ? 0 : blocks, rows, (int)n0, kz * kWgTile, (int)N,
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
  if (M < 0 || M > INT32_MAX || L < 0 || 3 + 6 * L > kPeMaxCols)
    throw ArgError("sizes out of range (L in [0, 10])");
  if (col0 < 0 || pad_to < 3 + 6 * L || pad_to > kPeMaxCols || col0 + pad_to > ldo)
    throw ArgError("columns out of range (pad_to in [3 + 6 L, 64])");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, out}, true);
  hipLaunchKernelGGL(pe_kernel, dim3((unsigned)((M + kPePts - 1) / kPePts)), dim3(kPePts),
                     (size_t)kPePts * (pad_to | 1) * sizeof(uint16_t), (hipStream_t)stream, x,
                     (int)M, (int)L, scale, out, ldo, (int)col0, (int)pad_to);
  return check_launch("pe_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mlp_pe_gradient(const float* x, int64_t M, int32_t L, float scale,
                                     const uint16_t* ga, int64_t lda, const uint16_t* gb,
                                     int64_t ldb, float* n, geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX || L < 0 || 3 + 6 * L > kPeMaxCols)
    throw ArgError("sizes out of range (L in [0, 10])");
  if (lda < 3 + 6 * L || ldb < 3 + 6 * L) throw ArgError("leading dimensions too small");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, ga, gb, n}, true);
  const size_t lds = (size_t)2 * kPePts * ((3 + 6 * L) | 1) * sizeof(uint16_t);
  allow_lds(pe_grad_kernel, lds);
  hipLaunchKernelGGL(pe_grad_kernel, dim3((unsigned)((M + kPePts - 1) / kPePts)), dim3(kPePts),
                     lds, (hipStream_t)stream, x, (int)M, (int)L, scale, ga, lda, gb, ldb, n);
  return check_launch("pe_grad_kernel");
  GERAF_TRY_END
}

extern "C" int geraf_mlp_pe_gradient_adjoint(const float* x, int64_t M, int32_t L, float scale,
                                             const float* dn, uint16_t* out, int64_t ldo,
                                             int32_t col0, int32_t pad_to,
                                             geraf_stream_t stream) {
  GERAF_TRY_BEGIN
  if (M < 0 || M > INT32_MAX || L < 0 || 3 + 6 * L > kPeMaxCols)
    throw ArgError("sizes out of range (L in [0, 10])");
  if (col0 < 0 || pad_to < 3 + 6 * L || pad_to > kPeMaxCols || col0 + pad_to > ldo)
    throw ArgError("columns out of range (pad_to in [3 + 6 L, 64])");
  if (M == 0) return GERAF_RF_OK;
  require_ptrs({x, dn, out}, true);
  hipLaunchKernelGGL(pe_grad_adj_kernel, dim3((unsigned)((M + kPePts - 1) / kPePts)),
                     dim3(kPePts), (size_t)kPePts * (pad_to | 1) * sizeof(uint16_t),
                     (hipStream_t)stream, x, (int)M, (int)L, scale, dn, out, ldo, (int)col0,
                     (int)pad_to);
  return check_launch("pe_grad_adj_kernel");
  GERAF_TRY_END
}
