// BF16 MFMA kernels of the per-edge closure MLP (gfx950).
//
// Matrix core: v_mfma_f32_16x16x32_bf16 (__builtin_amdgcn_mfma_f32_16x16x32_bf16,
// gfx950's double-K form) on wave64.  Operands per lane l (0..63), with
// q = l / 16, r = l % 16:
//   A (16 x 32, m x k):  A[r][8q + 0..7]
//   B (32 x 16, k x n):  B[8q + 0..7][r]
//   D (16 x 16 FP32):    D[4q + 0..3][r]
// (A and B are filled with the same k mapping, so the sum runs over the same
// 32 k whatever order the unit takes them in.)
//
// Every product is formed transposed, D' = W^T H^T: the weights are the A
// operand, the activation rows the B operand, so a lane's four results are
// four consecutive features of one row and go to HBM as one 8-byte store.
// Weights are repacked once per update as BF16 "packs" in the A layout (one
// 16-byte fragment per lane per MFMA): P(W) for the forward pass (m over a
// layer's outputs, k over its inputs) and Q(W) for the backward pass (m over
// inputs, k over outputs).  A workgroup (one per CU) copies the four packs
// it needs (84 KB) into LDS once and then loops over row tiles, one per wave
// at a time.  The forward pass keeps a tile's activations in registers
// between layers (two lane swaps per dword reorder a layer's output into the
// next layer's B operand; see k_mlp_fwd); the backward pass keeps them in
// the wave's own LDS tile.
//
// Layers: Z1 = X W1 + b1, H1 = GELU(Z1); Z2 = H1 W2 + b2, H2 = GELU(Z2);
// Z3 = H2 W3 + b3, H3 = GELU(Z3); Y = H3 W4 + b4.  GELU is the erf form of
// Hendrycks & Gimpel (2016), x * Phi(x).  The forward pass reads it from a
// table in LDS, the chord interpolant on 1024 intervals of [-6, 6] (error
// <= 1.37e-5); GELU' of the training forward and the table's nodes use erf
// from Abramowitz & Stegun (1964) eq. 7.1.26 (|error| <= 1.5e-7).  Both are
// far below the BF16 rounding every activation then takes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "mlp_dev.h"
#include "tmcore_gpu.h"

namespace tmk {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned short u16x4_t __attribute__((ext_vector_type(4)));

constexpr int kSub = 1;                      // 16-row MFMA column blocks per wave tile
constexpr int kTileRows = 16 * kSub;         // rows per wave tile
// waves per workgroup (one workgroup per CU): 15, so that the packs, the
// waves' activation tiles, the biases and the GELU table fit the 160 KB LDS
constexpr int kMlpWaves = 15;
static_assert(kSub == 1, "the LDS budget above is for 16-row tiles");
constexpr int kMlpBlock = 64 * kMlpWaves;
constexpr int kGradBlock = 512;              // k_mlp_grad: 8 waves, one 16-feature slice each
constexpr int kActStride = kMlpH + 8;        // bf16 per LDS activation row (conflict-free 16-byte reads)
constexpr int kIn1 = 64;                     // K of layer 1: 40 inputs padded to two 32-steps

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
// Two floats to BF16 (round to nearest even, as f2bf) in one instruction
// (v_cvt_pk_bf16_f32), a in the low half.
__device__ __forceinline__ uint32_t pk_bf16(float a, float b) {
  const bf16x2_t v = {(__bf16)a, (__bf16)b};
  return __builtin_bit_cast(uint32_t, v);
}

// 1 + erf(x / sqrt(2)) = 2 Phi(x) and exp(-x^2 / 2).  A&S 7.1.26 for u >= 0:
// erf(u) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5)))) exp(-u^2), t = 1 /
// (1 + p u); for x < 0 the complement Q itself is 2 Phi(x), without
// cancellation.  Explicit fused multiply-adds (the files are built without
// contraction), the hardware reciprocal and 2^y.  Used for the nodes of the
// forward pass's GELU table and for GELU' of the training forward.
struct Phi2 {
  float two_phi, e;
};
__device__ __forceinline__ Phi2 phi2(float x) {
  const float u = fabsf(x) * 0.70710678118654752f;
  const float t = __builtin_amdgcn_rcpf(fmaf(0.3275911f, u, 1.0f));
  const float poly =
      t * fmaf(t, fmaf(t, fmaf(t, fmaf(t, 1.061405429f, -1.453152027f), 1.421413741f), -0.284496736f), 0.254829592f);
  const float e = __builtin_amdgcn_exp2f((x * x) * -0.72134752044448170f);  // exp(-x^2 / 2) = exp(-u^2)
  const float q = poly * e;
  return Phi2{x >= 0.f ? 2.0f - q : q, e};
}
__device__ __forceinline__ float gelu(float x) { return (0.5f * x) * phi2(x).two_phi; }

// GELU of the forward pass: the chord interpolant of GELU on kGeluN equal
// intervals of [-6, 6], each interval's line a + b z kept in LDS.  Its error
// is at most h^2 / 8 max|GELU''| = h^2 / 8 * 2 phi(0) = 1.37e-5 (h = 12 /
// 1024), below the BF16 rounding that follows; outside [-6, 6] the end
// intervals' lines continue (GELU(z) - z resp. GELU(z) are below 1e-8 in
// magnitude at |z| = 6).  Three VALU operations, a clamp and one LDS read
// instead of the reciprocal, 2^y and polynomial of phi2.
constexpr int kGeluN = 1024;
constexpr float kGeluLo = -6.0f, kGeluH = 12.0f / kGeluN, kGeluInvH = kGeluN / 12.0f;
__device__ __forceinline__ void gelu_table(float2 *tab, int tid, int nthreads) {
  for (int i = tid; i < kGeluN; i += nthreads) {
    const float x0 = fmaf((float)i, kGeluH, kGeluLo), x1 = x0 + kGeluH;
    const float g0 = gelu(x0), b = (gelu(x1) - g0) * kGeluInvH;
    tab[i] = make_float2(fmaf(-b, x0, g0), b);
  }
}
__device__ __forceinline__ float gelu_tab(const float2 *tab, float z) {
  // truncation toward zero: z < -6 gives index <= 0, NaN gives 0 (and a NaN result)
  int j = (int)fmaf(z, kGeluInvH, -kGeluLo * kGeluInvH);
  j = min(max(j, 0), kGeluN - 1);
  const float2 ab = tab[j];
  return fmaf(ab.y, z, ab.x);
}
// GELU'(x) = Phi(x) + x phi(x), phi(x) = exp(-x^2 / 2) / sqrt(2 pi)
__device__ __forceinline__ float gelu_d(float x) {
  const Phi2 p = phi2(x);
  return fmaf(x * 0.39894228040143268f, p.e, 0.5f * p.two_phi);
}

// Pack shapes: KS 32-deep k-steps x MT 16-wide m-tiles, 64 lane fragments of
// 8 bf16 each; element (ks, mt, l, i) holds W[k][m] (P) or W[m][k] (Q) with
// k = 32 ks + 8 (l / 16) + i and m = 16 mt + l % 16, zero outside the matrix.
struct PackDesc {
  const float *W;  // row-major [Kin][Nout]
  int Kin, Nout, KS, MT, q_form;
  uint16_t *out;
};

__global__ void k_mlp_pack(PackDesc d) {
  const int total = d.KS * d.MT * 64 * 8;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int i = idx & 7, lane = (idx >> 3) & 63, frag = idx >> 9;
  const int mt = frag % d.MT, ks = frag / d.MT;
  const int k = ks * 32 + 8 * (lane >> 4) + i, m = mt * 16 + (lane & 15);
  float v = 0.f;
  if (!d.q_form) {  // P: k over inputs, m over outputs
    if (k < d.Kin && m < d.Nout) v = d.W[(size_t)k * d.Nout + m];
  } else {  // Q: m over inputs, k over outputs
    if (m < d.Kin && k < d.Nout) v = d.W[(size_t)m * d.Nout + k];
  }
  d.out[idx] = f2bf(v);
}

// Fragment counts (16-byte lane fragments = KS * MT * 64) of the packs.
constexpr int kP1 = 2 * 8 * 64, kP2 = 4 * 8 * 64, kP4 = 4 * 1 * 64;  // P(W1), P(W2) = P(W3), P(W4)
constexpr int kQ1 = 4 * 3 * 64, kQ2 = 4 * 8 * 64, kQ4 = 1 * 8 * 64;  // Q(W1), Q(W2) = Q(W3), Q(W4)
constexpr int kPackFrags = kP1 + 2 * kP2 + kP4;                         // per pass (P or Q): 5376 = 84 KB
static_assert(kQ1 + 2 * kQ2 + kQ4 == kPackFrags, "P and Q packs take the same LDS");

struct MlpPtrs {
  const bf16x8_t *p1, *p2, *p3, *p4, *q1, *q2, *q3, *q4;
  const float *params;
};

// Copies a pass's four packs (laid out back to back, kPackFrags fragments)
// into LDS with NT threads: every load is issued before the first store, so
// the ~6 loads of a thread wait once, not one after another.
template <int NT>
__device__ __forceinline__ void stage_packs(bf16x8_t *lds, const bf16x8_t *g) {
  constexpr int kPer = (kPackFrags + NT - 1) / NT;
  bf16x8_t v[kPer];
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < kPackFrags) v[j] = g[i];
  }
#pragma unroll
  for (int j = 0; j < kPer; ++j) {
    const int i = threadIdx.x + j * NT;
    if (i < kPackFrags) lds[i] = v[j];
  }
}

constexpr int kNBias = 3 * kMlpH + kMlpOut;  // b1, b2, b3, b4 back to back in LDS

// acc[s][mt] = pack(ks, mt) x act^T over the wave's 16-row blocks s.
template <int KS, int MT>
__device__ __forceinline__ void wave_mma(const bf16x8_t *pack, const uint16_t (*act)[kActStride],
                                         f32x4_t (&acc)[kSub][MT], int lane) {
#pragma unroll
  for (int s = 0; s < kSub; ++s)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[s][mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const int col = ks * 32 + 8 * (lane >> 4);
    bf16x8_t b[kSub];
#pragma unroll
    for (int s = 0; s < kSub; ++s) b[s] = *reinterpret_cast<const bf16x8_t *>(&act[16 * s + (lane & 15)][col]);
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const bf16x8_t a = pack[(ks * MT + mt) * 64 + lane];
#pragma unroll
      for (int s = 0; s < kSub; ++s) acc[s][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[s], acc[s][mt], 0, 0, 0);
    }
  }
}

// ---------------------------------------------------------------- forward
// Forward pass: each wave takes 32-row tiles (two 16-row MFMA column blocks,
// so every weight fragment read from LDS feeds two MFMAs) and keeps the
// activations in registers between layers.  A layer's result in lane (q, r)
// is features 4q + 0..3 of m-tile mt for row r; the next layer's B operand
// for k-step ks needs features 32 ks + 8q + 0..7 of row r, i.e. m-tile
// 2ks + q/2, features 8(q%2) + 0..7, which lanes (2(q%2), r) and
// (2(q%2) + 1, r) hold.  v_permlane32_swap then v_permlane16_swap on the
// packed m-tiles 2ks and 2ks+1 move exactly those: the first operand ends as
// [2ks of rows 0, 2 | 2ks+1 of rows 0, 2], the second as [2ks of rows 1, 3 |
// 2ks+1 of rows 1, 3] (rows of 16 lanes), the low and high halves of the B
// fragment.  LDS holds only the P packs, the biases and the GELU table.
// 16-row column blocks per wave tile: two for inference (126 VGPRs, 16 waves
// per workgroup, one workgroup per CU); one for the training form, which
// also forms GELU' and stores H and GELU'(Z) (measured: 16 waves of 16-row
// tiles faster than 12 waves of 32-row tiles, which need 150 VGPRs)
constexpr int fwd_sub(bool train) { return train ? 1 : 2; }
constexpr int kFwdWaves = 16;
constexpr int kFwdBlock = 64 * kFwdWaves;

typedef uint32_t u32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x8_t bf16x8_of(const f32x4_t &lo, const f32x4_t &hi) {
  return __builtin_bit_cast(bf16x8_t, (u32x4_t{pk_bf16(lo[0], lo[1]), pk_bf16(lo[2], lo[3]), pk_bf16(hi[0], hi[1]),
                                               pk_bf16(hi[2], hi[3])}));
}

// Layer-1 B fragments of one 16-row block: lane (q, r) takes X[row][32 ks +
// 8q + 0..7] (zero past the 40 inputs and past the last row).
__device__ __forceinline__ void x_fragments(const float *X, int n, int row, int q, bf16x8_t *b) {
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const int c = 32 * ks + 8 * q;
    f32x4_t lo = f32x4_t{0.f, 0.f, 0.f, 0.f}, hi = lo;
    if (c < kMlpIn && row < n) {
      const f32x4_t *src = reinterpret_cast<const f32x4_t *>(X + (size_t)row * kMlpIn + c);
      lo = src[0];
      hi = src[1];
    }
    b[ks] = bf16x8_of(lo, hi);
  }
}

// acc[s][mt] = pack(ks, mt) x b[s][ks] summed over ks < KS.
template <int KS, int MT, int SUB>
__device__ __forceinline__ void wave_mma_r(const bf16x8_t *pack, const bf16x8_t (&b)[SUB][4],
                                           f32x4_t (&acc)[SUB][MT], int lane) {
#pragma unroll
  for (int s = 0; s < SUB; ++s)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) acc[s][mt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks)
#pragma unroll
    for (int mt = 0; mt < MT; ++mt) {
      const bf16x8_t a = pack[(ks * MT + mt) * 64 + lane];
#pragma unroll
      for (int s = 0; s < SUB; ++s)
        acc[s][mt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[s][ks], acc[s][mt], 0, 0, 0);
    }
}

// Hidden-layer epilogue: bias + GELU (the LDS table), BF16, into the next
// layer's B fragments b; when training, H and GELU'(Z) (A&S) to HBM (BF16
// [n][128]).
template <bool TRAIN, int SUB = fwd_sub(TRAIN)>
__device__ __forceinline__ void fwd_epilogue(const f32x4_t (&acc)[SUB][8], const float *bias, const float2 *gtab,
                                             int lane, int row0, int n, uint16_t *H, uint16_t *D,
                                             bf16x8_t (&b)[SUB][4]) {
  const int q = lane >> 4, r = lane & 15;
  uint32_t t[SUB][8][2];
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    const int f = mt * 16 + 4 * q;
    const f32x4_t bv = *reinterpret_cast<const f32x4_t *>(bias + f);
#pragma unroll
    for (int s = 0; s < SUB; ++s) {
      float g[4], gd[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const float z = acc[s][mt][i] + bv[i];
        g[i] = gelu_tab(gtab, z);
        if constexpr (TRAIN) gd[i] = gelu_d(z);
      }
      t[s][mt][0] = pk_bf16(g[0], g[1]);
      t[s][mt][1] = pk_bf16(g[2], g[3]);
      const int row = row0 + 16 * s + r;
      if (TRAIN && row < n) {
        *reinterpret_cast<uint2 *>(&H[(size_t)row * kMlpH + f]) = make_uint2(t[s][mt][0], t[s][mt][1]);
        *reinterpret_cast<uint2 *>(&D[(size_t)row * kMlpH + f]) =
            make_uint2(pk_bf16(gd[0], gd[1]), pk_bf16(gd[2], gd[3]));
      }
    }
  }
#pragma unroll
  for (int s = 0; s < SUB; ++s)
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      uint32_t w[4];
#pragma unroll
      for (int d = 0; d < 2; ++d) {
        const auto p32 = __builtin_amdgcn_permlane32_swap(t[s][2 * ks][d], t[s][2 * ks + 1][d], false, false);
        const auto p16 = __builtin_amdgcn_permlane16_swap(p32[0], p32[1], false, false);
        w[d] = p16[0];
        w[2 + d] = p16[1];
      }
      b[s][ks] = __builtin_bit_cast(bf16x8_t, (u32x4_t{w[0], w[1], w[2], w[3]}));
    }
}

template <bool TRAIN>
__global__ __launch_bounds__(kFwdBlock) void k_mlp_fwd(MlpPtrs p, const float *__restrict__ X, int n,
                                                       float *__restrict__ Y, uint16_t *H, uint16_t *D) {
  constexpr int kFwdSub = fwd_sub(TRAIN), kFwdRows = 16 * kFwdSub;
  __shared__ bf16x8_t packs[kPackFrags];
  __shared__ __attribute__((aligned(16))) float bias[kNBias];
  __shared__ float2 gtab[kGeluN];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  stage_packs<kFwdBlock>(packs, p.p1);
  gelu_table(gtab, threadIdx.x, kFwdBlock);
  for (int i = threadIdx.x; i < kNBias; i += kFwdBlock) {
    const int l = i / kMlpH;
    bias[i] = p.params[(l == 0 ? kOB1 : l == 1 ? kOB2 : l == 2 ? kOB3 : kOB4) + i % kMlpH];
  }
  __syncthreads();
  const bf16x8_t *P1 = packs, *P2 = P1 + kP1, *P3 = P2 + kP2, *P4 = P3 + kP2;
  const size_t plane = (size_t)n * kMlpH;
  const int ntiles = (n + kFwdRows - 1) / kFwdRows;
  for (int tile = blockIdx.x * kFwdWaves + wave; tile < ntiles; tile += gridDim.x * kFwdWaves) {
    // the weight fragments are loop-invariant LDS reads: without this
    // compiler barrier they are hoisted out of the tile loop (84 fragments,
    // 336 VGPRs) and spill; read per tile, they stream through a few VGPRs
    __asm__ volatile("" ::: "memory");
    const int row0 = tile * kFwdRows;
    bf16x8_t b[kFwdSub][4];
#pragma unroll
    for (int s = 0; s < kFwdSub; ++s) x_fragments(X, n, row0 + 16 * s + (lane & 15), lane >> 4, b[s]);
    f32x4_t acc[kFwdSub][8];
    wave_mma_r<2, 8>(P1, b, acc, lane);
    fwd_epilogue<TRAIN>(acc, bias, gtab, lane, row0, n, H, D, b);
    wave_mma_r<4, 8>(P2, b, acc, lane);
    fwd_epilogue<TRAIN>(acc, bias + kMlpH, gtab, lane, row0, n, H + plane, D + plane, b);
    wave_mma_r<4, 8>(P3, b, acc, lane);
    fwd_epilogue<TRAIN>(acc, bias + 2 * kMlpH, gtab, lane, row0, n, H + 2 * plane, D + 2 * plane, b);
    f32x4_t out[kFwdSub][1];
    wave_mma_r<4, 1>(P4, b, out, lane);
    if (lane < 16) {  // outputs 0 and 1 are features 4q + 0, 1 of q = 0
#pragma unroll
      for (int s = 0; s < kFwdSub; ++s) {
        const int row = row0 + s * 16 + lane;
        if (row < n) {
          float2 y;
          y.x = out[s][0][0] + bias[3 * kMlpH];
          y.y = out[s][0][1] + bias[3 * kMlpH + 1];
          *reinterpret_cast<float2 *>(&Y[(size_t)row * kMlpOut]) = y;
        }
      }
    }
  }
}

// dZ of a hidden layer = (upstream x W^T) * GELU'(Z) (GELU' saved by the
// training forward), BF16, into the wave's LDS tile and to HBM for the
// weight gradients.
__device__ __forceinline__ void dz_epilogue(const f32x4_t (&acc)[kSub][8], const uint16_t *D, uint16_t (*act)[kActStride],
                                            int lane, int row0, int n, uint16_t *dZ) {
  const int q = lane >> 4, r = lane & 15;
#pragma unroll
  for (int mt = 0; mt < 8; ++mt) {
    const int f = mt * 16 + 4 * q;
#pragma unroll
    for (int s = 0; s < kSub; ++s) {
      const int row = s * 16 + r;
      const bool live = row0 + row < n;
      u16x4_t d = {0, 0, 0, 0};
      if (live) d = *reinterpret_cast<const u16x4_t *>(&D[(size_t)(row0 + row) * kMlpH + f]);
      float v[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] = acc[s][mt][i] * bf2f(d[i]);  // 0 on dead rows (D = 0)
      const uint2 z = make_uint2(pk_bf16(v[0], v[1]), pk_bf16(v[2], v[3]));
      *reinterpret_cast<uint2 *>(&act[row][f]) = z;
      if (live) *reinterpret_cast<uint2 *>(&dZ[(size_t)(row0 + row) * kMlpH + f]) = z;
    }
  }
}

__global__ __launch_bounds__(kMlpBlock) void k_mlp_bwd(MlpPtrs p, const float *__restrict__ dY, int n,
                                                       const uint16_t *D, uint16_t *dZ, float *__restrict__ dX) {
  __shared__ bf16x8_t packs[kPackFrags];
  __shared__ __attribute__((aligned(16))) uint16_t act[kMlpWaves][kTileRows][kActStride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  stage_packs<kMlpBlock>(packs, p.q1);
  __syncthreads();
  const bf16x8_t *Q1 = packs, *Q2 = Q1 + kQ1, *Q3 = Q2 + kQ2, *Q4 = Q3 + kQ2;
  uint16_t(*A)[kActStride] = act[wave];
  const size_t plane = (size_t)n * kMlpH;
  const int ntiles = (n + kTileRows - 1) / kTileRows;
  for (int tile = blockIdx.x * kMlpWaves + wave; tile < ntiles; tile += gridDim.x * kMlpWaves) {
    const int row0 = tile * kTileRows;
    // dY (FP32, 2 per row) -> BF16 tile columns 0, 1, zeros to 31
    for (int i = lane; i < kTileRows * 4; i += 64) {
      const int r = i / 4, c = 8 * (i % 4);
      u16x4_t lo = {0, 0, 0, 0};
      if (c == 0 && row0 + r < n) {
        const float2 y = *reinterpret_cast<const float2 *>(&dY[(size_t)(row0 + r) * kMlpOut]);
        lo[0] = f2bf(y.x);
        lo[1] = f2bf(y.y);
      }
      *reinterpret_cast<u16x4_t *>(&A[r][c]) = lo;
      *reinterpret_cast<u16x4_t *>(&A[r][c + 4]) = u16x4_t{0, 0, 0, 0};
    }
    f32x4_t acc[kSub][8];
    wave_mma<1, 8>(Q4, A, acc, lane);  // dH3 = dY W4^T
    dz_epilogue(acc, D + 2 * plane, A, lane, row0, n, dZ + 2 * plane);
    wave_mma<4, 8>(Q3, A, acc, lane);  // dH2 = dZ3 W3^T
    dz_epilogue(acc, D + plane, A, lane, row0, n, dZ + plane);
    wave_mma<4, 8>(Q2, A, acc, lane);  // dH1 = dZ2 W2^T
    dz_epilogue(acc, D, A, lane, row0, n, dZ);
    if (dX) {
      f32x4_t ax[kSub][3];
      wave_mma<4, 3>(Q1, A, ax, lane);  // dX = dZ1 W1^T
      const int q = lane >> 4, r = lane & 15;
#pragma unroll
      for (int mt = 0; mt < 3; ++mt) {
        const int f = mt * 16 + 4 * q;
#pragma unroll
        for (int s = 0; s < kSub; ++s) {
          const int row = row0 + s * 16 + r;
          if (f < kMlpIn && row < n) *reinterpret_cast<f32x4_t *>(&dX[(size_t)row * kMlpIn + f]) = ax[s][mt];
        }
      }
    }
  }
}

// Weight and bias gradients, deterministic: workgroup g sums its own fixed
// row range [g rps, (g + 1) rps) into part[g][kMlpNParams], and k_mlp_gsum
// adds the partials in workgroup order.  dW[k][j] = sum_rows A[row][k]
// dZ[row][j] is an MFMA with the rows as its k: 64-row chunks of every layer's
// input (X, H1, H2, H3) and output gradient (dZ1, dZ2, dZ3, dY) are staged in
// LDS transposed, feature-major, so both operands are 16-byte fragment reads.
// Wave w owns, for W2 and W3, the 16 inputs 16w..16w+15 (8 output tiles
// each), for W1 the outputs 16w..16w+15 (3 input tiles) and for W4 the
// inputs 16w..: 20 tiles, accumulated over the whole range in registers.
constexpr int kGradRows = 64;              // rows per staged chunk
constexpr int kTStride = kGradRows + 8;    // bf16 per transposed LDS row

struct GradArgs {
  const float *X;         // [n][40]
  const uint16_t *H;      // [3][n][128]
  const uint16_t *dZ;     // [3][n][128]
  const float *dY;        // [n][2]
  int n, rps;
  float *part;            // [gridDim.x][kMlpNParams]
};

// Rows [row0, row0 + 64) of a BF16 [n][128] array into T[128][kTStride]
// (feature-major); two rows per thread and item so that each LDS store is a
// 4-byte pair; rows past n are zero.
__device__ __forceinline__ void stage_t128(uint16_t (*T)[kTStride], const uint16_t *src, int row0, int n) {
  for (int it = threadIdx.x; it < (kGradRows / 2) * 16; it += kGradBlock) {
    const int pr = it % (kGradRows / 2), fc = it / (kGradRows / 2);
    const int ra = row0 + 2 * pr, rb = ra + 1;
    u16x4_t a0 = {0, 0, 0, 0}, a1 = a0, b0 = a0, b1 = a0;
    if (ra < n) {
      a0 = *reinterpret_cast<const u16x4_t *>(&src[(size_t)ra * kMlpH + 8 * fc]);
      a1 = *reinterpret_cast<const u16x4_t *>(&src[(size_t)ra * kMlpH + 8 * fc + 4]);
    }
    if (rb < n) {
      b0 = *reinterpret_cast<const u16x4_t *>(&src[(size_t)rb * kMlpH + 8 * fc]);
      b1 = *reinterpret_cast<const u16x4_t *>(&src[(size_t)rb * kMlpH + 8 * fc + 4]);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      *reinterpret_cast<uint32_t *>(&T[8 * fc + i][2 * pr]) = (uint32_t)a0[i] | ((uint32_t)b0[i] << 16);
      *reinterpret_cast<uint32_t *>(&T[8 * fc + 4 + i][2 * pr]) = (uint32_t)a1[i] | ((uint32_t)b1[i] << 16);
    }
  }
}

// acc[t] += A(16 x 64 rows, LDS feature-major at row a0) x B(64 rows x 16, at
// rows b0[t]) over the chunk's two 32-row k-steps.
template <int NT>
__device__ __forceinline__ void grad_mma(const uint16_t (*TA)[kTStride], int a0, const uint16_t (*TB)[kTStride],
                                         const int (&b0)[NT], f32x4_t (&acc)[NT], int lane) {
#pragma unroll
  for (int ks = 0; ks < kGradRows / 32; ++ks) {
    const int col = ks * 32 + 8 * (lane >> 4);
    const bf16x8_t a = *reinterpret_cast<const bf16x8_t *>(&TA[a0 + (lane & 15)][col]);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const bf16x8_t b = *reinterpret_cast<const bf16x8_t *>(&TB[b0[t] + (lane & 15)][col]);
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
}

// A 16 x 16 gradient tile (rows: inputs m0.., columns: outputs j0..) into
// the flat parameter layout W[m][j] at woff.
__device__ __forceinline__ void store_tile(float *part, int woff, int Kin, int Nout, int m0, int j0, const f32x4_t &t,
                                           int lane) {
  const int j = j0 + (lane & 15);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 4 * (lane >> 4) + i;
    if (m < Kin && j < Nout) part[woff + m * Nout + j] = t[i];
  }
}

__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(GradArgs g) {
  __shared__ __attribute__((aligned(16))) uint16_t TX[48][kTStride];      // X^T (40 inputs, padded)
  __shared__ __attribute__((aligned(16))) uint16_t TH[3][kMlpH][kTStride];  // H1^T, H2^T, H3^T
  __shared__ __attribute__((aligned(16))) uint16_t TZ[3][kMlpH][kTStride];  // dZ1^T, dZ2^T, dZ3^T
  __shared__ __attribute__((aligned(16))) uint16_t TY[16][kTStride];      // dY^T (2 outputs, padded)
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int r_begin = blockIdx.x * g.rps, r_end = min(g.n, r_begin + g.rps);
  const size_t plane = (size_t)g.n * kMlpH;
  f32x4_t a2[8], a3[8], a1[3], a4[1];
#pragma unroll
  for (int t = 0; t < 8; ++t) a2[t] = a3[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int t = 0; t < 3; ++t) a1[t] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  a4[0] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  // bias gradients: thread f < 386 sums one output feature over the rows
  const int f = threadIdx.x;
  float bsum = 0.f;
  for (int row0 = r_begin; row0 < r_end; row0 += kGradRows) {
    __syncthreads();  // the previous chunk's reads are done
    for (int l = 0; l < 3; ++l) {
      stage_t128(TH[l], g.H + l * plane, row0, r_end);
      stage_t128(TZ[l], g.dZ + l * plane, row0, r_end);
    }
    for (int it = threadIdx.x; it < kGradRows * 12; it += kGradBlock) {  // X: 64 rows x 12 float4 (40 + 8 zeros)
      const int r = it % kGradRows, c = 4 * (it / kGradRows);
      f32x4_t x = f32x4_t{0.f, 0.f, 0.f, 0.f};
      if (c < kMlpIn && row0 + r < r_end) x = *reinterpret_cast<const f32x4_t *>(&g.X[(size_t)(row0 + r) * kMlpIn + c]);
#pragma unroll
      for (int i = 0; i < 4; ++i) TX[c + i][r] = f2bf(x[i]);
    }
    for (int it = threadIdx.x; it < kGradRows * 16; it += kGradBlock) {  // dY: 64 rows x 16 (2 + 14 zeros)
      const int r = it % kGradRows, c = it / kGradRows;
      float y = 0.f;
      if (c < kMlpOut && row0 + r < r_end) y = g.dY[(size_t)(row0 + r) * kMlpOut + c];
      TY[c][r] = f2bf(y);
    }
    __syncthreads();
    {
      int b[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) b[t] = 16 * t;
      grad_mma<8>(TH[0], 16 * w, TZ[1], b, a2, lane);  // dW2 = H1^T dZ2, inputs 16w..
      grad_mma<8>(TH[1], 16 * w, TZ[2], b, a3, lane);  // dW3 = H2^T dZ3
      const int bx[1] = {0};
      grad_mma<1>(TH[2], 16 * w, TY, bx, a4, lane);  // dW4 = H3^T dY
    }
#pragma unroll
    for (int t = 0; t < 3; ++t) {  // dW1 = X^T dZ1, outputs 16w.. (input tiles t)
      const int b1[1] = {16 * w};
      f32x4_t one[1] = {a1[t]};
      grad_mma<1>(TX, 16 * t, TZ[0], b1, one, lane);
      a1[t] = one[0];
    }
    // bias sums in row order
    if (f < 3 * kMlpH + kMlpOut) {
      const uint16_t *z = f < 3 * kMlpH ? TZ[f / kMlpH][f % kMlpH] : TY[f - 3 * kMlpH];
      for (int r = 0; r < kGradRows; ++r) bsum += bf2f(z[r]);
    }
  }
  float *part = g.part + (size_t)blockIdx.x * kMlpNParams;
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    store_tile(part, kOW2, kMlpH, kMlpH, 16 * w, 16 * t, a2[t], lane);
    store_tile(part, kOW3, kMlpH, kMlpH, 16 * w, 16 * t, a3[t], lane);
  }
#pragma unroll
  for (int t = 0; t < 3; ++t) store_tile(part, kOW1, kMlpIn, kMlpH, 16 * t, 16 * w, a1[t], lane);
  store_tile(part, kOW4, kMlpH, kMlpOut, 16 * w, 0, a4[0], lane);
  if (f < 3 * kMlpH + kMlpOut) {
    const int boff = f < kMlpH ? kOB1 + f : f < 2 * kMlpH ? kOB2 + f - kMlpH : f < 3 * kMlpH ? kOB3 + f - 2 * kMlpH
                                                                                               : kOB4 + f - 3 * kMlpH;
    part[boff] = bsum;
  }
}

// grad[i] += sum over workgroups g = 0, 1, ... of part[g][i], in that order
// (the partials are loaded eight at a time ahead of the dependent adds).
__global__ void k_mlp_gsum(const float *part, int nparts, float *grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpNParams) return;
  float s = 0.f;
  for (int g0 = 0; g0 < nparts; g0 += 8) {
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] = g0 + k < nparts ? part[(size_t)(g0 + k) * kMlpNParams + i] : 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (g0 + k < nparts) s += v[k];
  }
  grad[i] += s;
}

__global__ void k_mlp_adam(float *p, float *g, float *m, float *v, int np, float lr, float b1, float b2, float eps,
                           float c1, float c2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  g[i] = 0.f;
}

__global__ void k_fill0(float *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// ------------------------------------------------------------------ host
namespace {
struct PackShape {
  int W_off, Kin, Nout, KS, MT, q_form;
};
// P(W1..W4) then Q(W1..W4), back to back (each pass's four packs contiguous,
// as k_mlp_fwd / k_mlp_bwd copy them into LDS in one sweep)
const PackShape kPacks[8] = {
    {kOW1, kMlpIn, kMlpH, 2, 8, 0},  {kOW2, kMlpH, kMlpH, 4, 8, 0},  {kOW3, kMlpH, kMlpH, 4, 8, 0},
    {kOW4, kMlpH, kMlpOut, 4, 1, 0}, {kOW1, kMlpIn, kMlpH, 4, 3, 1}, {kOW2, kMlpH, kMlpH, 4, 8, 1},
    {kOW3, kMlpH, kMlpH, 4, 8, 1},   {kOW4, kMlpH, kMlpOut, 1, 8, 1}};

MlpPtrs ptrs(const MlpState &s) {
  MlpPtrs p;
  const bf16x8_t *b = reinterpret_cast<const bf16x8_t *>(s.pack);
  const bf16x8_t **dst[8] = {&p.p1, &p.p2, &p.p3, &p.p4, &p.q1, &p.q2, &p.q3, &p.q4};
  for (int i = 0; i < 8; ++i) *dst[i] = b + s.pack_off[i] / 8;
  p.params = s.params;
  return p;
}

// Workgroups of the persistent MLP kernels: one per CU, fewer for small batches.
int mlp_blocks(const MlpState &s, int rows) {
  const int tiles = (rows + kTileRows - 1) / kTileRows;
  return std::max(1, std::min(s.ncu, (tiles + kMlpWaves - 1) / kMlpWaves));
}

int fwd_blocks(const MlpState &s, int rows, bool train) {
  const int tiles = (rows + 16 * fwd_sub(train) - 1) / (16 * fwd_sub(train));
  return std::max(1, std::min(s.ncu, (tiles + kFwdWaves - 1) / kFwdWaves));
}

// Row ranges of the gradient partials: one per CU, at least 64 rows each.
int grad_parts(const MlpState &s, int rows) { return std::max(1, std::min(s.ncu, (rows + 255) / 256)); }

int fail(MlpState &s, int code, const char *m) {
  s.err = m;
  return code;
}

void *salloc(MlpState &s, size_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
  s.allocs.push_back(p);
  return p;
}
}  // namespace

int mlp_reserve(MlpState &s, int rows) {
  if (rows <= s.cap) return TM_OK;
  // the old buffers may still be in use by queued work of any stream
  if (hipDeviceSynchronize() != hipSuccess) return fail(s, TM_EDEVICE, "hipDeviceSynchronize");
  for (void *q : s.batch_bufs) (void)hipFree(q);
  s.batch_bufs.clear();
  s.cap = 0;
  const int cap = ((rows + 63) / 64) * 64;
  const size_t plane = (size_t)cap * kMlpH;
  auto a = [&](size_t bytes) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(bytes, 16)) != hipSuccess) return (void *)nullptr;
    s.batch_bufs.push_back(q);
    return q;
  };
  if (!(s.X = (float *)a((size_t)cap * kMlpIn * 4)) || !(s.Y = (float *)a((size_t)cap * kMlpOut * 4)) ||
      !(s.dY = (float *)a((size_t)cap * kMlpOut * 4)) || !(s.dX = (float *)a((size_t)cap * kMlpIn * 4)) ||
      !(s.H = (uint16_t *)a(3 * plane * 2)) || !(s.D = (uint16_t *)a(3 * plane * 2)) ||
      !(s.dZ = (uint16_t *)a(3 * plane * 2)) ||
      !(s.gpart = (float *)a((size_t)grad_parts(s, cap) * kMlpNParams * 4)))
    return fail(s, TM_ENOMEM, "hipMalloc MLP batch buffers");
  s.cap = cap;
  return TM_OK;
}

int mlp_pack(MlpState &s, hipStream_t st) {
  for (int i = 0; i < 8; ++i) {
    const PackShape &sh = kPacks[i];
    PackDesc d{s.params + sh.W_off, sh.Kin, sh.Nout, sh.KS, sh.MT, sh.q_form, s.pack + s.pack_off[i]};
    const int total = sh.KS * sh.MT * 512;
    hipLaunchKernelGGL(k_mlp_pack, dim3((total + 255) / 256), dim3(256), 0, st, d);
  }
  return hipGetLastError() == hipSuccess ? TM_OK : fail(s, TM_EDEVICE, "k_mlp_pack launch");
}

int mlp_forward(MlpState &s, const float *X, int n, float *Y, bool train, hipStream_t st) {
  if (n <= 0) return TM_OK;
  if (train && n > s.cap) return fail(s, TM_EINVAL, "batch exceeds the reserved MLP batch buffers");
  if (train)
    hipLaunchKernelGGL(k_mlp_fwd<true>, dim3(fwd_blocks(s, n, true)), dim3(kFwdBlock), 0, st, ptrs(s), X, n, Y, s.H,
                       s.D);
  else
    hipLaunchKernelGGL(k_mlp_fwd<false>, dim3(fwd_blocks(s, n, false)), dim3(kFwdBlock), 0, st, ptrs(s), X, n, Y, s.H,
                       s.D);
  return hipGetLastError() == hipSuccess ? TM_OK : fail(s, TM_EDEVICE, "k_mlp_fwd launch");
}

int mlp_backward(MlpState &s, const float *X, const float *dY, int n, float *dX, hipStream_t st) {
  if (n <= 0) return TM_OK;
  if (n > s.cap) return fail(s, TM_EINVAL, "batch exceeds the reserved MLP batch buffers");
  hipLaunchKernelGGL(k_mlp_bwd, dim3(mlp_blocks(s, n)), dim3(kMlpBlock), 0, st, ptrs(s), dY, n, (const uint16_t *)s.D,
                     s.dZ, dX);
  const int parts = grad_parts(s, n);
  const int rps = ((n + parts - 1) / parts + kGradRows - 1) / kGradRows * kGradRows;
  GradArgs g{X, s.H, s.dZ, dY, n, rps, s.gpart};
  hipLaunchKernelGGL(k_mlp_grad, dim3((n + rps - 1) / rps), dim3(kGradBlock), 0, st, g);
  hipLaunchKernelGGL(k_mlp_gsum, dim3((kMlpNParams + 255) / 256), dim3(256), 0, st, (const float *)s.gpart,
                     (n + rps - 1) / rps, s.grad);
  return hipGetLastError() == hipSuccess ? TM_OK : fail(s, TM_EDEVICE, "MLP backward launch");
}

int mlp_zero_grad(MlpState &s, hipStream_t st) {
  hipLaunchKernelGGL(k_fill0, dim3((kMlpNParams + 255) / 256), dim3(256), 0, st, s.grad, kMlpNParams);
  return hipGetLastError() == hipSuccess ? TM_OK : fail(s, TM_EDEVICE, "k_fill0 launch");
}

int mlp_adam(MlpState &s, float lr, float b1, float b2, float eps, hipStream_t st) {
  s.adam_t += 1;
  const float c1 = 1.f - std::pow(b1, (float)s.adam_t), c2 = 1.f - std::pow(b2, (float)s.adam_t);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpNParams + 255) / 256), dim3(256), 0, st, s.params, s.grad, s.adam_m,
                     s.adam_v, kMlpNParams, lr, b1, b2, eps, c1, c2);
  if (hipGetLastError() != hipSuccess) return fail(s, TM_EDEVICE, "k_mlp_adam launch");
  return mlp_pack(s, st);
}

}  // namespace tmk

using namespace tmk;

struct tm_mlp {
  MlpState s;
};

#define MLPCHK(p, call)                                                                                    \
  do {                                                                                                     \
    hipError_t e_ = (call);                                                                                \
    if (e_ != hipSuccess) {                                                                                \
      (p)->s.err = std::string(#call) + ": " + hipGetErrorString(e_);                                     \
      return TM_EDEVICE;                                                                                   \
    }                                                                                                      \
  } while (0)

extern "C" int tm_mlp_create(tm_mlp **out, int device, uint64_t seed) {
  if (!out) return TM_EINVAL;
  *out = nullptr;
  tm_mlp *p = new tm_mlp();
  MlpState &s = p->s;
  s.device = device;
  auto bail = [&](int rc) {
    tm_mlp_destroy(p);
    return rc;
  };
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess ||
      hipDeviceGetAttribute(&s.ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || s.ncu < 1)
    return bail(TM_EDEVICE);
  size_t off = 0;
  for (int i = 0; i < 8; ++i) {
    s.pack_off[i] = off;
    off += (size_t)kPacks[i].KS * kPacks[i].MT * 512;
  }
  if (!(s.params = (float *)salloc(s, kMlpNParams * 4)) || !(s.grad = (float *)salloc(s, kMlpNParams * 4)) ||
      !(s.adam_m = (float *)salloc(s, kMlpNParams * 4)) || !(s.adam_v = (float *)salloc(s, kMlpNParams * 4)) ||
      !(s.pack = (uint16_t *)salloc(s, off * 2)) || !(s.norm = (float *)salloc(s, kNormLen * 4)))
    return bail(TM_ENOMEM);
  // He-uniform initialisation from the seed (host), zero biases
  std::vector<float> h(kMlpNParams, 0.f);
  std::mt19937_64 rng(seed);
  auto init = [&](int off_w, int fan_in, int count) {
    const float a = std::sqrt(6.0f / fan_in);
    std::uniform_real_distribution<float> U(-a, a);
    for (int i = 0; i < count; ++i) h[off_w + i] = U(rng);
  };
  init(kOW1, kMlpIn, kMlpIn * kMlpH);
  init(kOW2, kMlpH, kMlpH * kMlpH);
  init(kOW3, kMlpH, kMlpH * kMlpH);
  init(kOW4, kMlpH, kMlpH * kMlpOut);
  for (int i = 0; i < kMlpOut * kMlpH; ++i) h[kOW4 + i] *= 0.1f;  // start near the plain method
  if (hipMemcpy(s.params, h.data(), kMlpNParams * 4, hipMemcpyHostToDevice) != hipSuccess) return bail(TM_EDEVICE);
  std::vector<float> z(kMlpNParams, 0.f);
  if (hipMemcpy(s.grad, z.data(), kMlpNParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s.adam_m, z.data(), kMlpNParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s.adam_v, z.data(), kMlpNParams * 4, hipMemcpyHostToDevice) != hipSuccess)
    return bail(TM_EDEVICE);
  for (int i = 0; i < kMlpIn; ++i) s.norm_h[kNormScale + i] = 1.f;
  for (int i = 0; i < kMlpOut; ++i) s.norm_h[kNormOut + i] = 1.f;
  if (hipMemcpy(s.norm, s.norm_h, sizeof(s.norm_h), hipMemcpyHostToDevice) != hipSuccess) return bail(TM_EDEVICE);
  if (mlp_pack(s, s.stream) || hipStreamSynchronize(s.stream) != hipSuccess) return bail(TM_EDEVICE);
  *out = p;
  return TM_OK;
}

extern "C" int tm_mlp_destroy(tm_mlp *p) {
  if (!p) return TM_EINVAL;
  for (void *q : p->s.allocs) (void)hipFree(q);
  for (void *q : p->s.batch_bufs) (void)hipFree(q);
  if (p->s.stream) (void)hipStreamDestroy(p->s.stream);
  delete p;
  return TM_OK;
}

extern "C" const char *tm_mlp_last_error(const tm_mlp *p) { return p ? p->s.err.c_str() : "null mlp"; }

extern "C" int tm_mlp_get_params(tm_mlp *p, float *params) {
  if (!p || !params) return TM_EINVAL;
  MLPCHK(p, hipMemcpy(params, p->s.params, kMlpNParams * 4, hipMemcpyDeviceToHost));
  return TM_OK;
}

extern "C" int tm_mlp_set_params(tm_mlp *p, const float *params) {
  if (!p || !params) return TM_EINVAL;
  for (int i = 0; i < kMlpNParams; ++i)
    if (!std::isfinite(params[i])) {
      p->s.err = "non-finite parameter";
      return TM_ENONFINITE;
    }
  MLPCHK(p, hipMemcpy(p->s.params, params, kMlpNParams * 4, hipMemcpyHostToDevice));
  int rc = mlp_pack(p->s, p->s.stream);
  if (rc) return rc;
  MLPCHK(p, hipStreamSynchronize(p->s.stream));
  return TM_OK;
}

extern "C" int tm_mlp_set_scales(tm_mlp *p, const float *in_shift, const float *in_scale, const float *out_scale) {
  if (!p || !in_shift || !in_scale || !out_scale) return TM_EINVAL;
  float h[kNormLen];
  std::memcpy(h + kNormShift, in_shift, kMlpIn * 4);
  std::memcpy(h + kNormScale, in_scale, kMlpIn * 4);
  std::memcpy(h + kNormOut, out_scale, kMlpOut * 4);
  for (int i = 0; i < kNormLen; ++i)
    if (!std::isfinite(h[i]) || (i >= kNormOut && h[i] == 0.f)) {
      p->s.err = "scales must be finite and output scales non-zero";
      return TM_EINVAL;
    }
  // the device copy is read by kernels of other streams (a context's step):
  // a synchronous copy after the device is idle for this MLP
  MLPCHK(p, hipStreamSynchronize(p->s.stream));
  MLPCHK(p, hipMemcpy(p->s.norm, h, sizeof(h), hipMemcpyHostToDevice));
  std::memcpy(p->s.norm_h, h, sizeof(h));
  return TM_OK;
}

extern "C" int tm_mlp_get_scales(tm_mlp *p, float *in_shift, float *in_scale, float *out_scale) {
  if (!p || !in_shift || !in_scale || !out_scale) return TM_EINVAL;
  std::memcpy(in_shift, p->s.norm_h + kNormShift, kMlpIn * 4);
  std::memcpy(in_scale, p->s.norm_h + kNormScale, kMlpIn * 4);
  std::memcpy(out_scale, p->s.norm_h + kNormOut, kMlpOut * 4);
  return TM_OK;
}

extern "C" int tm_mlp_forward(tm_mlp *p, const float *x, int n, float *y) {
  if (!p || !x || !y || n < 0) return TM_EINVAL;
  MlpState &s = p->s;
  int rc = mlp_reserve(s, std::max(n, 1));
  if (rc) return rc;
  MLPCHK(p, hipMemcpyAsync(s.X, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, s.stream));
  if ((rc = mlp_forward(s, s.X, n, s.Y, false, s.stream))) return rc;
  MLPCHK(p, hipMemcpyAsync(y, s.Y, (size_t)n * kMlpOut * 4, hipMemcpyDeviceToHost, s.stream));
  MLPCHK(p, hipStreamSynchronize(s.stream));
  return TM_OK;
}

extern "C" int tm_mlp_backward(tm_mlp *p, const float *x, const float *dy, int n, float *grad, float *dx) {
  if (!p || !x || !dy || !grad || n < 1) return TM_EINVAL;
  MlpState &s = p->s;
  int rc = mlp_reserve(s, n);
  if (rc) return rc;
  MLPCHK(p, hipMemcpyAsync(s.X, x, (size_t)n * kMlpIn * 4, hipMemcpyHostToDevice, s.stream));
  MLPCHK(p, hipMemcpyAsync(s.dY, dy, (size_t)n * kMlpOut * 4, hipMemcpyHostToDevice, s.stream));
  if ((rc = mlp_zero_grad(s, s.stream)) || (rc = mlp_forward(s, s.X, n, s.Y, true, s.stream)) ||
      (rc = mlp_backward(s, s.X, s.dY, n, dx ? s.dX : nullptr, s.stream)))
    return rc;
  MLPCHK(p, hipMemcpyAsync(grad, s.grad, kMlpNParams * 4, hipMemcpyDeviceToHost, s.stream));
  if (dx) MLPCHK(p, hipMemcpyAsync(dx, s.dX, (size_t)n * kMlpIn * 4, hipMemcpyDeviceToHost, s.stream));
  MLPCHK(p, hipStreamSynchronize(s.stream));
  return TM_OK;
}

extern "C" int tm_mlp_adam_step(tm_mlp *p, const float *grad, float lr) {
  if (!p || !grad) return TM_EINVAL;
  MLPCHK(p, hipMemcpyAsync(p->s.grad, grad, kMlpNParams * 4, hipMemcpyHostToDevice, p->s.stream));
  int rc = mlp_adam(p->s, lr, 0.9f, 0.999f, 1e-8f, p->s.stream);
  if (rc) return rc;
  MLPCHK(p, hipStreamSynchronize(p->s.stream));
  return TM_OK;
}

// Internal accessor for the closure code in tmgpu.hip.
namespace tmk {
MlpState &mlp_state(tm_mlp *p) { return p->s; }
}  // namespace tmk
