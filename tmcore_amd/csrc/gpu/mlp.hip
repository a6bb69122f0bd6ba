// BF16 MFMA kernels of the per-edge closure MLP (gfx950).
//
// Matrix core: v_mfma_f32_16x16x16_bf16 (__builtin_amdgcn_mfma_f32_16x16x16bf16_1k)
// on wave64.  Operand layout per lane l (0..63), with q = l / 16, r = l % 16:
//   A (16x16, rows x k):  A[r][4q + 0..3]
//   B (16x16, k x cols):  B[4q + 0..3][r]
//   D (16x16 FP32):       D[4q + 0..3][r]
// Weights are stored once per update as BF16 "packs" in exactly the B layout
// (one 8-byte load per lane per MFMA), for both W (forward) and W^T
// (backward through the layer).  Activations of a 16-row tile live in LDS,
// row-major, so A fragments are one 8-byte LDS read per lane.
//
// Layers: Z1 = X W1 + b1, H1 = GELU(Z1); Z2 = H1 W2 + b2, H2 = GELU(Z2);
// Z3 = H2 W3 + b3, H3 = GELU(Z3); Y = H3 W4 + b4.  GELU is the exact form of
// Hendrycks & Gimpel (2016), x * Phi(x), with Phi from erff.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>

#include "mlp_dev.h"
#include "tmcore_gpu.h"

namespace tmk {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

constexpr int kMlpBlock = 256;  // 4 waves, 16 rows each
constexpr int kActStride = kMlpH + 8;

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(((uint32_t)b) << 16); }

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float x) {
  return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

// pack index: [ks][nt][lane][4]
struct PackDesc {
  const float *W;  // row-major [Kin][Nout]
  int Kin, Nout, KS, NT, transposed;
  uint16_t *out;
};

__global__ void k_mlp_pack(PackDesc d) {
  const int total = d.KS * d.NT * 64 * 4;
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int i = idx & 3, lane = (idx >> 2) & 63, tile = idx >> 8;
  const int nt = tile % d.NT, ks = tile / d.NT;
  const int k = ks * 16 + 4 * (lane >> 4) + i, n = nt * 16 + (lane & 15);
  float v = 0.f;
  if (!d.transposed) {
    if (k < d.Kin && n < d.Nout) v = d.W[(size_t)k * d.Nout + n];
  } else {
    if (k < d.Nout && n < d.Kin) v = d.W[(size_t)n * d.Nout + k];
  }
  d.out[idx] = f2bf(v);
}

// acc[nt] = A(16 x 16*KS, LDS row-major) x B(pack)
template <int KS, int NT>
__device__ __forceinline__ void wave_mma(const uint16_t (*A)[kActStride], const bf16x4_t *pack, f32x4_t (&acc)[NT],
                                         int lane) {
#pragma unroll
  for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4_t{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < KS; ++ks) {
    const uint16_t *ap = &A[lane & 15][ks * 16 + 4 * (lane >> 4)];
    const bf16x4_t a = *reinterpret_cast<const bf16x4_t *>(ap);
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
      const bf16x4_t b = pack[(ks * NT + nt) * 64 + lane];
      acc[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[nt], 0, 0, 0);
    }
  }
}

struct MlpPtrs {
  const bf16x4_t *f1, *f2, *f3, *f4, *t1, *t2, *t3, *t4;
  const float *params;
};

// Hidden layer epilogue: bias + GELU, H (and GELU') to global when training,
// H as the next A operand in LDS.
__device__ __forceinline__ void hidden_epilogue(const f32x4_t (&acc)[8], const float *bias, uint16_t (*act)[kActStride],
                                                int lane, int row0, int n, bool train, uint16_t *H, uint16_t *D) {
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) {
    const int col = nt * 16 + (lane & 15);
    const float b = bias[col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * (lane >> 4) + i;
      const float z = acc[nt][i] + b;
      const float h = gelu(z);
      const uint16_t hb = f2bf(h);
      act[r][col] = hb;
      if (train && row0 + r < n) {
        H[(size_t)(row0 + r) * kMlpH + col] = hb;
        D[(size_t)(row0 + r) * kMlpH + col] = f2bf(gelu_d(z));
      }
    }
  }
}

__global__ __launch_bounds__(kMlpBlock) void k_mlp_fwd(MlpPtrs p, const float *__restrict__ X, int n,
                                                       float *__restrict__ Y, int train, uint16_t *H, uint16_t *D) {
  __shared__ __attribute__((aligned(16))) uint16_t act[kMlpBlock / 64][16][kActStride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * (kMlpBlock / 64) + wave) * 16;
  uint16_t(*A)[kActStride] = act[wave];
  for (int i = lane; i < 16 * kMlpInP; i += 64) {
    const int r = i / kMlpInP, k = i % kMlpInP;
    const float v = (row0 + r < n && k < kMlpIn) ? X[(size_t)(row0 + r) * kMlpIn + k] : 0.f;
    A[r][k] = f2bf(v);
  }
  __syncthreads();
  const size_t plane = (size_t)n * kMlpH;
  f32x4_t acc[8];
  wave_mma<3, 8>(A, p.f1, acc, lane);
  hidden_epilogue(acc, p.params + kOB1, A, lane, row0, n, train, H, D);
  __syncthreads();
  wave_mma<8, 8>(A, p.f2, acc, lane);
  __syncthreads();
  hidden_epilogue(acc, p.params + kOB2, A, lane, row0, n, train, H + plane, D + plane);
  __syncthreads();
  wave_mma<8, 8>(A, p.f3, acc, lane);
  __syncthreads();
  hidden_epilogue(acc, p.params + kOB3, A, lane, row0, n, train, H + 2 * plane, D + 2 * plane);
  __syncthreads();
  f32x4_t out[1];
  wave_mma<8, 1>(A, p.f4, out, lane);
  const int col = lane & 15;
  if (col < kMlpOut) {
    const float b = p.params[kOB4 + col];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * (lane >> 4) + i;
      if (row0 + r < n) Y[(size_t)(row0 + r) * kMlpOut + col] = out[0][i] + b;
    }
  }
}

// dZ of a hidden layer = (upstream A x W^T) * GELU'(Z); kept in LDS and saved.
__device__ __forceinline__ void dz_epilogue(const f32x4_t (&acc)[8], const uint16_t *D, uint16_t (*act)[kActStride],
                                            int lane, int row0, int n, uint16_t *dZ) {
#pragma unroll
  for (int nt = 0; nt < 8; ++nt) {
    const int col = nt * 16 + (lane & 15);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = 4 * (lane >> 4) + i;
      float dz = 0.f;
      if (row0 + r < n) dz = acc[nt][i] * bf2f(D[(size_t)(row0 + r) * kMlpH + col]);
      const uint16_t b = f2bf(dz);
      act[r][col] = b;
      if (row0 + r < n) dZ[(size_t)(row0 + r) * kMlpH + col] = b;
    }
  }
}

__global__ __launch_bounds__(kMlpBlock) void k_mlp_bwd(MlpPtrs p, const float *__restrict__ dY, int n,
                                                       const uint16_t *D, uint16_t *dZ, uint16_t *dYb,
                                                       float *__restrict__ dX) {
  __shared__ __attribute__((aligned(16))) uint16_t act[kMlpBlock / 64][16][kActStride];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int row0 = (blockIdx.x * (kMlpBlock / 64) + wave) * 16;
  uint16_t(*A)[kActStride] = act[wave];
  for (int i = lane; i < 16 * 16; i += 64) {
    const int r = i / 16, k = i % 16;
    const float v = (row0 + r < n && k < kMlpOut) ? dY[(size_t)(row0 + r) * kMlpOut + k] : 0.f;
    const uint16_t b = f2bf(v);
    A[r][k] = b;
    if (row0 + r < n) dYb[(size_t)(row0 + r) * kMlpOutP + k] = b;
  }
  __syncthreads();
  const size_t plane = (size_t)n * kMlpH;
  f32x4_t acc[8];
  wave_mma<1, 8>(A, p.t4, acc, lane);  // dH3 = dY W4^T
  __syncthreads();
  dz_epilogue(acc, D + 2 * plane, A, lane, row0, n, dZ + 2 * plane);
  __syncthreads();
  wave_mma<8, 8>(A, p.t3, acc, lane);  // dH2 = dZ3 W3^T
  __syncthreads();
  dz_epilogue(acc, D + plane, A, lane, row0, n, dZ + plane);
  __syncthreads();
  wave_mma<8, 8>(A, p.t2, acc, lane);  // dH1 = dZ2 W2^T
  __syncthreads();
  dz_epilogue(acc, D, A, lane, row0, n, dZ);
  __syncthreads();
  if (dX) {
    f32x4_t ax[3];
    wave_mma<8, 3>(A, p.t1, ax, lane);  // dX = dZ1 W1^T
#pragma unroll
    for (int nt = 0; nt < 3; ++nt) {
      const int col = nt * 16 + (lane & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * (lane >> 4) + i;
        if (col < kMlpIn && row0 + r < n) dX[(size_t)(row0 + r) * kMlpIn + col] = ax[nt][i];
      }
    }
  }
}

// dW[k][j] += sum_rows A[row][k] dZ[row][j]  (MFMA over the row dimension).
// One wave per 16x16 output tile and row split; A is BF16 [n][lda] or, for
// layer 1, FP32 X [n][40].
__global__ __launch_bounds__(64) void k_mlp_wgrad(const uint16_t *A, const float *Xf, int lda, int Kin,
                                                  const uint16_t *dZ, int ldz, int Nout, int n, int rows_per_split,
                                                  float *dW) {
  const int ntk = (Kin + 15) / 16;
  const int tk = blockIdx.x % ntk, tj = blockIdx.x / ntk;
  const int lane = threadIdx.x;
  const int r_begin = blockIdx.y * rows_per_split, r_end = min(n, r_begin + rows_per_split);
  const int kin = tk * 16 + (lane & 15), col = tj * 16 + (lane & 15);
  f32x4_t acc{0.f, 0.f, 0.f, 0.f};
  for (int r0 = r_begin; r0 < r_end; r0 += 16) {
    bf16x4_t a, b;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int row = r0 + 4 * (lane >> 4) + i;
      const bool ok = row < r_end;
      uint16_t av = 0, bv = 0;
      if (ok && kin < Kin) av = Xf ? f2bf(Xf[(size_t)row * kMlpIn + kin]) : A[(size_t)row * lda + kin];
      if (ok && col < Nout) bv = dZ[(size_t)row * ldz + col];
      a[i] = (short)av;
      b[i] = (short)bv;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = tk * 16 + 4 * (lane >> 4) + i;
    if (m < Kin && col < Nout) atomicAdd(&dW[(size_t)m * Nout + col], acc[i]);
  }
}

// db[j] += sum_rows dZ[row][j] (BF16 dZ), FP32 atomics per block.
__global__ void k_mlp_bgrad(const uint16_t *dZ, int ldz, int N, int n, float *db) {
  const int j = threadIdx.x;
  if (j >= N) return;
  float s = 0.f;
  for (int row = blockIdx.x; row < n; row += gridDim.x) s += bf2f(dZ[(size_t)row * ldz + j]);
  atomicAdd(&db[j], s);
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
  int W_off, Kin, Nout, KS, NT, transposed;
};
const PackShape kPacks[8] = {
    {kOW1, kMlpIn, kMlpH, 3, 8, 0},  {kOW2, kMlpH, kMlpH, 8, 8, 0},  {kOW3, kMlpH, kMlpH, 8, 8, 0},
    {kOW4, kMlpH, kMlpOut, 8, 1, 0}, {kOW1, kMlpIn, kMlpH, 8, 3, 1}, {kOW2, kMlpH, kMlpH, 8, 8, 1},
    {kOW3, kMlpH, kMlpH, 8, 8, 1},   {kOW4, kMlpH, kMlpOut, 1, 8, 1}};

MlpPtrs ptrs(const MlpState &s) {
  MlpPtrs p;
  const bf16x4_t *b = reinterpret_cast<const bf16x4_t *>(s.pack);
  p.f1 = b + s.pack_off[0] / 4;
  p.f2 = b + s.pack_off[1] / 4;
  p.f3 = b + s.pack_off[2] / 4;
  p.f4 = b + s.pack_off[3] / 4;
  p.t1 = b + s.pack_off[4] / 4;
  p.t2 = b + s.pack_off[5] / 4;
  p.t3 = b + s.pack_off[6] / 4;
  p.t4 = b + s.pack_off[7] / 4;
  p.params = s.params;
  return p;
}

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
      !(s.dZ = (uint16_t *)a(3 * plane * 2)) || !(s.dYb = (uint16_t *)a((size_t)cap * kMlpOutP * 2)))
    return fail(s, TM_ENOMEM, "hipMalloc MLP batch buffers");
  s.cap = cap;
  return TM_OK;
}

int mlp_pack(MlpState &s, hipStream_t st) {
  for (int i = 0; i < 8; ++i) {
    const PackShape &sh = kPacks[i];
    PackDesc d{s.params + sh.W_off, sh.Kin, sh.Nout, sh.KS, sh.NT, sh.transposed, s.pack + s.pack_off[i]};
    const int total = sh.KS * sh.NT * 256;
    hipLaunchKernelGGL(k_mlp_pack, dim3((total + 255) / 256), dim3(256), 0, st, d);
  }
  return hipGetLastError() == hipSuccess ? TM_OK : fail(s, TM_EDEVICE, "k_mlp_pack launch");
}

int mlp_forward(MlpState &s, const float *X, int n, float *Y, bool train, hipStream_t st) {
  if (n <= 0) return TM_OK;
  if (train && n > s.cap) return fail(s, TM_EINVAL, "batch exceeds the reserved MLP batch buffers");
  const int rows_per_block = 16 * (kMlpBlock / 64);
  hipLaunchKernelGGL(k_mlp_fwd, dim3((n + rows_per_block - 1) / rows_per_block), dim3(kMlpBlock), 0, st, ptrs(s),
                     X, n, Y, train ? 1 : 0, s.H, s.D);
  return hipGetLastError() == hipSuccess ? TM_OK : fail(s, TM_EDEVICE, "k_mlp_fwd launch");
}

int mlp_backward(MlpState &s, const float *X, const float *dY, int n, float *dX, hipStream_t st) {
  if (n <= 0) return TM_OK;
  if (n > s.cap) return fail(s, TM_EINVAL, "batch exceeds the reserved MLP batch buffers");
  const int rows_per_block = 16 * (kMlpBlock / 64);
  hipLaunchKernelGGL(k_mlp_bwd, dim3((n + rows_per_block - 1) / rows_per_block), dim3(kMlpBlock), 0, st, ptrs(s),
                     dY, n, (const uint16_t *)s.D, s.dZ, s.dYb, dX);
  const size_t plane = (size_t)n * kMlpH;
  const int splits = std::max(1, std::min(256, (n + 255) / 256));
  const int rps = ((n + splits - 1) / splits + 15) / 16 * 16;
  // dW4 = H3^T dY, dW3 = H2^T dZ3, dW2 = H1^T dZ2, dW1 = X^T dZ1
  struct L {
    const uint16_t *A;
    const float *Xf;
    int Kin;
    const uint16_t *dZ;
    int ldz, Nout, woff, boff;
  } layers[4] = {{s.H + 2 * plane, nullptr, kMlpH, s.dYb, kMlpOutP, kMlpOut, kOW4, kOB4},
                 {s.H + plane, nullptr, kMlpH, s.dZ + 2 * plane, kMlpH, kMlpH, kOW3, kOB3},
                 {s.H, nullptr, kMlpH, s.dZ + plane, kMlpH, kMlpH, kOW2, kOB2},
                 {nullptr, X, kMlpIn, s.dZ, kMlpH, kMlpH, kOW1, kOB1}};
  for (const L &l : layers) {
    const int tiles = ((l.Kin + 15) / 16) * ((l.Nout + 15) / 16);
    hipLaunchKernelGGL(k_mlp_wgrad, dim3(tiles, splits), dim3(64), 0, st, l.A, l.Xf, kMlpH, l.Kin, l.dZ, l.ldz,
                       l.Nout, n, rps, s.grad + l.woff);
    hipLaunchKernelGGL(k_mlp_bgrad, dim3(std::min(n, 1024)), dim3(kMlpH), 0, st, l.dZ, l.ldz, l.Nout, n,
                       s.grad + l.boff);
  }
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
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking) != hipSuccess)
    return bail(TM_EDEVICE);
  size_t off = 0;
  for (int i = 0; i < 8; ++i) {
    s.pack_off[i] = off;
    off += (size_t)kPacks[i].KS * kPacks[i].NT * 256;
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
