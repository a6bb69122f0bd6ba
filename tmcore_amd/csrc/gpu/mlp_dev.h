// Per-edge MLP of the learned closure: 40 -> 128 -> 128 -> 128 -> 2, GELU,
// BF16 weights on MFMA with FP32 accumulation, FP32 master weights and Adam.
// Device entry points shared by mlp.hip (kernels) and tmgpu.hip (closure).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace tmk {

constexpr int kMlpIn = 40;
constexpr int kMlpH = 128;
constexpr int kMlpOut = 2;
// flat FP32 parameter vector: W1[40][128] b1 W2[128][128] b2 W3 b3 W4[128][2] b4
constexpr int kOW1 = 0;
constexpr int kOB1 = kOW1 + kMlpIn * kMlpH;
constexpr int kOW2 = kOB1 + kMlpH;
constexpr int kOB2 = kOW2 + kMlpH * kMlpH;
constexpr int kOW3 = kOB2 + kMlpH;
constexpr int kOB3 = kOW3 + kMlpH * kMlpH;
constexpr int kOW4 = kOB3 + kMlpH;
constexpr int kOB4 = kOW4 + kMlpH * kMlpOut;
constexpr int kMlpNParams = kOB4 + kMlpOut;  // 38530
constexpr int kNormShift = 0, kNormScale = kMlpIn, kNormOut = 2 * kMlpIn, kNormLen = 2 * kMlpIn + kMlpOut;

struct MlpState {
  int device = 0;
  hipStream_t stream = nullptr;
  float *params = nullptr;        // [kMlpNParams] FP32 master
  float *grad = nullptr;          // [kMlpNParams]
  float *adam_m = nullptr, *adam_v = nullptr;
  int64_t adam_t = 0;
  uint16_t *pack = nullptr;       // BF16 MFMA operand packs, forward P(W1..W4) then backward Q(W1..W4)
  size_t pack_off[8] = {};        // offsets (in uint16) of the eight packs
  int ncu = 1;                    // compute units of the device (workgroups of the persistent kernels)
  // per-batch buffers for up to cap rows
  int cap = 0;
  float *X = nullptr, *Y = nullptr, *dY = nullptr, *dX = nullptr;
  uint16_t *H = nullptr, *D = nullptr, *dZ = nullptr;  // [3][cap][128] each
  float *gpart = nullptr;  // [parts][kMlpNParams] per-row-range gradient partials
  // input normalisation and output scales, host copy and device copy
  // norm[kNormShift + j], norm[kNormScale + j], norm[kNormOut + k]
  float norm_h[2 * kMlpIn + kMlpOut] = {};
  float *norm = nullptr;
  std::string err;
  std::vector<void *> allocs;   // freed at destroy
  std::vector<void *> batch_bufs;  // per-batch buffers, replaced when the batch grows
};

// Launches go to stream st; device pointers.  Return 0 or a TM_E* code.
int mlp_reserve(MlpState &s, int rows);
int mlp_pack(MlpState &s, hipStream_t st);
// Y[n][2] = MLP(X[n][40]); train = 1 keeps H/D for the backward pass.
int mlp_forward(MlpState &s, const float *X, int n, float *Y, bool train, hipStream_t st);
// Given dL/dY[n][2] (FP32) after mlp_forward(train): accumulates dL/dparams
// into s.grad and, when dX != nullptr, writes dL/dX[n][40].
int mlp_backward(MlpState &s, const float *X, const float *dY, int n, float *dX, hipStream_t st);
int mlp_zero_grad(MlpState &s, hipStream_t st);
// Adam update of s.params from s.grad (which it clears), then repack.
int mlp_adam(MlpState &s, float lr, float beta1, float beta2, float eps, hipStream_t st);

}  // namespace tmk

struct tm_mlp;
namespace tmk {
MlpState &mlp_state(tm_mlp *p);
}  // namespace tmk
