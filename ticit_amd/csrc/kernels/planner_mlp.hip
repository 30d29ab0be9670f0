// The learned planner's cost model: a 12 -> 64 -> 64 -> 3 MLP in FP32
// (ReLU hidden layers). Outputs 0 and 1 are the logits of "dense" and
// "tensor components"; output 2 is the standardized log10 time per shot of
// the faster layout.
//
// Training kernel: one workgroup of 256 threads runs whole minibatches of
// B <= 64 samples (forward, loss, backward, Adam) with parameters,
// activations and gradients in LDS. Every gradient entry is summed over the
// minibatch by one thread in sample order, so gradients are deterministic.
// Inference kernel: one thread per sample.
//
// Parameter layout (floats): W1[12][64], b1[64], W2[64][64], b2[64],
// W3[64][3], b3[3] (row = input unit).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ticit {

constexpr int kIn = 12, kH = 64, kOut = 3;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kIn * kH, kOffW2 = kOffB1 + kH, kOffB2 = kOffW2 + kH * kH,
              kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kH * kOut, kParams = kOffB3 + kOut;
constexpr int kMaxBatch = 64;
constexpr int kThreads = 256;

struct MlpStepArgs {
  float* params;      // [kParams]
  float* m;           // Adam first moments
  float* v;           // Adam second moments
  const float* x;     // [n][kIn]
  const int32_t* y;   // [n] 0 dense, 1 components
  const float* t;     // [n] standardized log10 time target
  const uint32_t* order;  // [steps * batch] sample indices
  uint32_t steps, batch;
  float lr, beta1, beta2, eps, time_weight;
  uint32_t step0;     // Adam steps already taken
  int32_t update;     // 0: only gradients and loss of the first minibatch
  float* grads;       // [kParams] when update == 0
  float* losses;      // [steps] mean minibatch loss
};

__global__ __launch_bounds__(kThreads) void mlp_train_kernel(MlpStepArgs a) {
  __shared__ float P[kParams];
  __shared__ float G[kParams];
  __shared__ float X[kMaxBatch][kIn];
  __shared__ float H1[kMaxBatch][kH];
  __shared__ float H2[kMaxBatch][kH];
  __shared__ float D2[kMaxBatch][kH];  // dL/d(pre-activation 2)
  __shared__ float D1[kMaxBatch][kH];  // dL/d(pre-activation 1)
  __shared__ float DO[kMaxBatch][kOut];
  __shared__ float L[kMaxBatch];
  const int tid = threadIdx.x;
  const int B = int(a.batch);
  for (int i = tid; i < kParams; i += kThreads) P[i] = a.params[i];
  __syncthreads();
  for (uint32_t s = 0; s < a.steps; ++s) {
    const uint32_t* idx = a.order + size_t(s) * a.batch;
    for (int e = tid; e < B * kIn; e += kThreads) X[e / kIn][e % kIn] = a.x[size_t(idx[e / kIn]) * kIn + e % kIn];
    __syncthreads();
    // forward
    for (int e = tid; e < B * kH; e += kThreads) {
      const int b = e / kH, j = e % kH;
      float acc = P[kOffB1 + j];
      for (int i = 0; i < kIn; ++i) acc += X[b][i] * P[kOffW1 + i * kH + j];
      H1[b][j] = acc > 0.f ? acc : 0.f;
    }
    __syncthreads();
    for (int e = tid; e < B * kH; e += kThreads) {
      const int b = e / kH, j = e % kH;
      float acc = P[kOffB2 + j];
      for (int i = 0; i < kH; ++i) acc += H1[b][i] * P[kOffW2 + i * kH + j];
      H2[b][j] = acc > 0.f ? acc : 0.f;
    }
    __syncthreads();
    // outputs, loss and dL/d(outputs): softmax cross-entropy on the choice
    // plus time_weight * squared error on the time, averaged over the batch
    if (tid < B) {
      const int b = tid;
      float o[kOut];
      for (int k = 0; k < kOut; ++k) {
        float acc = P[kOffB3 + k];
        for (int i = 0; i < kH; ++i) acc += H2[b][i] * P[kOffW3 + i * kOut + k];
        o[k] = acc;
      }
      const size_t sample = idx[b];
      const int y = a.y[sample];
      const float mx = fmaxf(o[0], o[1]);
      const float e0 = expf(o[0] - mx), e1 = expf(o[1] - mx), z = e0 + e1;
      const float p0 = e0 / z, p1 = e1 / z;
      const float d = o[2] - a.t[sample];
      L[b] = -logf(y ? p1 : p0) + a.time_weight * d * d;
      const float inv = 1.f / float(B);
      DO[b][0] = (p0 - (y == 0 ? 1.f : 0.f)) * inv;
      DO[b][1] = (p1 - (y == 1 ? 1.f : 0.f)) * inv;
      DO[b][2] = 2.f * a.time_weight * d * inv;
    }
    __syncthreads();
    // backward: layer 3
    for (int e = tid; e < kH * kOut; e += kThreads) {
      const int i = e / kOut, k = e % kOut;
      float acc = 0.f;
      for (int b = 0; b < B; ++b) acc += H2[b][i] * DO[b][k];
      G[kOffW3 + e] = acc;
    }
    if (tid < kOut) {
      float acc = 0.f;
      for (int b = 0; b < B; ++b) acc += DO[b][tid];
      G[kOffB3 + tid] = acc;
    }
    for (int e = tid; e < B * kH; e += kThreads) {
      const int b = e / kH, j = e % kH;
      float acc = 0.f;
      for (int k = 0; k < kOut; ++k) acc += DO[b][k] * P[kOffW3 + j * kOut + k];
      D2[b][j] = H2[b][j] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    // layer 2
    for (int e = tid; e < kH * kH; e += kThreads) {
      const int i = e / kH, j = e % kH;
      float acc = 0.f;
      for (int b = 0; b < B; ++b) acc += H1[b][i] * D2[b][j];
      G[kOffW2 + e] = acc;
    }
    if (tid < kH) {
      float acc = 0.f;
      for (int b = 0; b < B; ++b) acc += D2[b][tid];
      G[kOffB2 + tid] = acc;
    }
    for (int e = tid; e < B * kH; e += kThreads) {
      const int b = e / kH, i = e % kH;
      float acc = 0.f;
      for (int j = 0; j < kH; ++j) acc += D2[b][j] * P[kOffW2 + i * kH + j];
      D1[b][i] = H1[b][i] > 0.f ? acc : 0.f;
    }
    __syncthreads();
    // layer 1
    for (int e = tid; e < kIn * kH; e += kThreads) {
      const int i = e / kH, j = e % kH;
      float acc = 0.f;
      for (int b = 0; b < B; ++b) acc += X[b][i] * D1[b][j];
      G[kOffW1 + e] = acc;
    }
    if (tid < kH) {
      float acc = 0.f;
      for (int b = 0; b < B; ++b) acc += D1[b][tid];
      G[kOffB1 + tid] = acc;
    }
    if (tid == 0) {
      float acc = 0.f;
      for (int b = 0; b < B; ++b) acc += L[b];
      a.losses[s] = acc / float(B);
    }
    __syncthreads();
    if (!a.update) {
      for (int i = tid; i < kParams; i += kThreads) a.grads[i] = G[i];
      return;
    }
    // Adam (Kingma & Ba, 2015), bias-corrected
    const float tstep = float(a.step0 + s + 1);
    const float c1 = 1.f - powf(a.beta1, tstep), c2 = 1.f - powf(a.beta2, tstep);
    for (int i = tid; i < kParams; i += kThreads) {
      const float g = G[i];
      const float m = a.beta1 * a.m[i] + (1.f - a.beta1) * g;
      const float v = a.beta2 * a.v[i] + (1.f - a.beta2) * g * g;
      a.m[i] = m;
      a.v[i] = v;
      P[i] -= a.lr * (m / c1) / (sqrtf(v / c2) + a.eps);
    }
    __syncthreads();
  }
  for (int i = tid; i < kParams; i += kThreads) a.params[i] = P[i];
}

__global__ __launch_bounds__(256) void mlp_infer_kernel(const float* __restrict__ params, const float* __restrict__ x,
                                                        uint32_t n, float* __restrict__ out) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float xin[kIn], h1[kH];
  for (int i = 0; i < kIn; ++i) xin[i] = x[size_t(s) * kIn + i];
  for (int j = 0; j < kH; ++j) {
    float acc = params[kOffB1 + j];
    for (int i = 0; i < kIn; ++i) acc += xin[i] * params[kOffW1 + i * kH + j];
    h1[j] = acc > 0.f ? acc : 0.f;
  }
  float o[kOut];
  for (int k = 0; k < kOut; ++k) o[k] = params[kOffB3 + k];
  for (int j = 0; j < kH; ++j) {
    float acc = params[kOffB2 + j];
    for (int i = 0; i < kH; ++i) acc += h1[i] * params[kOffW2 + i * kH + j];
    const float h2 = acc > 0.f ? acc : 0.f;
    for (int k = 0; k < kOut; ++k) o[k] += h2 * params[kOffW3 + j * kOut + k];
  }
  for (int k = 0; k < kOut; ++k) out[size_t(s) * kOut + k] = o[k];
}

int mlp_param_count() { return kParams; }
int mlp_max_batch() { return kMaxBatch; }

hipError_t launch_mlp_train(const MlpStepArgs& a, hipStream_t stream) {
  if (a.batch == 0 || a.batch > uint32_t(kMaxBatch)) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mlp_train_kernel, dim3(1), dim3(kThreads), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_mlp_infer(const float* params, const float* x, uint32_t n, float* out, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_infer_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, params, x, n, out);
  return hipGetLastError();
}

}  // namespace ticit
