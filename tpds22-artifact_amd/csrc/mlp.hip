// mlp.hip -- training and batch inference of the schedule MLP
// (8 -> 64 -> 64 -> 64 -> 8, ReLU, FP32) on the GPU.
//
// The same network is evaluated inside the multi-source BFS kernel (one hidden
// unit per lane of a wave64).  Every sample is one BFS level; its 8 outputs
// score the level's actions: direction (push / pull) x workgroups that expand
// (all / half) x barrier poll backoff (short / long), action = dir | half << 1 |
// long << 2.  Two losses:
//   labels: sample-weighted softmax cross-entropy against one action per sample;
//   costs:  the expected time of the softmax policy, sum_s sum_a p_sa * c_sa,
//           with c_sa the measured time of action a at level s (normalised by
//           the caller) -- training on the timings the plain method generates
//           minimises the schedule's expected total time directly.
// Gradients are accumulated per workgroup in a fixed sample order and folded
// across workgroups in a fixed order, so a training step is deterministic.
// Parameter layout (floats):
//   W1[64][8] b1[64] W2[64][64] b2[64] W3[64][64] b3[64] W4[8][64] b4[8]
#include <hip/hip_runtime.h>

#include "hs_internal.hpp"

namespace {

constexpr int kIn = 8, kH = 64, kOut = 8;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kH * kH, kW4 = kB3 + kH, kB4 = kW4 + kOut * kH, kParams = kB4 + kOut;
constexpr int kGradWgs = 64;  // workgroups (one wave each) accumulating gradients

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// He-uniform initialisation from the counter stream; biases zero
__global__ void k_mlp_init(float *w, uint64_t seed) {
  const uint64_t st = mix64(seed);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < kParams; i += gridDim.x * blockDim.x) {
    int fan_in = 0;
    if (i < kB1) fan_in = kIn;
    else if (i >= kW2 && i < kB2) fan_in = kH;
    else if (i >= kW3 && i < kB3) fan_in = kH;
    else if (i >= kW4 && i < kB4) fan_in = kH;
    if (!fan_in) { w[i] = 0.f; continue; }
    const double u = (double)(mix64(st + (uint64_t)(i + 1) * 0x9E3779B97F4A7C15ULL) >> 11) *
                     (1.0 / 9007199254740992.0);
    const double lim = sqrt(6.0 / fan_in);
    w[i] = (float)((2.0 * u - 1.0) * lim);
  }
}

struct Acts {
  float h1, h2, h3;  // this lane's hidden unit after ReLU
  float o[kOut];     // logits (uniform across the wave)
};

__device__ __forceinline__ Acts forward_lane(const float *w, const float *x) {
  const int j = threadIdx.x & 63;
  Acts a;
  float h = w[kB1 + j];
#pragma unroll
  for (int i = 0; i < kIn; ++i) h = fmaf(w[kW1 + j * kIn + i], x[i], h);
  a.h1 = fmaxf(h, 0.f);
  h = w[kB2 + j];
  for (int k = 0; k < kH; ++k) h = fmaf(w[kW2 + j * kH + k], __shfl(a.h1, k), h);
  a.h2 = fmaxf(h, 0.f);
  h = w[kB3 + j];
  for (int k = 0; k < kH; ++k) h = fmaf(w[kW3 + j * kH + k], __shfl(a.h2, k), h);
  a.h3 = fmaxf(h, 0.f);
#pragma unroll
  for (int q = 0; q < kOut; ++q) {
    float o = w[kW4 + q * kH + j] * a.h3;
    for (int k = 32; k >= 1; k >>= 1) o += __shfl_xor(o, k);  // fixed butterfly order
    a.o[q] = o + w[kB4 + q];
  }
  return a;
}

// logits of N samples; one wave per sample
__global__ void k_mlp_forward(const float *w, const float *x, int n, float *logits) {
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const Acts a = forward_lane(w, x + (size_t)s * kIn);
    if ((threadIdx.x & 63) < kOut) {
      float v = a.o[0];
#pragma unroll
      for (int q = 1; q < kOut; ++q) v = (int)(threadIdx.x & 63) == q ? a.o[q] : v;
      logits[(size_t)s * kOut + (threadIdx.x & 63)] = v;
    }
  }
}

// per-workgroup gradient partials of the loss.  Labels (cost == nullptr):
// L = sum_s wt[s] * CE(softmax(o_s), y_s) * inv_norm.  Costs: L = sum_s
// sum_a softmax(o_s)_a * cost[s][a] * inv_norm.
__global__ void __launch_bounds__(64) k_mlp_grad_partial(const float *w, const float *x, const int *y,
                                                        const float *wt, const float *cost, int n,
                                                        float inv_norm, float *partials, float *loss_partials) {
  __shared__ float g[kParams];
  const int j = threadIdx.x;
  for (int i = j; i < kParams; i += 64) g[i] = 0.f;
  __syncthreads();
  float loss = 0.f;
  for (int s = blockIdx.x; s < n; s += gridDim.x) {
    const float *xs = x + (size_t)s * kIn;
    const Acts a = forward_lane(w, xs);
    float m = a.o[0];
#pragma unroll
    for (int q = 1; q < kOut; ++q) m = fmaxf(m, a.o[q]);
    float e[kOut], z = 0.f;
#pragma unroll
    for (int q = 0; q < kOut; ++q) { e[q] = expf(a.o[q] - m); z += e[q]; }
    float p[kOut], d[kOut];
#pragma unroll
    for (int q = 0; q < kOut; ++q) p[q] = e[q] / z;
    if (!cost) {
      const float ws = wt[s] * inv_norm;
      const int lab = y[s];
      float pl = p[0];
#pragma unroll
      for (int q = 1; q < kOut; ++q) pl = q == lab ? p[q] : pl;
      loss += -ws * logf(pl);
#pragma unroll
      for (int q = 0; q < kOut; ++q) d[q] = ws * (p[q] - (q == lab ? 1.f : 0.f));
    } else {
      const float *cs = cost + (size_t)s * kOut;
      float ec = 0.f;
#pragma unroll
      for (int q = 0; q < kOut; ++q) ec = fmaf(p[q], cs[q], ec);
      loss += ec * inv_norm;
#pragma unroll
      for (int q = 0; q < kOut; ++q) d[q] = p[q] * (cs[q] - ec) * inv_norm;
    }
    // layer 4
    float dh3 = 0.f;
#pragma unroll
    for (int q = 0; q < kOut; ++q) {
      g[kW4 + q * kH + j] += d[q] * a.h3;
      dh3 = fmaf(w[kW4 + q * kH + j], d[q], dh3);
    }
    if (j == 0) {
#pragma unroll
      for (int q = 0; q < kOut; ++q) g[kB4 + q] += d[q];
    }
    const float dz3 = (a.h3 > 0.f) ? dh3 : 0.f;
    // layer 3: dW3[j][k] = dz3[j] * h2[k]; dh2[k] = sum_j W3[j][k] dz3[j]
    for (int k = 0; k < kH; ++k) g[kW3 + j * kH + k] += dz3 * __shfl(a.h2, k);
    g[kB3 + j] += dz3;
    float dh2 = 0.f;
    for (int q = 0; q < kH; ++q) dh2 = fmaf(w[kW3 + q * kH + j], __shfl(dz3, q), dh2);
    const float dz2 = (a.h2 > 0.f) ? dh2 : 0.f;
    for (int k = 0; k < kH; ++k) g[kW2 + j * kH + k] += dz2 * __shfl(a.h1, k);
    g[kB2 + j] += dz2;
    float dh1 = 0.f;
    for (int q = 0; q < kH; ++q) dh1 = fmaf(w[kW2 + q * kH + j], __shfl(dz2, q), dh1);
    const float dz1 = (a.h1 > 0.f) ? dh1 : 0.f;
#pragma unroll
    for (int i = 0; i < kIn; ++i) g[kW1 + j * kIn + i] += dz1 * xs[i];
    g[kB1 + j] += dz1;
  }
  __syncthreads();
  for (int i = j; i < kParams; i += 64) partials[(size_t)blockIdx.x * kParams + i] = g[i];
  if (j == 0) loss_partials[blockIdx.x] = loss;
}

// fixed-order fold of the partials; optional Adam update
__global__ void k_mlp_fold(const float *partials, int nparts, const float *loss_partials, float *grad,
                           float *loss, float *w, float *m1, float *m2, float lr, float b1, float b2,
                           float eps, int step) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && loss) {
    float l = 0.f;
    for (int p = 0; p < nparts; ++p) l += loss_partials[p];
    *loss = l;
  }
  if (i >= kParams) return;
  float gsum = 0.f;
  for (int p = 0; p < nparts; ++p) gsum += partials[(size_t)p * kParams + i];
  if (grad) grad[i] = gsum;
  if (w) {
    const float a = b1 * m1[i] + (1.f - b1) * gsum;
    const float v = b2 * m2[i] + (1.f - b2) * gsum * gsum;
    m1[i] = a;
    m2[i] = v;
    const float ah = a / (1.f - powf(b1, (float)step));
    const float vh = v / (1.f - powf(b2, (float)step));
    w[i] -= lr * ah / (sqrtf(vh) + eps);
  }
}

int grad_launch(const float *w, const float *x, const int *y, const float *wt, const float *cost, int n,
                float inv_norm, float *grad, float *loss, float *upd_w, float lr, int step, float *scratch,
                hipStream_t s) {
  float *parts = scratch, *lparts = scratch + (size_t)kGradWgs * kParams;
  float *m1 = lparts + kGradWgs, *m2 = m1 + kParams;
  hipLaunchKernelGGL(k_mlp_grad_partial, dim3(kGradWgs), dim3(64), 0, s, w, x, y, wt, cost, n, inv_norm, parts,
                     lparts);
  hipLaunchKernelGGL(k_mlp_fold, dim3((kParams + 255) / 256), dim3(256), 0, s, parts, kGradWgs, lparts, grad,
                     loss, upd_w, upd_w ? m1 : nullptr, upd_w ? m2 : nullptr, lr, 0.9f, 0.999f, 1e-8f, step);
  HS_TRY_HIP(hipGetLastError());
  return HS_OK;
}

}  // namespace

extern "C" {

int hs_mlp_param_count(void) { return kParams; }

int hs_mlp_actions(void) { return kOut; }

int hs_mlp_scratch_floats(void) { return kGradWgs * kParams + kGradWgs + 2 * kParams + 8; }

int hs_mlp_init(float *d_w, uint64_t seed, void *stream) {
  if (!d_w) return HS_ERR_INVALID;
  hipLaunchKernelGGL(k_mlp_init, dim3(64), dim3(256), 0, (hipStream_t)stream, d_w, seed);
  HS_TRY_HIP(hipGetLastError());
  return HS_OK;
}

int hs_mlp_forward(const float *d_w, const float *d_x, int n, float *d_logits, void *stream) {
  if (!d_w || (n && (!d_x || !d_logits)) || n < 0) return HS_ERR_INVALID;
  if (!n) return HS_OK;
  hipLaunchKernelGGL(k_mlp_forward, dim3(n < 4096 ? n : 4096), dim3(64), 0, (hipStream_t)stream, d_w,
                     d_x, n, d_logits);
  HS_TRY_HIP(hipGetLastError());
  return HS_OK;
}

// gradient of the weighted cross-entropy (no update): d_grad[kParams], d_loss[1]
int hs_mlp_grad(const float *d_w, const float *d_x, const int *d_y, const float *d_wt, int n,
                float wsum, float *d_grad, float *d_loss, float *d_scratch, void *stream) {
  if (!d_w || !d_x || !d_y || !d_wt || n <= 0 || !(wsum > 0.f) || !d_grad || !d_scratch)
    return HS_ERR_INVALID;
  return grad_launch(d_w, d_x, d_y, d_wt, nullptr, n, 1.f / wsum, d_grad, d_loss, nullptr, 0.f, 0, d_scratch,
                     (hipStream_t)stream);
}

// one Adam step on the weighted cross-entropy; d_scratch holds the moments after
// the partial buffers (hs_mlp_scratch_floats); zero it before the first step
int hs_mlp_train_step(float *d_w, const float *d_x, const int *d_y, const float *d_wt, int n, float wsum,
                      float lr, int step, float *d_loss, float *d_scratch, void *stream) {
  if (!d_w || !d_x || !d_y || !d_wt || n <= 0 || !(wsum > 0.f) || step < 1 || !d_scratch)
    return HS_ERR_INVALID;
  return grad_launch(d_w, d_x, d_y, d_wt, nullptr, n, 1.f / wsum, nullptr, d_loss, d_w, lr, step, d_scratch,
                     (hipStream_t)stream);
}

// the expected-cost loss: d_cost is [n][hs_mlp_actions()], the loss is
// sum_s sum_a softmax(o_s)_a * cost[s][a] / norm
int hs_mlp_grad_cost(const float *d_w, const float *d_x, const float *d_cost, int n, float norm,
                     float *d_grad, float *d_loss, float *d_scratch, void *stream) {
  if (!d_w || !d_x || !d_cost || n <= 0 || !(norm > 0.f) || !d_grad || !d_scratch) return HS_ERR_INVALID;
  return grad_launch(d_w, d_x, nullptr, nullptr, d_cost, n, 1.f / norm, d_grad, d_loss, nullptr, 0.f, 0,
                     d_scratch, (hipStream_t)stream);
}

int hs_mlp_train_step_cost(float *d_w, const float *d_x, const float *d_cost, int n, float norm, float lr,
                           int step, float *d_loss, float *d_scratch, void *stream) {
  if (!d_w || !d_x || !d_cost || n <= 0 || !(norm > 0.f) || step < 1 || !d_scratch) return HS_ERR_INVALID;
  return grad_launch(d_w, d_x, nullptr, nullptr, d_cost, n, 1.f / norm, nullptr, d_loss, d_w, lr, step,
                     d_scratch, (hipStream_t)stream);
}

}  // extern "C"
