// Fused Adam over a flat parameter store (the trainer's six parameter groups
// laid out back to back: means, sh_dc, sh_ac, opacity, svec, quaternion).
//
// Semantics of torch.optim.Adam(amsgrad=False, weight_decay=0) as the
// reference configures it (gaussians_handler.py:40, eps=1e-15):
//   m <- m + (1-b1)(g - m);  v <- b2 v + (1-b2) g^2
//   p <- p - lr/(1-b1^t) * m / (sqrt(v)/sqrt(1-b2^t) + eps)
// with a learning rate per group (optimizer_config.json) and one step count
// for the store.  One pass: 28 bytes per parameter (read p,g,m,v; write
// p,m,v), HBM-bound.
#pragma once
#include "w2w_device.hpp"

namespace w2w {

constexpr int ADAM_MAX_SEGMENTS = 8;

struct AdamArgs {
  float* p;
  const float* g;
  float* m;
  float* v;
  uint64_t n;
  int nseg;
  uint64_t seg_end[ADAM_MAX_SEGMENTS];  // exclusive end of each group
  float neg_step[ADAM_MAX_SEGMENTS];    // -lr / (1 - b1^t) per group
  float inv_bc2_sqrt;                   // 1 / sqrt(1 - b2^t)
  float one_m_b1, b2, one_m_b2, eps;
};

// One element's update with the step size ns = -lr / (1 - b1^t) and
// inv_bc2_sqrt = 1 / sqrt(1 - b2^t) (the store-wide kernels and the SH
// backward's fused step share it).
W2W_DEV void adam_one(float& p, float g, float& m, float& v, float ns, float inv_bc2_sqrt, float one_m_b1,
                      float b2, float one_m_b2, float eps) {
  m = m + one_m_b1 * (g - m);
  v = v * b2 + (one_m_b2 * g) * g;
  const float denom = __builtin_sqrtf(v) * inv_bc2_sqrt + eps;
  p = p + ns * (m / denom);
}

__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    int s = 0;
#pragma unroll
    for (int k = 0; k < ADAM_MAX_SEGMENTS - 1; k++)
      if (k < a.nseg - 1 && i >= a.seg_end[k]) s = k + 1;
    const float g = a.g[i];
    float m = a.m[i];
    float v = a.v[i];
    m = m + a.one_m_b1 * (g - m);
    v = v * a.b2 + (a.one_m_b2 * g) * g;
    const float denom = __builtin_sqrtf(v) * a.inv_bc2_sqrt + a.eps;
    a.p[i] = a.p[i] + a.neg_step[s] * (m / denom);
    a.m[i] = m;
    a.v[i] = v;
  }
}

}  // namespace w2w
