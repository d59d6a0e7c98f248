// Step-level kernels of the batched trainer: parameter activation and its
// backward, the L2 image loss, and Adam with hyper-parameters in device memory
// (so a captured HIP graph picks up the learning-rate schedule at replay).
//
// Activation as Renderer.py:164-167: opacity = sigmoid(logit), scale =
// exp(svec), rotation = q / max(|q|, 1e-12) (torch.nn.functional.normalize).
// Loss as the benchmark workloads define it: per frame, the mean over 3*H*W of
// (clamp(colour, 0, 1) - target)^2, summed over the frames that take part in
// the backward (GaussianTrainer.py:86 clamps the render before the loss).
#pragma once
#include "w2w_adam.hpp"
#include "w2w_backward.hpp"
#include "w2w_device.hpp"
#include "w2w_forward.hpp"

namespace w2w {

// activation (and its backward) over Gaussian-row ranges of a resident store
struct RowRange {
  uint64_t start, off, len;
};

W2W_DEV uint32_t row_range_of(const RowRange* r, uint32_t n, uint64_t w) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (r[mid].off <= w) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Work index w -> row of the store: identity without a range table.
W2W_DEV uint64_t row_of(const RowRange* r, uint32_t n, uint64_t w) {
  if (!r) return w;
  const RowRange& x = r[row_range_of(r, n, w)];
  return x.start + (w - x.off);
}

__global__ __launch_bounds__(256) void k_activate(const float* __restrict__ svec,
                                                  const float* __restrict__ logit,
                                                  const float* __restrict__ quat, uint64_t G,
                                                  const RowRange* __restrict__ ranges,
                                                  uint32_t n_ranges, float* __restrict__ scales,
                                                  float* __restrict__ opac,
                                                  float* __restrict__ rots) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < G; w += stride) {
    const uint64_t i = row_of(ranges, n_ranges, w);
#pragma unroll
    for (int k = 0; k < 3; k++) scales[3 * i + k] = expf(svec[3 * i + k]);
    opac[i] = 1.0f / (1.0f + expf(-logit[i]));
    const float q0 = quat[4 * i], q1 = quat[4 * i + 1], q2 = quat[4 * i + 2], q3 = quat[4 * i + 3];
    const float n = fmaxf(__builtin_sqrtf(((q0 * q0 + q1 * q1) + q2 * q2) + q3 * q3), 1e-12f);
    rots[4 * i] = q0 / n;
    rots[4 * i + 1] = q1 / n;
    rots[4 * i + 2] = q2 / n;
    rots[4 * i + 3] = q3 / n;
  }
}

// d svec = d scale * scale; d logit = d o * o (1 - o);
// d q = (d qhat - qhat (qhat . d qhat)) / |q|  (|q| > 1e-12)
__global__ __launch_bounds__(256) void k_activate_bwd(
    const float* __restrict__ scales, const float* __restrict__ opac,
    const float* __restrict__ quat, const float* __restrict__ g_scales,
    const float* __restrict__ g_opac, const float* __restrict__ g_rots, uint64_t G,
    const RowRange* __restrict__ ranges, uint32_t n_ranges, float* __restrict__ g_svec,
    float* __restrict__ g_logit, float* __restrict__ g_quat) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t w = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; w < G; w += stride) {
    const uint64_t i = row_of(ranges, n_ranges, w);
#pragma unroll
    for (int k = 0; k < 3; k++) g_svec[3 * i + k] = g_scales[3 * i + k] * scales[3 * i + k];
    const float o = opac[i];
    g_logit[i] = g_opac[i] * (o * (1.0f - o));
    const float q0 = quat[4 * i], q1 = quat[4 * i + 1], q2 = quat[4 * i + 2], q3 = quat[4 * i + 3];
    const float nr = __builtin_sqrtf(((q0 * q0 + q1 * q1) + q2 * q2) + q3 * q3);
    const float n = fmaxf(nr, 1e-12f);
    const float h0 = q0 / n, h1 = q1 / n, h2 = q2 / n, h3 = q3 / n;
    const float d0 = g_rots[4 * i], d1 = g_rots[4 * i + 1], d2 = g_rots[4 * i + 2], d3 = g_rots[4 * i + 3];
    if (nr > 1e-12f) {
      const float dot = ((h0 * d0 + h1 * d1) + h2 * d2) + h3 * d3;
      g_quat[4 * i] = (d0 - h0 * dot) / n;
      g_quat[4 * i + 1] = (d1 - h1 * dot) / n;
      g_quat[4 * i + 2] = (d2 - h2 * dot) / n;
      g_quat[4 * i + 3] = (d3 - h3 * dot) / n;
    } else {
      g_quat[4 * i] = d0 / n;
      g_quat[4 * i + 1] = d1 / n;
      g_quat[4 * i + 2] = d2 / n;
      g_quat[4 * i + 3] = d3 / n;
    }
  }
}

// L2 loss over the packed frame outputs; dL/dcolor for every pixel of every
// frame (zero for frames without the backward flag), per-frame loss sums
// accumulated into loss[frame] (pre-zeroed).
// L2 image loss, one wave per tile: dL/dcolour of the tile's pixels (zero
// where the clamp to [0, 1] is active or the frame has no backward) and the
// tile's share of its frame's mean squared error, summed in a fixed order
// (each lane's pixels in order, then DPP across the wave).  A wave keeps
// several pixels' loads in flight and needs no barrier.
__global__ __launch_bounds__(256) void k_l2_loss_tiles(const void* __restrict__ table,
                                                       uint32_t scenes_cap, uint32_t tiles_cap,
                                                       int ts, const float* __restrict__ color,
                                                       const float* __restrict__ target,
                                                       float* __restrict__ dL_dcolor,
                                                       float* __restrict__ tile_loss) {
  const BatchHeader* hdr = static_cast<const BatchHeader*>(table);
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t T = min(hdr->n_tiles, tiles_cap);
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < T; t += waves) {
    const FrameDesc& fd = frames[tile_frame(frames, hdr->n_frames, t)];
    const int lt = (int)(t - fd.tile_off);
    const int x0 = (lt % fd.tiles_x) * ts, y0 = (lt / fd.tiles_x) * ts;
    const uint64_t HW = (uint64_t)fd.W * fd.H;
    const float w = 1.0f / (3.0f * (float)HW);
    const bool on = (fd.flags & FRAME_BACKWARD) != 0;
    float part = 0.0f;
#pragma unroll 4
    for (int k = lane; k < ts * ts; k += 64) {
      const int px = x0 + k % ts, py = y0 + k / ts;
      if (px >= fd.W || py >= fd.H) continue;
      const uint64_t lp = (uint64_t)py * fd.W + px;
      float sq = 0.0f;
#pragma unroll
      for (int ch = 0; ch < 3; ch++) {
        const uint64_t e = 3 * fd.pix_off + ch * HW + lp;
        const float c = color[e];
        const float cc = fminf(fmaxf(c, 0.0f), 1.0f);
        const float d = cc - target[fd.target_off + ch * HW + lp];
        sq += d * d;
        dL_dcolor[e] = (on && c >= 0.0f && c <= 1.0f) ? 2.0f * d * w : 0.0f;
      }
      part += sq * w;
    }
    part = wave_reduce_dpp(part);
    if (lane == 63) tile_loss[t] = on ? part : 0.0f;
  }
}

// Per-frame loss: a workgroup per frame sums its tiles' shares in a fixed
// order (thread t the tiles t, t + 256, ... in turn, then the threads' sums
// in a fixed tree through LDS).  A single wave walking 4096 tiles was a chain
// of dependent loads (0.017 ms for four frames).
__global__ __launch_bounds__(256) void k_loss_frames(const void* __restrict__ table,
                                                     uint32_t scenes_cap, uint32_t frames_cap,
                                                     const float* __restrict__ tile_loss,
                                                     float* __restrict__ loss) {
  __shared__ float s_part[256];
  const BatchHeader* hdr = static_cast<const BatchHeader*>(table);
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t nf = min(hdr->n_frames, frames_cap);
  const int tid = (int)threadIdx.x;
  for (uint32_t f = blockIdx.x; f < nf; f += gridDim.x) {  // uniform
    const FrameDesc& fd = frames[f];
    const uint32_t n = (uint32_t)(fd.tiles_x * fd.tiles_y);
    float part = 0.0f;
    for (uint32_t k = (uint32_t)tid; k < n; k += 256) part += tile_loss[fd.tile_off + k];
    s_part[tid] = part;
    __syncthreads();
#pragma unroll
    for (int w = 128; w >= 1; w >>= 1) {
      if (tid < w) s_part[tid] += s_part[tid + w];
      __syncthreads();
    }
    if (tid == 0) loss[f] = s_part[0];
    __syncthreads();  // s_part is rewritten for the next frame
  }
}

// Adam with hyper-parameters in device memory: hyper = {neg_step[8] (-lr_k /
// (1 - b1^t) per group), inv_bc2_sqrt, 1 - b1, b2, 1 - b2, eps}.  A group may
// alternate two learning rates along its elements: element e of the group
// uses neg_step[k] when (e mod period) < split and neg_step2[k] otherwise
// (the SH block of a Gaussian: DC coefficients and the rest).
struct AdamGraphArgs {
  float* p;
  const float* g;
  float* m;
  float* v;
  uint64_t n;
  int nseg;
  uint64_t seg_end[ADAM_MAX_SEGMENTS];
  uint32_t period[ADAM_MAX_SEGMENTS];
  uint32_t split[ADAM_MAX_SEGMENTS];
  int lr2[ADAM_MAX_SEGMENTS];  // index into hyper for the second rate, -1 if none
  const float* hyper;
  const uint32_t* skip_if;  // optional: no update when *skip_if != 0 (an overflowed step)
  uint32_t* sticky;         // optional: ORed with 1 when the update was skipped
};

// Step size of element i of the store: its group's rate, or the group's
// second rate past `split` within each period.  `r` carries (i - group
// start) mod period from the previous element of the same thread.
struct AdamRate {
  int s = -1;
  uint32_t r = 0;
  W2W_DEV float at(const AdamGraphArgs& a, uint64_t i) {
    if (s >= 0 && (s == a.nseg - 1 || i < a.seg_end[s])) {
      if (a.lr2[s] >= 0 && ++r == a.period[s]) r = 0;
    } else {
      s = 0;
#pragma unroll
      for (int k = 0; k < ADAM_MAX_SEGMENTS - 1; k++)
        if (k < a.nseg - 1 && i >= a.seg_end[k]) s = k + 1;
      if (a.lr2[s] >= 0) {
        const uint64_t start = s == 0 ? 0 : a.seg_end[s - 1];
        r = (uint32_t)((i - start) % a.period[s]);
      }
    }
    return (a.lr2[s] >= 0 && r >= a.split[s]) ? a.hyper[a.lr2[s]] : a.hyper[s];
  }
};

// VEC: four consecutive elements per thread with 16-byte accesses (all four
// arrays 16-byte aligned); the group and its period position are found once
// per thread and then advanced.
template <bool VEC>
__global__ __launch_bounds__(256) void k_adam_hyper(AdamGraphArgs a) {
  if (a.skip_if) {
    const uint32_t skip = *a.skip_if;
    if (skip) {
      if (a.sticky && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(a.sticky, 1u);
      return;
    }
  }
  const float inv_bc2_sqrt = a.hyper[16], one_m_b1 = a.hyper[17], b2 = a.hyper[18];
  const float one_m_b2 = a.hyper[19], eps = a.hyper[20];
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t n4 = VEC ? a.n / 4 : 0;
  for (uint64_t q = t0; q < n4; q += stride) {
    float4 P = reinterpret_cast<float4*>(a.p)[q];
    const float4 G = reinterpret_cast<const float4*>(a.g)[q];
    float4 M = reinterpret_cast<float4*>(a.m)[q];
    float4 V = reinterpret_cast<float4*>(a.v)[q];
    AdamRate rate;
    const uint64_t i0 = 4 * q;
    adam_one(P.x, G.x, M.x, V.x, rate.at(a, i0), inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
    adam_one(P.y, G.y, M.y, V.y, rate.at(a, i0 + 1), inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
    adam_one(P.z, G.z, M.z, V.z, rate.at(a, i0 + 2), inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
    adam_one(P.w, G.w, M.w, V.w, rate.at(a, i0 + 3), inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
    reinterpret_cast<float4*>(a.p)[q] = P;
    reinterpret_cast<float4*>(a.m)[q] = M;
    reinterpret_cast<float4*>(a.v)[q] = V;
  }
  for (uint64_t i = 4 * n4 + t0; i < a.n; i += stride) {  // the last n mod 4 elements
    AdamRate rate;
    float p = a.p[i], m = a.m[i], v = a.v[i];
    adam_one(p, a.g[i], m, v, rate.at(a, i), inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
    a.p[i] = p;
    a.m[i] = m;
    a.v[i] = v;
  }
}

// ---- range tables for stores with resident scenes ----
// A range is a contiguous run of the flat store; `off` is its start in the
// concatenation of all ranges of the table (the work index space).
struct ParamRange {
  uint64_t start, off, len;
  int32_t hyper;            // hyper block: {neg_step, neg_step2, inv_bc2_sqrt, 0}
  uint32_t period, split;   // second rate when (e mod period) >= split (period 0: none)
  uint32_t pad;
};

W2W_DEV uint32_t range_of(const ParamRange* r, uint32_t n, uint64_t w) {
  uint32_t lo = 0, hi = n - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (r[mid].off <= w) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Each thread takes four work elements 64 apart inside a 256-element block
// (loads stay coalesced across the wave): one range search for the first,
// then the range only advances (ranges are long runs of the store).
__global__ __launch_bounds__(256) void k_adam_ranges(float* __restrict__ p, const float* __restrict__ g,
                                                     float* __restrict__ m, float* __restrict__ v,
                                                     const ParamRange* __restrict__ ranges,
                                                     uint32_t n_ranges, uint64_t total,
                                                     const float* __restrict__ hyper, float one_m_b1,
                                                     float b2, float one_m_b2, float eps) {
  const uint64_t nblk = (total + 255) / 256;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (uint64_t blk = (uint64_t)blockIdx.x * 4 + wave; blk < nblk; blk += (uint64_t)gridDim.x * 4) {
    const uint64_t w0 = blk * 256 + lane;
    if (w0 >= total) continue;
    uint32_t ri = range_of(ranges, n_ranges, w0);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const uint64_t w = w0 + 64 * j;
      if (w >= total) break;
      while (w >= ranges[ri].off + ranges[ri].len) ri++;
      const ParamRange& r = ranges[ri];
      const uint64_t e = w - r.off;
      const uint64_t i = r.start + e;
      const float* h = hyper + 4 * r.hyper;
      float ns = h[0];
      if (r.period) {
        const uint32_t pos = (e >> 32) ? (uint32_t)(e % r.period) : (uint32_t)e % r.period;
        if (pos >= r.split) ns = h[1];
      }
      const float gi = g[i];
      float mi = m[i];
      float vi = v[i];
      mi = mi + one_m_b1 * (gi - mi);
      vi = vi * b2 + (one_m_b2 * gi) * gi;
      const float denom = __builtin_sqrtf(vi) * h[2] + eps;
      p[i] = p[i] + ns * (mi / denom);
      m[i] = mi;
      v[i] = vi;
    }
  }
}

}  // namespace w2w
