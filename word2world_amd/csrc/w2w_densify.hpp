// On-device adaptive density control: one adc_cycle of the reference trainer
// (gaussians_handler.py:185-194) over the flat parameter store, with the Adam
// state remap of OptimizerManager.py:23-45.
//
// The reference runs three phases, each a torch mask + cat/index of every
// parameter and moment tensor:
//   clone  (Densification.py:33-52): rows with |accum/(count+1e-5)| >= thr and
//          any scale <= clone_max are appended as copies (zero moments, zero
//          statistics);
//   split  (Densification.py:56-107), on the rows after cloning: rows with
//          |g| >= thr and any scale > split_min get `factor` children,
//          appended block by block (block k holds child k of every split row,
//          in row order), mean = R(q) (scale * z) + mean with z ~ N(0, 1),
//          svec = log(scale / (factor * shrink)), other parameters copied,
//          zero moments; the split rows are then removed;
//   prune  (Densification.py:167-187), on the rows after splitting: rows with
//          sigmoid(logit) < min_opacity or max scale > prune_max are removed.
//
// Here the three phases collapse into one stream compaction over the virtual
// sequence [original rows | clones | child block 0 | ... | block factor-1]:
// every row's fate (kept, cloned and kept, split, children kept) is decided
// by one flags kernel from the row alone, because a clone and the children
// of a row carry the row's own opacity and a scale derived from its own.
// A block-count scan then gives every kept row its output position, and one
// copy kernel moves parameters and moments, coalesced along each group.
//
// The store is described group by group (DensStore): the trainer's flat
// store [means | sh_dc | sh_ac | opacity | svec | quaternion], the batched
// store [means | sh | opacity | svec | quaternion], or one scene's rows of it.
#pragma once
#include "w2w_device.hpp"

namespace w2w {

struct DensifyCfg {
  float grad_thresh;      // on |accum / (count + 1e-5)|
  float min_opacity;      // prune: sigmoid(logit) < min_opacity
  float clone_scale_max;  // clone: any scale <= this (0.04)
  float split_scale_min;  // split: any scale > this (4)
  float prune_scale_max;  // prune: max scale > this (0.1 * 4)
  float split_div;        // children's scale divisor (factor * 0.8)
  uint32_t factor;        // children per split row (2)
};

// per-row flags (one byte per original row)
enum : uint32_t {
  DF_KEEP_O = 1u,     // the original row survives
  DF_KEEP_C = 2u,     // it is cloned and the clone survives
  DF_SPLIT_O = 4u,    // the original row is split
  DF_SPLIT_C = 8u,    // its clone is split (only when thr <= 0: a clone's statistics are zero)
  DF_CH_O = 16u,      // the original is split and its children survive the prune
  DF_CH_C = 32u,      // the clone is split and its children survive
  DF_CLONE = 64u,     // the row is cloned
};
constexpr int DF_NFIELDS = 7;
constexpr int DENS_BLOCK = 256;
constexpr uint32_t DENS_NONE = 0xFFFFFFFFu;

// counts[] written for the host (int64): clones, splits, rows removed by the
// prune phase, rows out, rows in; fields[] totals of the per-row flags
enum { DC_CLONE = 0, DC_SPLIT = 1, DC_PRUNE = 2, DC_OUT = 3, DC_IN = 4, DC_N = 8 };

constexpr int DENS_MAX_GROUPS = 8;

// A store (or one scene's rows of a store): per group a base pointer to its
// row 0 and the floats per row; which groups hold the means, opacity logit,
// log-scales and quaternion.
struct DensStore {
  float* p[DENS_MAX_GROUPS];
  float* m[DENS_MAX_GROUPS];
  float* v[DENS_MAX_GROUPS];
  uint32_t w[DENS_MAX_GROUPS];
  int ng, means, opacity, svec, quat;
};

// torch.max over a row propagates NaN
W2W_DEV float max3_nan(float x, float y, float z) {
  if (x != x || y != y || z != z) return __builtin_nanf("");
  return fmaxf(fmaxf(x, y), z);
}

struct DensFlagsArgs {
  DensStore st;
  uint64_t n;
  const float* accum;
  const float* count;
  DensifyCfg c;
  uint8_t* flags;
  uint32_t* blockcnt;  // [blocks][8]
};

__global__ __launch_bounds__(DENS_BLOCK) void k_dens_flags(DensFlagsArgs A) {
  __shared__ uint32_t s_cnt[DENS_BLOCK / 64][8];
  const uint64_t i = (uint64_t)blockIdx.x * DENS_BLOCK + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t f = 0;
  if (i < A.n) {
    const DensStore& st = A.st;
    const DensifyCfg& c = A.c;
    const float g = A.accum[i] / (A.count[i] + 1e-5f);
    const bool gm = fabsf(g) >= c.grad_thresh;
    const bool gm0 = 0.0f >= c.grad_thresh;  // statistics of a fresh clone
    float e[3], ec[3];
    bool small = false, big = false;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      e[k] = expf(st.p[st.svec][3 * i + k]);
      small |= e[k] <= c.clone_scale_max;
      big |= e[k] > c.split_scale_min;
      ec[k] = expf(logf(e[k] / c.split_div));  // a child's scale as the prune sees it
    }
    const float op = 1.0f / (1.0f + expf(-st.p[st.opacity][i]));
    const bool low = op < c.min_opacity;
    const bool prune = low || max3_nan(e[0], e[1], e[2]) > c.prune_scale_max;
    const bool prune_ch = low || max3_nan(ec[0], ec[1], ec[2]) > c.prune_scale_max;
    const bool clone = gm && small;
    const bool split_o = gm && big;
    const bool split_c = clone && gm0 && big;
    if (!split_o && !prune) f |= DF_KEEP_O;
    if (clone && !split_c && !prune) f |= DF_KEEP_C;
    if (split_o) f |= DF_SPLIT_O;
    if (split_c) f |= DF_SPLIT_C;
    if (split_o && !prune_ch) f |= DF_CH_O;
    if (split_c && !prune_ch) f |= DF_CH_C;
    if (clone) f |= DF_CLONE;
    A.flags[i] = (uint8_t)f;
  }
#pragma unroll
  for (int b = 0; b < DF_NFIELDS; b++) {
    const uint64_t m = __ballot((f >> b) & 1u);
    if (lane == 0) s_cnt[wave][b] = (uint32_t)__popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    uint32_t t = 0;
    if (threadIdx.x < DF_NFIELDS)
      for (int w = 0; w < DENS_BLOCK / 64; w++) t += s_cnt[w][threadIdx.x];
    A.blockcnt[(uint64_t)blockIdx.x * 8 + threadIdx.x] = t;
  }
}

// One workgroup: wave b scans field b of the block counts in place
// (exclusive) and writes the field total; then the host counts.
__global__ __launch_bounds__(512) void k_dens_scan(uint32_t* blockcnt, uint32_t nblocks,
                                                   uint32_t* totals, long long* counts, uint64_t n,
                                                   uint32_t factor) {
  __shared__ uint32_t s_tot[8];
  const int lane = threadIdx.x & 63, field = threadIdx.x >> 6;
  if (field < DF_NFIELDS) {
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nblocks; b0 += 64) {
      const uint32_t b = b0 + lane;
      const uint32_t v = b < nblocks ? blockcnt[(uint64_t)b * 8 + field] : 0u;
      const uint32_t incl = wave_incl_scan_u32(v, lane);
      if (b < nblocks) blockcnt[(uint64_t)b * 8 + field] = carry + incl - v;
      carry += __shfl(incl, 63, 64);
    }
    if (lane == 0) {
      totals[field] = carry;
      s_tot[field] = carry;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const uint64_t keep_o = s_tot[0], keep_c = s_tot[1], split = (uint64_t)s_tot[2] + s_tot[3];
    const uint64_t kept_ch = (uint64_t)s_tot[4] + s_tot[5], clones = s_tot[6];
    const uint64_t out = keep_o + keep_c + (uint64_t)factor * kept_ch;
    counts[DC_CLONE] = (long long)clones;
    counts[DC_SPLIT] = (long long)split;
    // rows after the split phase minus rows out
    counts[DC_PRUNE] = (long long)(n + clones + (uint64_t)factor * split - split - out);
    counts[DC_OUT] = (long long)out;
    counts[DC_IN] = (long long)n;
    counts[5] = counts[6] = counts[7] = 0;
  }
}

struct DensApplyArgs {
  DensStore in;
  uint64_t n;
  const uint8_t* flags;
  const uint32_t* blockoff;  // exclusive block offsets [blocks][8]
  const uint32_t* totals;
  const float* normals;      // [rows][3] standard normal samples
  uint64_t normal_rows;
  DensifyCfg c;
  DensStore out;
  uint64_t n_out;
};

__global__ __launch_bounds__(DENS_BLOCK) void k_dens_apply(DensApplyArgs A) {
  __shared__ uint32_t s_cnt[DENS_BLOCK / 64][8];
  __shared__ uint32_t s_do[DENS_BLOCK], s_dc[DENS_BLOCK], s_cho[DENS_BLOCK], s_chc[DENS_BLOCK];
  __shared__ uint32_t s_so[DENS_BLOCK], s_sc[DENS_BLOCK];
  const uint32_t T0 = A.totals[0], T1 = A.totals[1], T2 = A.totals[2], T3 = A.totals[3];
  const uint32_t T4 = A.totals[4], T5 = A.totals[5];
  const uint32_t S = T2 + T3, KC = T4 + T5;
  // the output layout depends on the row count: write nothing unless the
  // caller sized the output from this plan
  if ((uint64_t)T0 + T1 + (uint64_t)A.c.factor * KC != A.n_out) return;
  const uint64_t r0 = (uint64_t)blockIdx.x * DENS_BLOCK;
  const uint64_t i = r0 + threadIdx.x;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t f = i < A.n ? A.flags[i] : 0u;
  const uint64_t lt = (1ull << lane) - 1ull;
  uint32_t rank[DF_NFIELDS];
#pragma unroll
  for (int b = 0; b < DF_NFIELDS; b++) {
    const uint64_t m = __ballot((f >> b) & 1u);
    rank[b] = (uint32_t)__popcll(m & lt);
    if (lane == 0) s_cnt[wave][b] = (uint32_t)__popcll(m);
  }
  __syncthreads();
#pragma unroll
  for (int b = 0; b < DF_NFIELDS; b++) {
    uint32_t pre = A.blockoff[(uint64_t)blockIdx.x * 8 + b];
    for (int w = 0; w < wave; w++) pre += s_cnt[w][b];
    rank[b] += pre;
  }
  const int t = threadIdx.x;
  s_do[t] = (f & DF_KEEP_O) ? rank[0] : DENS_NONE;
  s_dc[t] = (f & DF_KEEP_C) ? T0 + rank[1] : DENS_NONE;
  s_cho[t] = (f & DF_CH_O) ? T0 + T1 + rank[4] : DENS_NONE;
  s_chc[t] = (f & DF_CH_C) ? T0 + T1 + T4 + rank[5] : DENS_NONE;
  s_so[t] = rank[2];       // split rank of the original among all splits
  s_sc[t] = T2 + rank[3];  // split rank of the clone
  __syncthreads();

  const DensStore &I = A.in, &O = A.out;
  const uint32_t rows = (uint32_t)min<uint64_t>(DENS_BLOCK, A.n - r0);
  const uint32_t F = A.c.factor;
  for (int g = 0; g < I.ng; g++) {
    const uint32_t w = I.w[g];
    if (w == 0) continue;
    const float *ip = I.p[g] + r0 * w, *im = I.m[g] + r0 * w, *iv = I.v[g] + r0 * w;
    float *op = O.p[g], *om = O.m[g], *ov = O.v[g];
    const bool derived = (g == I.means || g == I.svec);  // children's values are computed below
    for (uint32_t e = threadIdx.x; e < rows * w; e += DENS_BLOCK) {
      const uint32_t r = e / w, col = e - r * w;
      const float x = ip[e];
      const uint32_t d_o = s_do[r], d_c = s_dc[r], cho = s_cho[r], chc = s_chc[r];
      if (d_o != DENS_NONE) {
        op[(uint64_t)d_o * w + col] = x;
        om[(uint64_t)d_o * w + col] = im[e];
        ov[(uint64_t)d_o * w + col] = iv[e];
      }
      if (d_c != DENS_NONE) {
        op[(uint64_t)d_c * w + col] = x;
        om[(uint64_t)d_c * w + col] = 0.0f;
        ov[(uint64_t)d_c * w + col] = 0.0f;
      }
      for (uint32_t k = 0; k < F; k++) {
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const uint32_t base = h ? chc : cho;
          if (base == DENS_NONE) continue;
          const uint64_t o = (uint64_t)(base + k * KC) * w + col;
          if (!derived) op[o] = x;
          om[o] = 0.0f;
          ov[o] = 0.0f;
        }
      }
    }
  }

  // children's means and svec (Densification.py:68-96)
  if (i < A.n && (s_cho[t] != DENS_NONE || s_chc[t] != DENS_NONE)) {
    const float* q = I.p[I.quat] + 4 * i;
    const float qw0 = q[0], qx0 = q[1], qy0 = q[2], qz0 = q[3];
    const float nrm = __builtin_sqrtf(qw0 * qw0 + qx0 * qx0 + qy0 * qy0 + qz0 * qz0);
    const float qw = qw0 / nrm, qx = qx0 / nrm, qy = qy0 / nrm, qz = qz0 / nrm;
    float R[9];
    R[0] = 1.0f - 2.0f * (qy * qy + qz * qz);
    R[1] = 2.0f * (qx * qy - qw * qz);
    R[2] = 2.0f * (qx * qz + qw * qy);
    R[3] = 2.0f * (qx * qy + qw * qz);
    R[4] = 1.0f - 2.0f * (qx * qx + qz * qz);
    R[5] = 2.0f * (qy * qz - qw * qx);
    R[6] = 2.0f * (qx * qz - qw * qy);
    R[7] = 2.0f * (qy * qz + qw * qx);
    R[8] = 1.0f - 2.0f * (qx * qx + qy * qy);
    float sd[3], mu[3], sv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      sd[k] = expf(I.p[I.svec][3 * i + k]);
      mu[k] = I.p[I.means][3 * i + k];
      sv[k] = logf(sd[k] / A.c.split_div);
    }
    for (uint32_t k = 0; k < F; k++) {
      for (int h = 0; h < 2; h++) {
        const uint32_t base = h ? s_chc[t] : s_cho[t];
        if (base == DENS_NONE) continue;
        const uint64_t zr = (uint64_t)k * S + (h ? s_sc[t] : s_so[t]);
        if (zr >= A.normal_rows) continue;  // too few samples: the host checks first
        const float s0 = sd[0] * A.normals[3 * zr + 0];
        const float s1 = sd[1] * A.normals[3 * zr + 1];
        const float s2 = sd[2] * A.normals[3 * zr + 2];
        const uint64_t row = base + (uint64_t)k * KC;
#pragma unroll
        for (int c = 0; c < 3; c++) {
          const float rs = (R[3 * c] * s0 + R[3 * c + 1] * s1) + R[3 * c + 2] * s2;
          O.p[O.means][3 * row + c] = rs + mu[c];
          O.p[O.svec][3 * row + c] = sv[c];
        }
      }
    }
  }
}

// Densification statistics (gaussians_handler.py:116-118, collected per
// rendered view with update_filter = radii > 0, Renderer.py:172): for every
// Gaussian of the batch and every frame of its scene with the backward flag,
// in frame order, where the item's radius is positive: count += 1 and
// accum += |dL/dmeans2D[item, 0:2]|.  Items of a scene are Gaussian-major
// (item_off + g * n_views + view); accum/count rows are g_base + g.
__global__ __launch_bounds__(256) void k_dens_stats(const void* table, uint32_t scenes_cap,
                                                    const float* __restrict__ m2d,
                                                    const int32_t* __restrict__ radii,
                                                    float* __restrict__ accum,
                                                    float* __restrict__ count) {
  const BatchHeader* h = static_cast<const BatchHeader*>(table);
  const uint32_t n_g = h->n_gaussians, n_scenes = h->n_scenes;
  const SceneDesc* scenes = table_scenes(table);
  const FrameDesc* frames = table_frames(table, scenes_cap);
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < n_g; g += gridDim.x * blockDim.x) {
    const uint32_t sc = upper_index(n_scenes, g, [&](uint32_t k) { return scenes[k].g_off; });
    const SceneDesc& sd = scenes[sc];
    const uint32_t local = g - sd.g_off, row = sd.g_base + local;
    float a = accum[row], c = count[row];
    for (uint32_t v = 0; v < sd.n_views; v++) {
      if (!(frames[sd.frame_off + v].flags & FRAME_BACKWARD)) continue;
      const uint64_t item = (uint64_t)sd.item_off + (uint64_t)local * sd.n_views + v;
      if (radii[item] > 0) {
        const float x = m2d[3 * item], y = m2d[3 * item + 1];
        c += 1.0f;
        a += __builtin_sqrtf(x * x + y * y);
      }
    }
    accum[row] = a;
    count[row] = c;
  }
}

}  // namespace w2w
