// Occupancy field of the export path (compute_voxel_field.cu:25-79).
//
// The reference runs one thread per grid point that walks every Gaussian,
// keeping those whose mean lies in the relaxed box of the point's block:
// O(res^3 * N) box tests.  Here the Gaussians are binned once by the block
// cell their mean falls in (a stable sort of cell ids, so each cell lists its
// Gaussians in index order), and a workgroup takes one block: it walks the
// cells around the block (every Gaussian of the relaxed box lies in them),
// keeps, with the reference's own float comparisons, the Gaussians inside the
// relaxed box, stages them in LDS with their inverse covariance and adds each
// to every grid point of the block.  Sums are in a fixed order (cells in
// z, y, x order, Gaussians by index within a cell): deterministic, but not the
// reference's plain index order, so the parity with it is within rounding.
//
// Grid point (ix, iy, iz): coordinate -1 + 2 i / (res - 1) per axis, block
// (int)((c + 1) / (2 / nb)) per axis (nb + 1 values: the last grid plane
// c = 1 falls into block nb), relaxed box of block b:
// [-1 + b s - s r, -1 + (b + 1) s + s r] with s = 2 / nb, r = relax ratio;
// output occ[ix + iy res + iz res^2].
#pragma once
#include "w2w_device.hpp"

namespace w2w {

struct VoxelGeom {
  int res, nb;
  float block_size, relax;
};

W2W_HD float vox_coord(int i, int res) { return -1.0f + 2.0f * (float)i / (float)(res - 1); }
W2W_HD int vox_block(float c, float block_size) { return (int)((c + 1.0f) / block_size); }
W2W_HD float vox_lo(int b, const VoxelGeom& g) {
  return -1.0f + (float)b * g.block_size - g.block_size * g.relax;
}
W2W_HD float vox_hi(int b, const VoxelGeom& g) {
  return -1.0f + (float)(b + 1) * g.block_size + g.block_size * g.relax;
}

// cell of a mean along one axis: its block, clamped to [0, nb]
W2W_DEV int vox_cell(float m, const VoxelGeom& g) {
  const float q = (m + 1.0f) / g.block_size;
  if (!(q >= 0.0f)) return 0;  // also NaN
  if (q >= (float)g.nb) return g.nb;
  return (int)q;
}

__global__ __launch_bounds__(256) void k_vox_cells(const float* __restrict__ means, uint32_t n, VoxelGeom g,
                                                   uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const uint32_t nc = (uint32_t)g.nb + 1;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t cx = vox_cell(means[3 * i], g), cy = vox_cell(means[3 * i + 1], g);
    const uint32_t cz = vox_cell(means[3 * i + 2], g);
    keys[i] = (cz * nc + cy) * nc + cx;
    vals[i] = i;
  }
}

// [start, end) of every cell in the sorted keys (cells without Gaussians keep
// the (0, 0) they were filled with)
__global__ __launch_bounds__(256) void k_vox_cell_ranges(const uint32_t* __restrict__ keys, uint32_t n,
                                                         uint2* __restrict__ ranges) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const uint32_t k = keys[i];
    if (i == 0 || keys[i - 1] != k) ranges[k].x = i;
    if (i + 1 == n || keys[i + 1] != k) ranges[k].y = i + 1;
  }
}

// first and one-past-last grid index of every block along an axis (the same
// for all three axes); lo filled with res, hi with 0 beforehand
__global__ __launch_bounds__(256) void k_vox_axis(VoxelGeom g, int* __restrict__ lo, int* __restrict__ hi) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < g.res; i += gridDim.x * blockDim.x) {
    int b = vox_block(vox_coord(i, g.res), g.block_size);
    b = b < 0 ? 0 : (b > g.nb ? g.nb : b);
    atomicMin(&lo[b], i);
    atomicMax(&hi[b], i + 1);
  }
}

// compute_3d_gaussian_coefficient_device (compute_voxel_field.cu:6-23): the
// inverse covariance depends on the Gaussian only, so it is formed once per
// Gaussian with the same expressions
struct VoxSplat {
  float mx, my, mz, ia, ib, ic, id, ie, iff, op;
};

W2W_DEV void vox_inverse(float a, float b, float c, float d, float e, float f, float* inv) {
  const float det = a * d * f + 2 * e * c * b - e * e * a - c * c * d - b * b * f;
  const float inv_det = 1.0f / (det + 1e-24f);
  inv[0] = (d * f - e * e) * inv_det;
  inv[1] = (e * c - b * f) * inv_det;
  inv[2] = (e * b - c * d) * inv_det;
  inv[3] = (a * f - c * c) * inv_det;
  inv[4] = (b * c - e * a) * inv_det;
  inv[5] = (a * d - b * b) * inv_det;
}

W2W_DEV float vox_weight(const VoxSplat& q, float x, float y, float z) {
  const float dx = x - q.mx, dy = y - q.my, dz = z - q.mz;
  float power = -0.5f * (dx * dx * q.ia + dy * dy * q.id + dz * dz * q.iff) - dx * dy * q.ib - dx * dz * q.ic -
                dy * dz * q.ie;
  if (power > 0) power = -1e10f;
  return expf(power);
}

constexpr int VOX_THREADS = 256;
constexpr int VOX_PER_THREAD = 8;  // grid points per thread: blocks of up to 2048 points
constexpr int VOX_STAGE = 256;

struct VoxArgs {
  const float* means;
  const float* covs;
  const float* opac;
  const uint32_t* cell_ids;  // Gaussians sorted by cell (index order within a cell)
  const uint2* cell_ranges;
  const int* axis_lo;
  const int* axis_hi;
  VoxelGeom g;
  int reach;                 // cells searched on each side of a block
  float* occ;
};

__global__ __launch_bounds__(VOX_THREADS) void k_vox_field(VoxArgs a) {
  __shared__ VoxSplat s_q[VOX_STAGE];
  __shared__ uint32_t s_wc[VOX_THREADS / 64];
  const VoxelGeom& g = a.g;
  const int nc = g.nb + 1;
  const int blk = blockIdx.x;
  const int bx = blk % nc, by = (blk / nc) % nc, bz = blk / (nc * nc);
  const int x0 = a.axis_lo[bx], x1 = a.axis_hi[bx], y0 = a.axis_lo[by], y1 = a.axis_hi[by];
  const int z0 = a.axis_lo[bz], z1 = a.axis_hi[bz];
  if (x1 <= x0 || y1 <= y0 || z1 <= z0) return;  // no grid point in this block (uniform)
  const int nxv = x1 - x0, nyv = y1 - y0;
  const int nvox = nxv * nyv * (z1 - z0);
  const float vmin_x = vox_lo(bx, g), vmax_x = vox_hi(bx, g);
  const float vmin_y = vox_lo(by, g), vmax_y = vox_hi(by, g);
  const float vmin_z = vox_lo(bz, g), vmax_z = vox_hi(bz, g);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cz0 = max(bz - a.reach, 0), cz1 = min(bz + a.reach, g.nb);
  const int cy0 = max(by - a.reach, 0), cy1 = min(by + a.reach, g.nb);
  const int cx0 = max(bx - a.reach, 0), cx1 = min(bx + a.reach, g.nb);
  // the block's grid points in batches of VOX_THREADS * VOX_PER_THREAD
  for (int vb = 0; vb < nvox; vb += VOX_THREADS * VOX_PER_THREAD) {
    float px[VOX_PER_THREAD], py[VOX_PER_THREAD], pz[VOX_PER_THREAD], acc[VOX_PER_THREAD];
#pragma unroll
    for (int k = 0; k < VOX_PER_THREAD; k++) {
      const int v = vb + (int)threadIdx.x + k * VOX_THREADS;
      const int vx = v % nxv, vy = (v / nxv) % nyv, vz = v / (nxv * nyv);
      px[k] = vox_coord(x0 + vx, g.res);
      py[k] = vox_coord(y0 + vy, g.res);
      pz[k] = vox_coord(z0 + vz, g.res);
      acc[k] = 0.0f;
    }
    const int kmax = min(VOX_PER_THREAD, (nvox - vb - (int)threadIdx.x + VOX_THREADS - 1) / VOX_THREADS);
    for (int cz = cz0; cz <= cz1; cz++)
      for (int cy = cy0; cy <= cy1; cy++)
        for (int cx = cx0; cx <= cx1; cx++) {
          const uint2 r = a.cell_ranges[(cz * nc + cy) * nc + cx];
          for (uint32_t base = r.x; base < r.y; base += VOX_STAGE) {
            const uint32_t i = base + threadIdx.x;
            bool keep = false;
            VoxSplat q;
            if (i < r.y) {
              const uint32_t gi = a.cell_ids[i];
              q.mx = a.means[3 * gi]; q.my = a.means[3 * gi + 1]; q.mz = a.means[3 * gi + 2];
              // compute_voxel_field.cu:55-58
              keep = !(q.mx < vmin_x || q.mx > vmax_x || q.my < vmin_y || q.my > vmax_y || q.mz < vmin_z ||
                       q.mz > vmax_z);
              if (keep) {
                float inv[6];
                vox_inverse(a.covs[6 * gi], a.covs[6 * gi + 1], a.covs[6 * gi + 2], a.covs[6 * gi + 3],
                            a.covs[6 * gi + 4], a.covs[6 * gi + 5], inv);
                q.ia = inv[0]; q.ib = inv[1]; q.ic = inv[2]; q.id = inv[3]; q.ie = inv[4]; q.iff = inv[5];
                q.op = a.opac[gi];
              }
            }
            // stage the kept Gaussians of this chunk in index order
            const uint64_t m = __ballot(keep);
            const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            if (lane == 0) s_wc[wave] = (uint32_t)__popcll(m);
            __syncthreads();
            uint32_t before = 0, total = 0;
#pragma unroll
            for (int w = 0; w < VOX_THREADS / 64; w++) {
              const uint32_t c = s_wc[w];
              before += w < wave ? c : 0u;
              total += c;
            }
            if (keep) s_q[before + rank] = q;
            __syncthreads();
            for (uint32_t j = 0; j < total; j++) {
              const VoxSplat qj = s_q[j];
#pragma unroll
              for (int k = 0; k < VOX_PER_THREAD; k++)
                if (k < kmax) acc[k] += qj.op * vox_weight(qj, px[k], py[k], pz[k]);
            }
            __syncthreads();  // s_q and s_wc are rewritten by the next chunk
          }
        }
    const size_t res = (size_t)g.res;
#pragma unroll
    for (int k = 0; k < VOX_PER_THREAD; k++) {
      const int v = vb + (int)threadIdx.x + k * VOX_THREADS;
      if (k < kmax && v < nvox) {
        const int vx = v % nxv, vy = (v / nxv) % nyv, vz = v / (nxv * nyv);
        a.occ[(size_t)(x0 + vx) + (size_t)(y0 + vy) * res + (size_t)(z0 + vz) * res * res] = acc[k];
      }
    }
  }
}

}  // namespace w2w
