// Forward rasterizer kernels for gfx950, over a batch of frames.
//
// Stage order (gaussianRasterizer.cu:118-280 restructured for HBM traffic):
//   k_preprocess   projection.cu:7-154 + ADRculling.cu:7-109 +
//                  gaussianForwardPass.cu:172-264 fused, one thread per item
//                  ((frame, Gaussian) pair)
//   depth sort     (depth bits, item) over all items with the onesweep sort
//   k_scan         tile counts gathered in depth order, single-pass decoupled
//                  look-back exclusive scan -> offsets, K, each item's offset
//   k_emit         instance keys in depth order, block-level load balancing
//                  (prepareSort.cu:239-301 emitted one Gaussian per thread
//                  with a serial tile loop)
//   tile sort      stable sort of the 32-bit global tile ids
//   k_ranges       gaussianTiles.cu:6-33
//   k_blend        pixelGaussiansFull.cu:3-98 / pixelGaussiansDev.cuh:20-79
//
// Depth-sliced binning (an option of the batched path; same outputs): the
// emission, sort and ranges run first over the near lists (the nearest items,
// plus the far items on the tiles the previous forward left unfinished), the
// blend stores unfinished pixels' state (BLEND_NEAR), and only the tiles left
// open get far lists (k_count_tiles, k_scan, k_emit_far, sort, ranges) and a
// resumed blend (BLEND_FAR).  A tile's lists, near then far, are its full list
// in depth order cut where every pixel is done.
//
// Sorting the depth keys first and then stably sorting the tile ids gives
// exactly the order of the reference's stable sort of 64-bit keys
// (tile << 32 | depth bits) emitted in Gaussian order: instances end up
// ordered by (tile, depth bits, Gaussian id), per frame.
//
// Every count that depends on the data (items, tiles, K) is read from device
// memory and grids are sized by the workspace capacities, so one launch
// sequence serves any batch that fits and can be captured in a HIP graph.
#pragma once
#include "w2w_device.hpp"
#include "w2w_sort.hpp"

namespace w2w {

// Culled items' depth key: the largest key below 2.0f, so culled items sort
// behind every kept item of depth bits under 2.0f (ndc depth in front of the
// far plane) and share the top byte with them, which lets the depth sort skip
// its top pass.  Where culled items land among kept ones changes nothing:
// they have no instances.
constexpr uint32_t CULLED_DEPTH_KEY = 0x3fffffffu;

// status block words (first 256 bytes of the geometry buffer)
enum StatusWord : int {
  ST_K = 0,            // instances emitted (the reference's num_rendered)
  ST_OVERFLOW = 1,     // nonzero when K exceeded the binning capacity
  ST_SPLIT = 2,        // depth-sliced binning: depth positions in the near slice
  ST_KA = 3,           //   instances of the near slice
  ST_KB = 4,           //   instances of the far slice on open tiles
  ST_OPEN_UNITS = 5,   //   8x8 blocks left unfinished by the near slice
  ST_KH = 6,           //   far-slice instances on hinted tiles (binned with the near slice)
  ST_KNEAR = 7,        //   the near lists' instances (ST_KA + ST_KH)
  ST_BIG_H = 31,       //   far items with many hinted / open tiles, emitted a wave each
  ST_BIG_F = 32,
  ST_DTICKET = 8,      // 4 depth-sort tickets
  ST_TTICKET = 16,     // up to 4 tile-sort tickets
  ST_SCAN_TICKET = 24,
  ST_BLEND_TICKET = 26,
  ST_BWD_TICKET = 27,
  ST_SCANB_TICKET = 28,
  ST_BLENDB_TICKET = 30,
  ST_ORDER_N = 33,     // backward: tiles with work, in the heaviest-first order
  ST_COST_HIST = 34,   //   tiles per cost bucket (kCostBuckets words)
  ST_COST_CURSOR = 67, //   slots taken per cost bucket (kCostBuckets words)
  ST_REC_TICKET = 100, // the scan of k_preprocess's tile sums
  ST_KNEAR_RAW = 101,  // depth-sliced binning: the near lists' instances before the capacity cap
  ST_DTOP = 102,       // nonzero when a depth key's top byte differs from CULLED_DEPTH_KEY's
  ST_WORDS = 128,
};

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Several word ranges set in one launch (a step's status words, look-back
// words, histograms and masks are many small fills; each launch of its own
// costs more than its stores).  Ranges are filled one after the other by the
// whole grid; the range index is uniform, so every field is a kernel
// argument read.
constexpr int kFillRanges = 8;
struct FillRanges {
  uint32_t* ptr[kFillRanges];
  uint64_t len[kFillRanges];
  uint32_t val[kFillRanges];
  int n;
};

__global__ __launch_bounds__(256) void k_fill_ranges(FillRanges f) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
#pragma unroll
  for (int r = 0; r < kFillRanges; r++) {
    if (r >= f.n) break;
    uint32_t* const p = f.ptr[r];
    const uint32_t v = f.val[r];
    for (uint64_t i = t; i < f.len[r]; i += stride) p[i] = v;
  }
}

// Item -> (scene, Gaussian, view, frame).
struct ItemRef {
  uint32_t scene, g, view, frame, gg;
};

W2W_DEV ItemRef item_ref(const SceneDesc* scenes, uint32_t n_scenes, uint32_t i) {
  ItemRef r;
  r.scene = upper_index(n_scenes, i, [&](uint32_t k) { return scenes[k].item_off; });
  const SceneDesc& sd = scenes[r.scene];
  const uint32_t local = i - sd.item_off;
  r.g = local / sd.n_views;
  r.view = local - r.g * sd.n_views;
  r.frame = sd.frame_off + r.view;
  r.gg = sd.g_base + r.g;
  return r;
}

// members of the tile rectangle [x0, x1) x [y0, y1) (non-empty) from the
// prefix counts
W2W_DEV uint32_t sat_rect(const uint32_t* __restrict__ sat, uint32_t t0, uint32_t tx, uint32_t x0, uint32_t x1,
                          uint32_t y0, uint32_t y1) {
  const uint32_t* r1 = sat + t0 + (y1 - 1) * tx;
  uint32_t c = r1[x1 - 1];
  if (x0 > 0) c -= r1[x0 - 1];
  if (y0 > 0) {
    const uint32_t* r0 = sat + t0 + (y0 - 1) * tx;
    c -= r0[x1 - 1];
    if (x0 > 0) c += r0[x0 - 1];
  }
  return c;
}

// Look-back by a whole wave: lane l reads the word of tile p - l (a tile
// before the first counts as inclusive 0) and the wave adds the words up to
// the nearest inclusive one once all of them are published.  Sums saturate
// at LB_VAL (above every capacity), so a total too large for the 30-bit
// field still reads as an overflow.
// lb_publish: the tile's own total, as soon as it is known (tile 0's is
// already its inclusive prefix); lb_resolve: the look-back proper, then the
// tile's inclusive prefix.  scan_lookback_wave does both in a row.
W2W_DEV void lb_publish(uint32_t* lb, uint32_t tile, uint32_t tot, int lane) {
  if (lane == 0) st_release_dev(lb + tile, (tile == 0 ? LB_INC : LB_AGG) | tot);
}

W2W_DEV uint32_t lb_resolve(uint32_t* lb, uint32_t tile, uint32_t tot, int lane) {
  if (tile == 0) return 0u;
  uint32_t excl = 0;
  int p = (int)tile - 1;
  for (;;) {
    const int q = p - lane;
    const uint32_t w = q >= 0 ? ld_acquire_dev(lb + q) : LB_INC;
    const uint32_t flag = w & ~LB_VAL;
    const uint64_t inc = __ballot(flag == LB_INC);
    const uint64_t unready = __ballot(flag == 0u);
    const int first = inc ? __builtin_ctzll(inc) : 64;
    const uint64_t need = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
    if (unready & need) continue;  // a word up to the nearest inclusive one is not published yet
    uint64_t v = lane <= first ? (uint64_t)(w & LB_VAL) : 0ull;  // 64 words of up to 30 bits
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    excl = (uint32_t)min<uint64_t>((uint64_t)excl + v, LB_VAL);
    if (first < 64) break;
    p -= 64;
  }
  if (lane == 0) st_release_dev(lb + tile, LB_INC | min(excl + tot, LB_VAL));
  return excl;
}

W2W_DEV uint32_t scan_lookback_wave(uint32_t* lb, uint32_t tile, uint32_t tot, int lane) {
  lb_publish(lb, tile, tot, lane);
  return lb_resolve(lb, tile, tot, lane);
}

struct PreArgs {
  const void* table;
  uint32_t scenes_cap;
  uint32_t items_cap;
  const float* means;    // packed [G,3]
  const float* scales;   // [G,3] activated
  const float* rots;     // [G,4] normalised
  const float* opac;     // [G] activated
  const float* sh;       // packed SH blocks (SceneDesc::sh_off)
  int ts;
  SplatRec* splats;      // per item
  int32_t* radii;
  uint2* rect;           // (x0 | y0 << 16, x1 | y1 << 16)
  uint32_t* counts;
  uint32_t* dkey;        // depth bits, CULLED_DEPTH_KEY when culled
  uint32_t* ids;         // iota, the values of the depth sort
  uint32_t* item_frame;
  // depth-sliced binning: prefix counts of the hinted tiles (null: none)
  // and each item's hinted tiles out
  const uint32_t* hint_sat;
  uint32_t* hcount;
  // each tile of PRE_TILE items: its tile counts summed (scanned after this
  // kernel into the tiles' first backward-record positions)
  uint32_t* tile_tot;
  // depth-sliced binning: the near slice's depth positions are written to
  // status[ST_SPLIT] when split_permille > 0
  uint32_t* status;
  uint32_t split_permille;
};

constexpr uint32_t PRE_TILE = 256;

// One thread per item: the splat record, radius, rectangle, tile count,
// depth key.  A workgroup takes tiles of PRE_TILE consecutive items and scans
// their tile counts: the record holds the exclusive prefix inside the tile,
// and tile_tot the tile's sum (saturating at LB_VAL), so an item's first
// backward-record position (an item's records are contiguous, in item
// order) is the scanned tile_tot of its tile plus the record's word.  Every
// tile of the capacity is written (zero past the batch's items).
__global__ __launch_bounds__(PRE_TILE) void k_preprocess(PreArgs a) {
  __shared__ uint32_t s_tmp[PRE_TILE / 64];
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t n_items = min(hdr->n_items, a.items_cap);
  const uint32_t n_scenes = min(hdr->n_scenes, a.scenes_cap);
  const SceneDesc* scenes = table_scenes(a.table);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  const int tid = (int)threadIdx.x;
  const uint32_t ntiles = (a.items_cap + PRE_TILE - 1) / PRE_TILE;
  if (a.split_permille && blockIdx.x == 0 && tid == 0)
    a.status[ST_SPLIT] = (uint32_t)min<uint64_t>(((uint64_t)n_items * a.split_permille + 999u) / 1000u, n_items);
  for (uint32_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const uint32_t i = tile * PRE_TILE + (uint32_t)tid;
    const bool live = i < n_items;
    ItemRef ir{};
    Cam c;
    float p[3] = {0.0f, 0.0f, 0.0f};
    float o = 0.0f;
    Geom g{};
    bool keep = false;
    float skip_below = 0.0f;
    uint32_t ext = kExtentInf;
    uint32_t x0 = 0, y0 = 0, x1 = 0, y1 = 0;
    int32_t radius = 0;
    uint32_t cnt = 0;
    if (live) {
    ir = item_ref(scenes, n_scenes, i);
    const FrameDesc& fd = frames[ir.frame];
    load_cam(c, fd);
    const uint32_t gi = ir.gg;
    p[0] = a.means[3 * gi]; p[1] = a.means[3 * gi + 1]; p[2] = a.means[3 * gi + 2];
    const float s[3] = {a.scales[3 * gi], a.scales[3 * gi + 1], a.scales[3 * gi + 2]};
    const float q[4] = {a.rots[4 * gi], a.rots[4 * gi + 1], a.rots[4 * gi + 2], a.rots[4 * gi + 3]};
    o = a.opac[gi];
    geometry(c, p, s, q, g);
    keep = g.visible;
    float r = 0.0f;
    if (keep) {
      float half_tr = (g.cxx + g.cyy) * 0.5f;
      float ds = half_tr * half_tr - g.det;
      float root = __builtin_sqrtf(fmaxf(0.1f, ds));
      float lmax = fmaxf(half_tr + root, half_tr - root);
      float r3 = 3.0f * __builtin_sqrtf(lmax);
      keep = o > 0.0039f;
      if (keep) {
        float lnr = log_pos(o * 255.0f);
        // exponents below ln(1/(255 o)) - 1e-5 give o*exp(power) < 1/255 with a
        // margin far above the rounding of exp and of this log, so the blend
        // kernels may skip such pairs before evaluating the exponential
        skip_below = -lnr - 1e-5f;
        float radp = __builtin_sqrtf((2.0f * lmax) * lnr);
        r = fminf(radp, r3);
        keep = r > 0.0f;
      }
    }
    ext = keep ? cull_extent(g.ca, g.cb, g.cc, skip_below) : kExtentInf;
    if (keep) {
      const float tsf = (float)a.ts, tsm1 = (float)(a.ts - 1);
      x0 = tile_coord((g.px - r) / tsf, (uint32_t)fd.tiles_x);
      y0 = tile_coord((g.py - r) / tsf, (uint32_t)fd.tiles_y);
      x1 = tile_coord(((g.px + r) + tsm1) / tsf, (uint32_t)fd.tiles_x);
      y1 = tile_coord(((g.py + r) + tsm1) / tsf, (uint32_t)fd.tiles_y);
      if (fd.flags & FRAME_TIGHT) {
        // tiles holding a pixel centre within the extent box (half a pixel of
        // slack covers the rounding of these quotients)
        const float2 e = unpack_extent(__builtin_bit_cast(float, ext));
        if (e.x < 60000.0f && e.y < 60000.0f) {
          const uint32_t bx0 = tile_coord(((g.px - e.x) - 0.5f) / tsf, (uint32_t)fd.tiles_x);
          const uint32_t by0 = tile_coord(((g.py - e.y) - 0.5f) / tsf, (uint32_t)fd.tiles_y);
          const uint32_t bx1 = min(tile_coord(((g.px + e.x) + 0.5f) / tsf, (uint32_t)fd.tiles_x) + 1u,
                                   (uint32_t)fd.tiles_x);
          const uint32_t by1 = min(tile_coord(((g.py + e.y) + 0.5f) / tsf, (uint32_t)fd.tiles_y) + 1u,
                                   (uint32_t)fd.tiles_y);
          x0 = max(x0, bx0); y0 = max(y0, by0);
          x1 = min(x1, bx1); y1 = min(y1, by1);
        }
      }
      cnt = (x1 > x0 && y1 > y0) ? (x1 - x0) * (y1 - y0) : 0u;
      radius = (int32_t)__builtin_ceilf(r);
    }
    }
    SplatRec rec;
    uint32_t hcnt = 0, dkey = CULLED_DEPTH_KEY;
    if (live) {
    const SceneDesc& sd = scenes[ir.scene];
    const FrameDesc& fd = frames[ir.frame];
    // SH colour, direction normalize(mean - campos), +0.5, unclamped
    float rgb[3];
    {
      float dx = p[0] - c.cp[0], dy = p[1] - c.cp[1], dz = p[2] - c.cp[2];
      float len = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
      float x = dx / len, y = dy / len, z = dz / len;
      float Y[16];
      sh_basis(sd.D, x, y, z, Y);
      const int nb = (sd.D + 1) * (sd.D + 1);
      const float* c0 = a.sh + sd.sh_off + (size_t)ir.g * sd.M * 3;
      // the 3 nb coefficients used, with 16-byte loads when the row allows
      // (the loads stay inside those 3 nb floats; the array stays in registers)
      float cf[48];
      const int nf = 3 * nb;
      if ((reinterpret_cast<uintptr_t>(c0) & 15u) == 0) {
#pragma unroll
        for (int qd = 0; qd < 12; qd++) {
          if (4 * qd + 3 < nf) {
            const float4 v = reinterpret_cast<const float4*>(c0)[qd];
            cf[4 * qd] = v.x; cf[4 * qd + 1] = v.y; cf[4 * qd + 2] = v.z; cf[4 * qd + 3] = v.w;
          } else {
#pragma unroll
            for (int e = 0; e < 4; e++) cf[4 * qd + e] = 4 * qd + e < nf ? c0[4 * qd + e] : 0.0f;
          }
        }
      } else {
#pragma unroll
        for (int k = 0; k < 48; k++) cf[k] = k < nf ? c0[k] : 0.0f;
      }
      float r0 = Y[0] * cf[0], r1 = Y[0] * cf[1], r2 = Y[0] * cf[2];
#pragma unroll
      for (int k = 1; k < 16; k++) {
        if (k < nb) {
          r0 = __builtin_fmaf(Y[k], cf[3 * k], r0);
          r1 = __builtin_fmaf(Y[k], cf[3 * k + 1], r1);
          r2 = __builtin_fmaf(Y[k], cf[3 * k + 2], r2);
        }
      }
      rgb[0] = r0 + 0.5f;
      rgb[1] = r1 + 0.5f;
      rgb[2] = r2 + 0.5f;
    }

    Splat sp;
    if (keep) {
      sp.a = make_float4(g.px, g.py, g.ca, g.cb);
      sp.b = make_float4(g.cc, o, g.ndcz, skip_below);
    } else {
      sp.a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
      sp.b = make_float4(0.0f, o, 0.0f, 0.0f);
    }
    sp.c = make_float4(rgb[0], rgb[1], rgb[2], __builtin_bit_cast(float, ext));
    rec.s = sp;
    rec.q1 = keep ? cull_threshold(g.ca, g.cb, g.cc, skip_below) : __builtin_inff();
    rec.rect0 = x0 | (y0 << 16);
    rec.rect1 = x1 | (y1 << 16);
    if (a.hint_sat) hcnt = cnt ? sat_rect(a.hint_sat, fd.tile_off, (uint32_t)fd.tiles_x, x0, x1, y0, y1) : 0u;
    dkey = keep ? __float_as_uint(g.ndcz) : CULLED_DEPTH_KEY;
    }
    // the depth sort counts its top digit only when some key's top byte
    // differs from the culled key's (an atomic per such wave until the flag
    // reads as set; none in the usual case)
    if (__ballot((dkey >> 24) != (CULLED_DEPTH_KEY >> 24)) && (tid & 63) == 0 &&
        __hip_atomic_load(a.status + ST_DTOP, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
      atomicOr(a.status + ST_DTOP, 1u);
    uint32_t btot;
    const uint32_t bexcl = block_excl_scan_256(cnt, s_tmp, &btot);
    if (tid == 0) a.tile_tot[tile] = min(btot, LB_VAL);
    if (live) {
      rec.item_off = bexcl;
      a.splats[i] = rec;
      a.radii[i] = radius;
      a.rect[i] = make_uint2(rec.rect0, rec.rect1);
      a.counts[i] = cnt;
      if (a.hint_sat) a.hcount[i] = hcnt;
      a.dkey[i] = dkey;
      a.ids[i] = i;
      a.item_frame[i] = ir.frame;
    }
  }
}

// Exclusive scan of in[perm[i]] (or in[i] without a permutation) with a
// single-pass decoupled look-back; writes the offsets, the total and, when
// asked, an overflow flag (total > capacity).  Tiles are claimed through a
// ticket until none is left, so any grid size works; look-back words lb[]
// and the ticket must be zero on entry.  n comes from *d_n when given.
// Tiles are large (16384 values: few look-back links), loads and stores are
// striped over the workgroup (coalesced) and transposed through LDS so that
// each thread scans 16 consecutive values, and one wave looks back over 64
// predecessors per step.  Optional (GATHER only): with in_far, positions at
// or past *d_split take in_far[perm[i]] instead of in[perm[i]]; with
// *gate == 0 the scan is skipped (total 0, out untouched).
constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

template <bool GATHER>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan(
    const uint32_t* __restrict__ in, const uint32_t* __restrict__ perm,
    uint32_t* __restrict__ out, uint32_t* __restrict__ lb, uint32_t* __restrict__ ticket,
    uint32_t* __restrict__ total_out, uint32_t* __restrict__ overflow_out,
    const uint32_t* __restrict__ d_n, uint32_t n_cap, uint32_t capacity,
    const uint32_t* __restrict__ in_far = nullptr, const uint32_t* __restrict__ d_split = nullptr,
    const uint32_t* __restrict__ gate = nullptr) {
  __shared__ uint32_t s_v[SCAN_TILE + SCAN_TILE / 32];  // one pad word per 32: conflict-free transposes
  __shared__ uint32_t s_tmp[SCAN_THREADS / 64];
  __shared__ uint32_t s_tile, s_excl, s_tot;
  const int tid = threadIdx.x;
  if (gate && *gate == 0u) {
    if (blockIdx.x == 0 && tid == 0) {
      *total_out = 0u;
      if (overflow_out) *overflow_out = 0u;
    }
    return;
  }
  const uint32_t n = sort_count(d_n, n_cap);
  const uint32_t split = (GATHER && in_far) ? *d_split : 0xFFFFFFFFu;
  const uint32_t ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  auto pad = [](int i) { return i + (i >> 5); };
  for (;;) {
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= ntiles) return;
    const uint32_t tb = tile * SCAN_TILE;
    uint32_t pr[SCAN_ITEMS];  // GATHER: the permuted index of each striped position
    if constexpr (GATHER) {
#pragma unroll
      for (int j = 0; j < SCAN_ITEMS; j++) {
        const uint32_t idx = tb + j * SCAN_THREADS + tid;
        pr[j] = idx < n ? perm[idx] : 0xFFFFFFFFu;
      }
#pragma unroll
      for (int j = 0; j < SCAN_ITEMS; j++) {
        const uint32_t idx = tb + j * SCAN_THREADS + tid;
        s_v[pad(j * SCAN_THREADS + tid)] = pr[j] != 0xFFFFFFFFu ? (idx < split ? in : in_far)[pr[j]] : 0u;
      }
    } else {
#pragma unroll
      for (int j = 0; j < SCAN_ITEMS; j++) {
        const uint32_t idx = tb + j * SCAN_THREADS + tid;
        s_v[pad(j * SCAN_THREADS + tid)] = idx < n ? in[idx] : 0u;
      }
    }
    __syncthreads();
    uint32_t v[SCAN_ITEMS];
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
      const uint32_t c = s_v[pad(tid * SCAN_ITEMS + j)];
      v[j] = run;
      run += c;
    }
    const uint32_t texcl = block_excl_scan<SCAN_THREADS>(run, s_tmp);
    // each thread rewrites only the slots it read
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) s_v[pad(tid * SCAN_ITEMS + j)] = texcl + v[j];
    if (tid == SCAN_THREADS - 1) s_tot = texcl + run;
    __syncthreads();
    if (tid < 64) {
      const uint32_t tot = min(s_tot, LB_VAL);
      const uint32_t excl = scan_lookback_wave(lb, tile, tot, tid);
      if (tid == 0) {
        s_excl = excl;
        if (tile == ntiles - 1) {
          const uint32_t K = min(excl + tot, LB_VAL);
          *total_out = K;
          if (overflow_out) *overflow_out = K > capacity ? 1u : 0u;
        }
      }
    }
    __syncthreads();
    const uint32_t e0 = s_excl;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; j++) {
      const uint32_t idx = tb + j * SCAN_THREADS + tid;
      if (idx < n) out[idx] = e0 + s_v[pad(j * SCAN_THREADS + tid)];
    }
    __syncthreads();  // s_v and s_tile are rewritten by the next tile
  }
}

// Tile sets of depth-sliced binning, as bitmaps over the global tile ids
// (bit t of bm[t >> 5]): the tiles a set `a` holds and a set `b` (null: the
// empty set) does not.
W2W_DEV uint32_t tileset_word(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t w) {
  return b ? (a[w] & ~b[w]) : a[w];
}

// members among tiles [b0, b0 + n)
W2W_DEV uint32_t tileset_count(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t b0,
                               uint32_t n) {
  uint32_t c = 0;
  while (n > 0) {
    const uint32_t o = b0 & 31u, k = min(32u - o, n);
    uint32_t m = tileset_word(a, b, b0 >> 5) >> o;
    if (k < 32u) m &= (1u << k) - 1u;
    c += (uint32_t)__builtin_popcount(m);
    b0 += k;
    n -= k;
  }
  return c;
}

// Inclusive 2-D prefix counts of a tile set per frame, indexed like the
// tiles: sat[tile_off + y tiles_x + x] = members in columns 0..x of rows
// 0..y.  One workgroup per frame: row prefixes, then column prefixes, in LDS
// when the frame fits (rows padded by one word: the row pass reads a column
// per step, conflict-free), in global memory otherwise.
constexpr uint32_t kSatLds = 16384;
__global__ __launch_bounds__(256) void k_tile_sat(const void* __restrict__ table, uint32_t scenes_cap,
                                                  uint32_t frames_cap, const uint32_t* __restrict__ set,
                                                  uint32_t* __restrict__ sat) {
  __shared__ uint32_t s[kSatLds];
  const BatchHeader* hdr = static_cast<const BatchHeader*>(table);
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t nf = min(hdr->n_frames, frames_cap);
  for (uint32_t f = blockIdx.x; f < nf; f += gridDim.x) {
    const FrameDesc& fd = frames[f];
    const uint32_t tx = (uint32_t)fd.tiles_x, ty = (uint32_t)fd.tiles_y, t0 = fd.tile_off;
    const uint32_t rs = tx + 1u;  // padded row stride
    if ((uint64_t)rs * ty <= kSatLds) {
      for (uint32_t t = threadIdx.x; t < tx * ty; t += blockDim.x) {
        const uint32_t g = t0 + t;
        s[(t / tx) * rs + t % tx] = (set[g >> 5] >> (g & 31u)) & 1u;
      }
      __syncthreads();
      for (uint32_t y = threadIdx.x; y < ty; y += blockDim.x) {
        uint32_t run = 0;
        for (uint32_t x = 0; x < tx; x++) {
          run += s[y * rs + x];
          s[y * rs + x] = run;
        }
      }
      __syncthreads();
      for (uint32_t x = threadIdx.x; x < tx; x += blockDim.x) {
        uint32_t run = 0;
        for (uint32_t y = 0; y < ty; y++) {
          run += s[y * rs + x];
          s[y * rs + x] = run;
        }
      }
      __syncthreads();
      for (uint32_t t = threadIdx.x; t < tx * ty; t += blockDim.x) sat[t0 + t] = s[(t / tx) * rs + t % tx];
      __syncthreads();  // s is rewritten by the next frame
    } else {
      for (uint32_t y = threadIdx.x; y < ty; y += blockDim.x) {
        uint32_t run = 0;
        for (uint32_t x = 0; x < tx; x++) {
          const uint32_t t = t0 + y * tx + x;
          run += (set[t >> 5] >> (t & 31u)) & 1u;
          sat[t] = run;
        }
      }
      __syncthreads();
      for (uint32_t x = threadIdx.x; x < tx; x += blockDim.x) {
        uint32_t run = 0;
        for (uint32_t y = 0; y < ty; y++) {
          const uint32_t t = t0 + y * tx + x;
          run += sat[t];
          sat[t] = run;
        }
      }
      __syncthreads();
    }
  }
}

// Depth-sliced binning: the near slice is the first S = ceil(P * permille /
// 1000) depth positions (set by k_preprocess).  One scan in depth order took
// each near item's tile count and each far item's hinted-tile count, so its
// offsets place both in the near lists; its total is the near lists' size.
// Split into the near slice's instances (the offset at S) and the far
// items' hinted instances.
__global__ void k_near_total(const void* __restrict__ table, uint32_t items_cap,
                             const uint32_t* __restrict__ offs, uint32_t* __restrict__ status,
                             uint32_t capacity) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const uint32_t P = min(static_cast<const BatchHeader*>(table)->n_items, items_cap);
  const uint32_t S = status[ST_SPLIT];
  const uint32_t raw = status[ST_KNEAR_RAW];
  const uint32_t KA = S < P ? offs[S] : raw;
  status[ST_KA] = min(KA, capacity);
  status[ST_KH] = raw - KA;
  status[ST_KNEAR] = min(raw, capacity);
}

// Per depth position of the far slice (0 in the near slice): how many tiles
// of its item's rectangle are in the tile set (a minus b).  With *gate == 0
// the set is known to be empty and nothing is written (the scan and the
// emission behind it are gated the same way).
__global__ __launch_bounds__(256) void k_count_tiles(const void* __restrict__ table, uint32_t scenes_cap,
                                                     uint32_t items_cap, const uint32_t* __restrict__ status,
                                                     const uint32_t* __restrict__ perm,
                                                     const uint2* __restrict__ rect,
                                                     const uint32_t* __restrict__ item_frame,
                                                     const uint32_t* __restrict__ set_a,
                                                     const uint32_t* __restrict__ set_b,
                                                     uint32_t* __restrict__ count,
                                                     const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t P = min(static_cast<const BatchHeader*>(table)->n_items, items_cap);
  const uint32_t S = status[ST_SPLIT];
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    uint32_t c = 0;
    if (i >= S) {
      const uint32_t g = perm[i];
      const uint2 r = rect[g];
      const uint32_t x0 = r.x & 0xffffu, y0 = r.x >> 16, x1 = r.y & 0xffffu, y1 = r.y >> 16;
      if (x1 > x0 && y1 > y0) {
        const FrameDesc& fd = frames[item_frame[g]];
        for (uint32_t ty = y0; ty < y1; ty++)
          c += tileset_count(set_a, set_b, fd.tile_off + ty * (uint32_t)fd.tiles_x + x0, x1 - x0);
      }
    }
    count[i] = c;
  }
}

// Tile keys are the global tile ids: 16-bit when the batch has at most
// 65536 tiles (the emission, the sort and the ranges then move two bytes per
// key), 32-bit otherwise.
template <typename KT>
W2W_DEV void store_keys4(KT* p, uint32_t k0, uint32_t k1, uint32_t k2, uint32_t k3) {
  if constexpr (sizeof(KT) == 2)
    *reinterpret_cast<uint2*>(p) = make_uint2(k0 | (k1 << 16), k2 | (k3 << 16));
  else
    *reinterpret_cast<uint4*>(p) = make_uint4(k0, k1, k2, k3);
}

template <typename KT>
W2W_DEV void load_keys4(const KT* p, uint32_t& k0, uint32_t& k1, uint32_t& k2, uint32_t& k3) {
  if constexpr (sizeof(KT) == 2) {
    const uint2 v = *reinterpret_cast<const uint2*>(p);
    k0 = v.x & 0xFFFFu; k1 = v.x >> 16; k2 = v.y & 0xFFFFu; k3 = v.y >> 16;
  } else {
    const uint4 v = *reinterpret_cast<const uint4*>(p);
    k0 = v.x; k1 = v.y; k2 = v.z; k3 = v.w;
  }
}

constexpr uint32_t kFarSerial = 32;
constexpr uint32_t kSetLds = 4096;  // tile-set words k_emit_far keeps in LDS (131072 tiles)

// Far items' instances on the tiles of a set (a minus b): one thread per far
// depth position (most have none and stop at once), the item's member tiles
// in row-major order from *d_base (null: 0) + its compact offset coff[i]
// (an exclusive scan whose total is *d_total: counts are differences).  An
// instance's key is its tile's bit index.  Adds the keys' digits to the tile
// sort's histogram (hist[pass][256]) like k_emit.  Skipped when *gate == 0.
// An item with more than kFarSerial instances is left to k_emit_far_big
// (its depth position appended to big[], *n_big counts them): one thread
// would hold up its wave.
template <typename KT>
__global__ __launch_bounds__(256) void k_emit_far(const void* __restrict__ table, uint32_t scenes_cap,
                                                  uint32_t tiles_cap, uint32_t items_cap, uint32_t capacity,
                                                  const uint32_t* __restrict__ status,
                                                  const uint32_t* __restrict__ perm,
                                                  const uint32_t* __restrict__ coff,
                                                  const uint32_t* __restrict__ d_total,
                                                  const uint2* __restrict__ rect,
                                                  const uint32_t* __restrict__ item_frame,
                                                  const uint32_t* __restrict__ set_a,
                                                  const uint32_t* __restrict__ set_b,
                                                  const uint32_t* __restrict__ d_base,
                                                  KT* __restrict__ tile_keys,
                                                  uint32_t* __restrict__ ids,
                                                  uint32_t* __restrict__ hist, int hist_passes,
                                                  int hist_last_bits, const uint32_t* __restrict__ gate,
                                                  uint32_t* __restrict__ big, uint32_t* __restrict__ n_big) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t s_hist[4][256];
  __shared__ uint32_t s_set[kSetLds];
  const BatchHeader* hdr = static_cast<const BatchHeader*>(table);
  // the tile set in LDS when it fits (the walk below reads a word per row
  // of each item's rectangle, one dependent load after another)
  const uint32_t set_words = (min(hdr->n_tiles, tiles_cap) + 31u) / 32u;
  const bool in_lds = set_words <= kSetLds;
  for (int k = threadIdx.x; k < 4 * 256; k += blockDim.x) (&s_hist[0][0])[k] = 0u;
  if (in_lds)
    for (uint32_t k = threadIdx.x; k < set_words; k += blockDim.x) s_set[k] = tileset_word(set_a, set_b, k);
  __syncthreads();
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t P = min(hdr->n_items, items_cap);
  const uint32_t S = status[ST_SPLIT];
  const uint32_t base = d_base ? *d_base : 0u;
  const uint32_t total = *d_total;
  for (uint32_t i = S + blockIdx.x * blockDim.x + threadIdx.x; i < P; i += gridDim.x * blockDim.x) {
    const uint32_t g = perm[i];  // loaded with the offsets: one dependent load less on the item's path
    const uint32_t c = (i + 1 < P ? coff[i + 1] : total) - coff[i];
    if (c == 0) continue;
    if (c > kFarSerial) {
      big[atomicAdd(n_big, 1u)] = i;
      continue;
    }
    const uint2 r = rect[g];
    const uint32_t x0 = r.x & 0xffffu, y0 = r.x >> 16, x1 = r.y & 0xffffu, y1 = r.y >> 16;
    const FrameDesc& fd = frames[item_frame[g]];
    uint32_t o = min(base + coff[i], capacity);
    const uint32_t oe = min(o + c, capacity);
    for (uint32_t ty = y0; ty < y1 && o < oe; ty++) {
      uint32_t pos = fd.tile_off + ty * (uint32_t)fd.tiles_x + x0, rem = x1 - x0;
      while (rem > 0 && o < oe) {
        const uint32_t sh = pos & 31u, k = min(32u - sh, rem);
        uint32_t m = (in_lds ? s_set[pos >> 5] : tileset_word(set_a, set_b, pos >> 5)) >> sh;
        if (k < 32u) m &= (1u << k) - 1u;
        while (m && o < oe) {
          const uint32_t key = pos + (uint32_t)__builtin_ctz(m);
          tile_keys[o] = (KT)key;
          ids[o] = g;
          o++;
          m &= m - 1u;
          for (int p = 0; p < hist_passes; p++) {
            const int bits = (p == hist_passes - 1) ? hist_last_bits : 8;
            atomicAdd(&s_hist[p][(key >> (8 * p)) & ((1u << bits) - 1u)], 1u);
          }
        }
        pos += k;
        rem -= k;
      }
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < hist_passes * 256; k += blockDim.x) {
    const uint32_t v = (&s_hist[0][0])[k];
    if (v) atomicAdd(hist + (blockIdx.x % kHistCopies) * kHistCopyStride + k, v);
  }
}

// The far items k_emit_far left (more than kFarSerial instances): a wave per
// item walks its rectangle 64 tiles at a time in row-major order; member
// tiles are placed by ballot prefix counts.  Same outputs as the serial walk.
template <typename KT>
__global__ __launch_bounds__(256) void k_emit_far_big(const void* __restrict__ table, uint32_t scenes_cap,
                                                      uint32_t capacity, const uint32_t* __restrict__ perm,
                                                      const uint32_t* __restrict__ coff,
                                                      const uint2* __restrict__ rect,
                                                      const uint32_t* __restrict__ item_frame,
                                                      const uint32_t* __restrict__ set_a,
                                                      const uint32_t* __restrict__ set_b,
                                                      const uint32_t* __restrict__ d_base,
                                                      const uint32_t* __restrict__ big,
                                                      const uint32_t* __restrict__ n_big, uint32_t big_cap,
                                                      KT* __restrict__ tile_keys,
                                                      uint32_t* __restrict__ ids,
                                                      uint32_t* __restrict__ hist, int hist_passes,
                                                      int hist_last_bits, const uint32_t* __restrict__ gate) {
  if (gate && *gate == 0u) return;
  __shared__ uint32_t s_hist[4][256];
  for (int k = threadIdx.x; k < 4 * 256; k += blockDim.x) (&s_hist[0][0])[k] = 0u;
  __syncthreads();
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t nb = min(*n_big, big_cap);
  const uint32_t base = d_base ? *d_base : 0u;
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); w < nb; w += waves) {
    const uint32_t i = big[w];
    const uint32_t g = perm[i];
    const uint2 r = rect[g];
    const uint32_t x0 = r.x & 0xffffu, y0 = r.x >> 16, x1 = r.y & 0xffffu, y1 = r.y >> 16;
    const uint32_t rw = x1 - x0, n = rw * (y1 - y0);
    const FrameDesc& fd = frames[item_frame[g]];
    uint32_t o = base + coff[i];
    for (uint32_t t0 = 0; t0 < n; t0 += 64) {
      const uint32_t t = t0 + lane;
      bool member = false;
      uint32_t key = 0;
      if (t < n) {
        key = fd.tile_off + (y0 + t / rw) * (uint32_t)fd.tiles_x + x0 + t % rw;
        member = (tileset_word(set_a, set_b, key >> 5) >> (key & 31u)) & 1u;
      }
      const uint64_t m = __ballot(member);
      const uint32_t at = o + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
      if (member && at < capacity) {
        tile_keys[at] = (KT)key;
        ids[at] = g;
        for (int p = 0; p < hist_passes; p++) {
          const int bits = (p == hist_passes - 1) ? hist_last_bits : 8;
          atomicAdd(&s_hist[p][(key >> (8 * p)) & ((1u << bits) - 1u)], 1u);
        }
      }
      o += (uint32_t)__popcll(m);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < hist_passes * 256; k += blockDim.x) {
    const uint32_t v = (&s_hist[0][0])[k];
    if (v) atomicAdd(hist + (blockIdx.x % kHistCopies) * kHistCopyStride + k, v);
  }
}

// Instance emission in depth order.  A workgroup owns depth-sorted positions
// [256b, 256b+256) and the contiguous instance range they cover; each thread
// strides over instances and finds its owner by binary search in LDS.
// Workgroups take chunks in a fixed order (any grid size): a ticket would
// cost a contended atomic per chunk, which at a few thousand chunks per
// launch costs more than the imbalance it would even out.
// Emission also histograms the tile keys for the tile sort (every digit of
// every pass, hist[pass][256], pre-zeroed), so the sort does not read the
// keys an extra time.
// Depth positions [0, *d_end) (null: all; the near slice of depth-sliced
// binning otherwise).
template <typename KT>
__global__ __launch_bounds__(256) void k_emit(const void* __restrict__ table, uint32_t scenes_cap,
                                              const uint32_t* __restrict__ perm,
                                              const uint32_t* __restrict__ offs,
                                              const uint32_t* __restrict__ counts,
                                              const uint2* __restrict__ rect,
                                              const uint32_t* __restrict__ item_frame,
                                              uint32_t items_cap, uint32_t capacity,
                                              const uint32_t* __restrict__ d_end,
                                              KT* __restrict__ tile_keys,
                                              uint32_t* __restrict__ ids,
                                              uint32_t* __restrict__ hist, int hist_passes,
                                              int hist_last_bits) {
  __shared__ uint32_t s_start[256], s_g[256], s_x0[256], s_y0[256], s_w[256], s_tx[256], s_toff[256];
  __shared__ uint32_t s_end;
  __shared__ uint32_t s_hist[4][256];
  const BatchHeader* hdr = static_cast<const BatchHeader*>(table);
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t P = min(hdr->n_items, items_cap);
  const uint32_t end = d_end ? min(*d_end, P) : P;
  const uint32_t nchunks = (end + 255) / 256;
  const int tid = threadIdx.x;
  for (int i = tid; i < 4 * 256; i += 256) (&s_hist[0][0])[i] = 0u;
  for (uint32_t chunk = blockIdx.x; chunk < nchunks; chunk += gridDim.x) {
    const uint32_t cbase = chunk * 256;
    const uint32_t i = cbase + tid;
    const uint32_t last = min(end, cbase + 256) - 1;
    if (i <= last) {
      uint32_t g = perm[i];
      uint2 r = rect[g];
      uint32_t c = counts[g];
      const FrameDesc& fd = frames[item_frame[g]];
      s_start[tid] = offs[i];
      s_g[tid] = g;
      s_x0[tid] = r.x & 0xffffu;
      s_y0[tid] = r.x >> 16;
      s_w[tid] = (r.y & 0xffffu) - (r.x & 0xffffu);
      s_tx[tid] = (uint32_t)fd.tiles_x;
      s_toff[tid] = fd.tile_off;
      if (i == last) s_end = offs[i] + c;
    }
    __syncthreads();
    const int nb = (int)(last - cbase + 1);
    const uint32_t kb = s_start[0], ke = min(s_end, capacity);
    // each thread emits an aligned group of four consecutive instances (one
    // 16-byte store of keys and one of ids when the group is whole): one
    // binary search per group, then the owner and the tile advance in place
    for (uint32_t q = (kb & ~3u) + 4 * tid; q < ke; q += 4 * 256) {
      const uint32_t kf = max(q, kb);
      int lo = 0, hi = nb - 1;  // largest j with s_start[j] <= kf
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (s_start[mid] <= kf) lo = mid; else hi = mid - 1;
      }
      uint32_t w = s_w[lo], x0 = s_x0[lo];
      uint32_t local = kf - s_start[lo];
      uint32_t ty = s_y0[lo] + local / w, tx = x0 + local % w;
      uint32_t key4[4], id4[4];
#pragma unroll
      for (int m = 0; m < 4; m++) {
        const uint32_t k = q + m;
        key4[m] = 0u;
        id4[m] = 0u;
        if (k < kb || k >= ke) continue;
        if (k > kf) {  // next instance: same item's next tile, or a later item's first
          if (lo + 1 < nb && s_start[lo + 1] <= k) {
            do lo++; while (lo + 1 < nb && s_start[lo + 1] <= k);
            w = s_w[lo];
            x0 = s_x0[lo];
            ty = s_y0[lo];
            tx = x0;
          } else if (++tx == x0 + w) {
            tx = x0;
            ty++;
          }
        }
        key4[m] = s_toff[lo] + ty * s_tx[lo] + tx;
        id4[m] = s_g[lo];
      }
      // the group's digits, runs of equal digits counted with one atomic
      // (the upper digits of a row's consecutive tiles mostly agree)
      for (int p = 0; p < hist_passes; p++) {
        const int bits = (p == hist_passes - 1) ? hist_last_bits : 8;
        const uint32_t dmask = (1u << bits) - 1u;
        uint32_t rd = 0u, rc = 0u;
#pragma unroll
        for (int m = 0; m < 4; m++) {
          if (q + m < kb || q + m >= ke) continue;
          const uint32_t d = (key4[m] >> (8 * p)) & dmask;
          if (rc && d == rd) {
            rc++;
          } else {
            if (rc) atomicAdd(&s_hist[p][rd], rc);
            rd = d;
            rc = 1u;
          }
        }
        if (rc) atomicAdd(&s_hist[p][rd], rc);
      }
      if (q >= kb && q + 3 < ke) {
        store_keys4(tile_keys + q, key4[0], key4[1], key4[2], key4[3]);
        *reinterpret_cast<uint4*>(ids + q) = make_uint4(id4[0], id4[1], id4[2], id4[3]);
      } else {
#pragma unroll
        for (int m = 0; m < 4; m++) {
          if (q + m >= kb && q + m < ke) {
            tile_keys[q + m] = (KT)key4[m];
            ids[q + m] = id4[m];
          }
        }
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < hist_passes * 256; i += 256) {
    const uint32_t v = (&s_hist[0][0])[i];
    if (v) atomicAdd(hist + (blockIdx.x % kHistCopies) * kHistCopyStride + i, v);
  }
}

// Per-tile [start, end) of the sorted tile keys; four keys per thread with
// one 8- or 16-byte load (the key array is 256-byte aligned).
template <typename KT>
__global__ __launch_bounds__(256) void k_ranges(const KT* __restrict__ keys,
                                                const uint32_t* __restrict__ d_n,
                                                const void* __restrict__ table, uint32_t capacity,
                                                uint32_t tiles_cap, uint2* __restrict__ ranges) {
  uint32_t n = *d_n;
  n = n < capacity ? n : capacity;
  const uint32_t T = min(static_cast<const BatchHeader*>(table)->n_tiles, tiles_cap);
  const uint32_t n4 = (n + 3) / 4;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
    const uint32_t i0 = 4 * q;
    uint32_t k[6];  // keys i0-1 .. i0+4; out of range = all ones (no tile)
    k[0] = i0 > 0 ? (uint32_t)keys[i0 - 1] : 0xFFFFFFFFu;
    if (i0 + 4 <= n) {
      load_keys4(keys + i0, k[1], k[2], k[3], k[4]);
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) k[1 + j] = i0 + j < n ? (uint32_t)keys[i0 + j] : 0xFFFFFFFFu;
    }
    k[5] = i0 + 4 < n ? (uint32_t)keys[i0 + 4] : 0xFFFFFFFFu;
#pragma unroll
    for (int j = 1; j <= 4; j++) {
      const uint32_t t = k[j];
      if (t >= T) continue;
      if (k[j - 1] != t) ranges[t].x = i0 + j - 1;
      if (k[j + 1] != t) ranges[t].y = i0 + j;
    }
  }
}

// frame owning global tile t
W2W_DEV uint32_t tile_frame(const FrameDesc* frames, uint32_t n_frames, uint32_t t) {
  return upper_index(n_frames, t, [&](uint32_t k) { return frames[k].tile_off; });
}

// Depth-sliced binning: the next forward's hint, the tiles its near lists
// left open grown by one tile within their frame (a tile that stays open is
// usually next to one that was): bit t = any open tile among t and its eight
// neighbours.  One thread per tile; a wave's ballot gives two 32-tile words.
__global__ __launch_bounds__(256) void k_dilate_tiles(const void* __restrict__ table, uint32_t scenes_cap,
                                                      uint32_t tiles_cap, const uint32_t* __restrict__ open_bm,
                                                      uint32_t* __restrict__ hint_bm) {
  const BatchHeader* hdr = static_cast<const BatchHeader*>(table);
  const FrameDesc* frames = table_frames(table, scenes_cap);
  const uint32_t T = min(hdr->n_tiles, tiles_cap);
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t Tw = (T + 63u) & ~63u;  // whole waves: every lane reaches the ballot
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < Tw; t += gridDim.x * blockDim.x) {
    uint32_t any = 0;
    if (t < T) {
      const FrameDesc& fd = frames[tile_frame(frames, hdr->n_frames, t)];
      const int txn = fd.tiles_x, tyn = fd.tiles_y;
      const int lt = (int)(t - fd.tile_off), tx = lt % txn, ty = lt / txn;
      for (int ny = max(ty - 1, 0); ny <= min(ty + 1, tyn - 1); ny++)
        for (int nx = max(tx - 1, 0); nx <= min(tx + 1, txn - 1); nx++) {
          const uint32_t nt = fd.tile_off + (uint32_t)(ny * txn + nx);
          any |= (open_bm[nt >> 5] >> (nt & 31u)) & 1u;
        }
    }
    const uint64_t b = __ballot(any != 0u);
    const uint32_t w = t >> 5;  // lanes 0 and 32 hold the first tile of a word
    if ((lane & 31u) == 0u && w < (T + 31u) / 32u) hint_bm[w] = (uint32_t)(b >> lane);
  }
}

// An entry of a tile order (k_tile_order: heaviest tiles first by a per-tile
// cost): the tile in bits 0-29, the issue priority of the wave that takes it
// in bits 30-31.
constexpr uint32_t kOrderTile = (1u << 30) - 1u;

struct BlendArgs {
  const void* table;
  uint32_t scenes_cap, tiles_cap, items_cap;
  uint32_t* ticket;           // zeroed; tiles are claimed in order through it
  const uint2* ranges;
  const uint32_t* ids;
  const SplatRec* splats;     // per item: splat and block-test constants
  unsigned long long* stats;  // optional [3]: evaluated pairs, contributing pairs, loaded instances
  float* color;               // per frame [3,H,W] at 3*pix_off
  float* alpha;               // per frame [H,W] at pix_off
  float* depth;
  uint32_t* n_contrib;
  float* final_T;
  // depth-sliced binning (modes BLEND_NEAR / BLEND_FAR): the far slice's
  // ranges and sorted ids, and the open-tile bitmap
  const uint2* ranges_far;
  const uint32_t* ids_far;
  uint32_t* open_bm;        // tiles with unfinished pixels after their near list
  const uint32_t* hint_bm;  // tiles whose near list is their whole list
  uint32_t* open_units;     // list of the unfinished 8x8 blocks (units) of other tiles, any order
  uint32_t* n_open_units;
  // optional, zeroed: per tile, the number of contributing (pixel, splat)
  // pairs, for frames that take part in the backward (which visits tiles
  // heaviest first).  The same in every emission and binning mode: each
  // pixel's sequence of contributors is.
  uint32_t* tile_cost;
  // optional (modes FULL and NEAR): every tile, heaviest first by the last
  // forward's costs; a quad of blocks is then four blocks in this order
  const uint32_t* order;
};

// BLEND_FULL: one list per tile, final outputs.  BLEND_NEAR: the near lists;
// a block with unfinished pixels marks its tile open, and unless the tile is
// hinted (its near list is its whole list, so its outputs are final) it
// stores every pixel's state unfinished (colour without background,
// n_contrib with bit 31 set once the pixel is done) and appends itself to
// the open-block list; other blocks write final outputs.  BLEND_FAR: the
// waves take the open blocks of that list in turn (each on its own: their
// work is uneven, so no workgroup barrier, and the list is short, so no
// ticket), resume the state with the far list (instance indices continue
// after the near list's) and write the final outputs.
enum BlendMode : int { BLEND_FULL = 0, BLEND_NEAR = 1, BLEND_FAR = 2 };
constexpr uint32_t kDoneBit = 1u << 31;

// Blend work unit: an 8x8 pixel block of a tile, one wave (lane = pixel).
constexpr int kUnit = 8;

template <int TS>
constexpr uint32_t units_per_tile() {
  return (uint32_t)(TS / kUnit) * (uint32_t)(TS / kUnit);
}

// Pixel block of unit `unit`; returns the frame.
template <int TS>
W2W_DEV const FrameDesc& unit_block(const FrameDesc* frames, uint32_t n_frames, uint32_t unit,
                                    uint32_t& tile, int& ux, int& uy) {
  constexpr uint32_t UPT = units_per_tile<TS>();
  constexpr int UX = TS / kUnit;
  tile = unit / UPT;
  const int q = (int)(unit % UPT);
  const FrameDesc& fd = frames[tile_frame(frames, n_frames, tile)];
  const int lt = (int)(tile - fd.tile_off);
  ux = (lt % fd.tiles_x) * TS + (q % UX) * kUnit;
  uy = (lt / fd.tiles_x) * TS + (q / UX) * kUnit;
  return fd;
}

W2W_DEV float bcast(float v, int j) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), j));
}

W2W_DEV bool wave_all(bool p) { return __ballot(!p) == 0ull; }

// Forward blend (gaussianForwardPass.cu:300-385 semantics): every pixel walks
// its tile's depth-ordered list front to back and stops before the splat that
// would take its transmittance below 1e-4.
//
// Each wave owns one 8x8 pixel block and walks the list in chunks of 64 * CW
// instances (CW = 1, 2, 4: the counterpart of the reference's 64/128/256
// batch, pixelGaussiansFull.cu:41-76): lane l gathers instances base + l,
// base + 64 + l, ..., and per 64-instance sub-chunk tests each one's cull
// extent against the block; a ballot leaves the splats that may touch any
// pixel of the block, and the wave visits those in order, broadcasting each
// from its lane.  A dropped splat is one every pixel of the block would skip
// (its exponent is below skip_below), so each pixel sees the same sequence of
// contributing splats, in the same order and with the same arithmetic, as a
// walk over the whole list, whatever CW is.  The next chunk's splat records
// and the chunk after's ids are in flight while a chunk is processed: CW
// sets how many gathers each lane keeps in flight.  A workgroup claims four
// consecutive blocks (a quad) at a time and its waves take them one by one,
// each moving on to the next quad on its own (no workgroup barrier).
// Occupancy: CW 1 runs 5 waves per SIMD (at most 102 VGPRs; 4 waves: 0.457
// ms per launch at the benchmark size, 5 waves: 0.427 ms), CW 2 up to 4 (it
// would spill at 5), CW 4 up to 2 (it holds four records per lane).
template <int TS, int MODE, int CW>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CW == 1 ? 5 : 1, CW == 1 ? 5 : (CW == 2 ? 4 : 2)))) void k_blend(
    BlendArgs a) {
  __shared__ Splat s_chunk[4][64];  // each wave's current sub-chunk, for broadcast reads
  constexpr int CH = 64 * CW;
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t T = min(hdr->n_tiles, a.tiles_cap);
  const uint32_t P = min(hdr->n_items, a.items_cap);
  const uint32_t U = T * units_per_tile<TS>();
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  const int lane = (int)(threadIdx.x & 63u);
  Splat* const stage = s_chunk[threadIdx.x >> 6];
  const float a_min = 0x1.010102p-8f;
  uint32_t n_eval = 0, n_con = 0, n_loaded = 0;
  // the workgroup's current quad of four consecutive blocks (one tile of
  // edge 16) and how many times its waves asked for a block since it was
  // claimed: (quad << 4) | count.  Counts 0-3 are its blocks; the wave that
  // reads 4 claims the next quad, waves reading 5 or more wait for it.  Each
  // of the four waves asks at most once past 4 before the claim resets the
  // count to 1, so the count stays below 16.  Starts "used up", so the first
  // wave claims a quad.
  __shared__ uint32_t s_claim;
  if (MODE != BLEND_FAR && threadIdx.x == 0) s_claim = 4u;
  if (MODE != BLEND_FAR) __syncthreads();
  const uint32_t nquads = (U + 3) / 4;
  const uint32_t n_open = MODE == BLEND_FAR ? min(*a.n_open_units, U) : 0u;
  uint32_t far_next = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);  // this wave's next list entry
  for (;;) {
    uint32_t unit;
    if constexpr (MODE == BLEND_FAR) {
      const uint32_t w = far_next;
      far_next += gridDim.x * (blockDim.x >> 6);
      if (w >= n_open) break;
      unit = a.open_units[w];
    } else {
      // each wave takes the next block of its workgroup's quad (so the
      // blocks of a tile gather its list through one CU's caches, without
      // waiting for each other); the wave that finds the quad used up
      // claims the next one for the workgroup, and waves that find it being
      // replaced wait for the new one
      uint32_t got = 0;
      if ((threadIdx.x & 63u) == 0) {
        for (;;) {
          const uint32_t v = atomicAdd(&s_claim, 1u);
          const uint32_t k = v & 15u;
          if (k < 4u) {
            got = (v >> 4) < nquads ? (v >> 4) * 4u + k : 0xFFFFFFFFu;
            break;
          }
          if (k == 4u) {
            const uint32_t qn = atomicAdd(a.ticket, 1u);
            atomicExch(&s_claim, (min(qn, 0x0FFFFFFFu) << 4) | 1u);  // block 0 of it is this wave's
            got = qn < nquads ? qn * 4u : 0xFFFFFFFFu;
            break;
          }
          while ((__atomic_load_n(&s_claim, __ATOMIC_RELAXED) & 15u) >= 4u) {
          }
        }
      }
      unit = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)got, 0, 64));
      if (unit == 0xFFFFFFFFu) break;
      if (unit >= U) continue;
      if (a.order) {
        // position unit / UPT of the order, block unit % UPT of its tile; the
        // heaviest tiles' waves take issue priority
        constexpr uint32_t UPT = units_per_tile<TS>();
        const uint32_t e = a.order[unit / UPT];
        unit = (e & kOrderTile) * UPT + unit % UPT;
        switch (e >> 30) {
          case 3: __builtin_amdgcn_s_setprio(3); break;
          case 2: __builtin_amdgcn_s_setprio(2); break;
          case 1: __builtin_amdgcn_s_setprio(1); break;
          default: __builtin_amdgcn_s_setprio(0); break;
        }
        if (unit >= U) continue;
      }
    }
    uint32_t tile;
    int ux, uy;
    const FrameDesc& fd = unit_block<TS>(frames, hdr->n_frames, unit, tile, ux, uy);
    const int px = ux + (lane & 7), py = uy + (lane >> 3);
    const bool inside = px < fd.W && py < fd.H;
    const float pxf = (float)px, pyf = (float)py;
    const float x0 = (float)ux, x1 = (float)(ux + kUnit - 1);
    const float y0 = (float)uy, y1 = (float)(uy + kUnit - 1);
    const uint2 range_near = a.ranges[tile];
    // the tile's hint bit, read now so the load is done by the block's end
    const uint32_t hint_word = MODE == BLEND_NEAR ? a.hint_bm[tile >> 5] : 0u;
    const uint2 range = MODE == BLEND_FAR ? a.ranges_far[tile] : range_near;
    const uint32_t* const ids = MODE == BLEND_FAR ? a.ids_far : a.ids;
    const uint32_t lbase = MODE == BLEND_FAR ? range_near.y - range_near.x : 0u;
    const int n = (int)(range.y - range.x);
    bool done = !inside;
    float T_ = 1.0f, C0 = 0.0f, C1 = 0.0f, C2 = 0.0f, A = 0.0f, Dd = 0.0f;
    uint32_t last = 0;
    uint32_t pairs = 0;  // contributing (pixel, splat) pairs of this pass (wave-uniform)
    const size_t HW = (size_t)fd.W * fd.H;
    const size_t lp = (size_t)py * fd.W + px;
    const size_t pid = fd.pix_off + lp;
    float* col = a.color + 3 * fd.pix_off;
    if (MODE == BLEND_FAR && inside) {
      C0 = col[lp];
      C1 = col[HW + lp];
      C2 = col[2 * HW + lp];
      A = a.alpha[pid];
      Dd = a.depth[pid];
      T_ = a.final_T[pid];
      const uint32_t v = a.n_contrib[pid];
      last = v & ~kDoneBit;
      done = (v & kDoneBit) != 0u;
    }

    // software pipeline: ids two chunks ahead, splat records one chunk ahead
    // (lanes past the end gather splat 0, which exists whenever n > 0)
    uint32_t id_n[CW];
    Splat cur[CW];
    float qcur[CW];
#pragma unroll
    for (int c = 0; c < CW; c++) {
      id_n[c] = 0u;
      cur[c] = Splat{};
      qcur[c] = 0.0f;
    }
    if (n > 0 && !wave_all(done)) {
#pragma unroll
      for (int c = 0; c < CW; c++) {
        const int k = 64 * c + lane;
        const uint32_t id = k < n ? ids[range.x + k] : 0u;
        const SplatRec& r = a.splats[id < P ? id : 0u];
        cur[c] = r.s;
        qcur[c] = r.q1;
        const int k1 = CH + k;
        id_n[c] = k1 < n ? ids[range.x + k1] : 0u;
      }
    }
    for (int base = 0; base < n; base += CH) {
      if (wave_all(done)) break;
      Splat sc[CW];
      float kq[CW];
#pragma unroll
      for (int c = 0; c < CW; c++) {
        sc[c] = cur[c];
        kq[c] = qcur[c];
      }
      if (base + CH < n) {
#pragma unroll
        for (int c = 0; c < CW; c++) {
          const SplatRec& r = a.splats[id_n[c] < P ? id_n[c] : 0u];
          cur[c] = r.s;
          qcur[c] = r.q1;
          const int k2 = base + 2 * CH + 64 * c + lane;
          id_n[c] = k2 < n ? ids[range.x + k2] : 0u;
        }
      }
#pragma unroll
      for (int c = 0; c < CW; c++) {
        const int sb = base + 64 * c;  // this sub-chunk's first instance (uniform)
        if (sb >= n) break;
        if (c > 0 && wave_all(done)) break;
        const Splat& s = sc[c];
        n_loaded += (uint32_t)min(64, n - sb);
        const bool valid = sb + lane < n;
        const bool touch = valid && extent_touches(s.a.x, s.a.y, unpack_extent(s.c.w), x0, x1, y0, y1) &&
                           ellipse_touches_q(s.a.x, s.a.y, s.a.z, s.a.w, s.b.x, kq[c], x0, x1, y0, y1);
        uint64_t m = __ballot(touch);
        if (m == 0ull) continue;
        // the wave's sub-chunk goes to its LDS slot; survivors are then read
        // back with one uniform address (a broadcast), the next one while the
        // current one is evaluated.  LDS operations of one wave complete in
        // order, so no barrier is needed.
        stage[lane] = s;
        __builtin_amdgcn_wave_barrier();
        // straight-line per lane: every lane evaluates and selects commit the
        // update (the tests of the reference's loop, in its order: skipped
        // pair, alpha below 1/255, transmittance below 1e-4), so a survivor
        // costs no exec-mask branches
        auto eval = [&](const Splat& q, int j) {
          const float dx = q.a.x - pxf, dy = q.a.y - pyf;
          const float pw = power_of(q.a.z, q.a.w, q.b.x, dx, dy);
          const bool live = !done;
          n_eval += live ? 1u : 0u;
          const bool in_range = live && !(pw > 0.0f || pw < q.b.w);
          const float G = exp_neg(pw);
          const float al = fminf(0.99f, q.b.y * G);
          const bool hit = in_range && !(al < a_min);
          const float nT = T_ * (1.0f - al);
          const bool saturate = nT < 1e-4f;
          const bool apply = hit && !saturate;
          done = done || (hit && saturate);
          const float aT = al * T_;
          C0 = apply ? __builtin_fmaf(q.c.x, aT, C0) : C0;
          C1 = apply ? __builtin_fmaf(q.c.y, aT, C1) : C1;
          C2 = apply ? __builtin_fmaf(q.c.z, aT, C2) : C2;
          A = apply ? A + aT : A;
          Dd = apply ? __builtin_fmaf(q.b.z, aT, Dd) : Dd;
          T_ = apply ? nT : T_;
          last = apply ? lbase + (uint32_t)(sb + j + 1) : last;
          n_con += apply ? 1u : 0u;
          pairs += (uint32_t)__builtin_popcountll(__ballot(apply));  // scalar
        };
        // two survivors in flight: the next one's read overlaps this one's work
        int ja = __builtin_ctzll(m);
        m &= m - 1;
        Splat qa = stage[ja], qb;
        for (;;) {
          const int jb = m ? __builtin_ctzll(m) : -1;
          if (jb >= 0) {
            qb = stage[jb];
            m &= m - 1;
          }
          eval(qa, ja);
          if (jb < 0 || wave_all(done)) break;
          ja = m ? __builtin_ctzll(m) : -1;
          if (ja >= 0) {
            qa = stage[ja];
            m &= m - 1;
          }
          eval(qb, jb);
          if (ja < 0 || wave_all(done)) break;
        }
        __builtin_amdgcn_wave_barrier();  // reads of this sub-chunk precede the next one's writes
      }
    }
    const bool open = MODE == BLEND_NEAR && !wave_all(done);
    const bool unfinished = open && !((hint_word >> (tile & 31u)) & 1u);
    if (a.tile_cost && pairs && lane == 0 && (fd.flags & FRAME_BACKWARD))
      atomicAdd(a.tile_cost + tile, pairs);  // this pass's contributing pairs of the block
    if (open && lane == 0) atomicOr(a.open_bm + (tile >> 5), 1u << (tile & 31u));
    if (unfinished) {
      if (inside) {
        col[lp] = C0;
        col[HW + lp] = C1;
        col[2 * HW + lp] = C2;
        a.alpha[pid] = A;
        a.depth[pid] = Dd;
        a.n_contrib[pid] = last | (done ? kDoneBit : 0u);
        a.final_T[pid] = T_;
      }
      if (lane == 0) a.open_units[atomicAdd(a.n_open_units, 1u)] = unit;
    } else if (inside) {
      col[lp] = __builtin_fmaf(T_, fd.bg[0], C0);
      col[HW + lp] = __builtin_fmaf(T_, fd.bg[1], C1);
      col[2 * HW + lp] = __builtin_fmaf(T_, fd.bg[2], C2);
      a.alpha[pid] = A;
      a.depth[pid] = Dd;
      a.n_contrib[pid] = last;
      a.final_T[pid] = T_;
    }
  }
  __builtin_amdgcn_s_setprio(0);
  if (a.stats) {  // evaluated / contributing pairs and loaded instances
    unsigned long long e = n_eval, c = n_con, kf = (lane == 0) ? n_loaded : 0u;
    for (int off = 32; off >= 1; off >>= 1) {
      e += __shfl_xor(e, off, 64);
      c += __shfl_xor(c, off, 64);
      kf += __shfl_xor(kf, off, 64);
    }
    if (lane == 0) {
      atomicAdd(a.stats + 0, e);
      atomicAdd(a.stats + 1, c);
      if (kf) atomicAdd(a.stats + 2, kf);
    }
  }
}

}  // namespace w2w
