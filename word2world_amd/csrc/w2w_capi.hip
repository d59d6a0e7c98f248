// C ABI of the rasterizer (declarations and contract: include/w2w_raster.h).
//
// Host side only sequences launches on the caller's stream: it carves the
// caller-owned workspace, never allocates, frees or synchronises, and never
// reads K back (kernels that depend on K read it from the device status
// word).  Grids are sized by the workspace capacities and every data-dependent
// count is read on the device, so a whole forward or backward can be captured
// into a HIP graph and replayed for any batch that fits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "../../include/w2w_raster.h"
#include "w2w_adam.hpp"
#include "w2w_backward.hpp"
#include "w2w_densify.hpp"
#include "w2w_device.hpp"
#include "w2w_forward.hpp"
#include "w2w_sort.hpp"
#include "w2w_step.hpp"
#include "w2w_voxel.hpp"

using namespace w2w;

namespace {

constexpr size_t kAlign = 256;
inline size_t align_up(size_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }
inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// ---- optional per-kernel timing with HIP events on the launch stream ----
// Records an event pair around every launch whose kernel id is enabled in
// the mask, into a fixed pool set up by w2w_prof_enable (host-side state,
// not thread-safe; meant for benchmarks and profiling runs).
struct Prof {
  int capacity = 0;
  int used = 0;
  uint32_t mask = 0;
  hipEvent_t* start = nullptr;
  hipEvent_t* stop = nullptr;
  int* kid = nullptr;
} g_prof;

// optional device counters for the blend kernels (w2w_stats_enable)
unsigned long long* g_stats = nullptr;

struct ProfSpan {
  int slot = -1;
  hipStream_t s;
  ProfSpan(int kid, hipStream_t stream) : s(stream) {
    if (g_prof.capacity && (g_prof.mask >> kid & 1u) && g_prof.used < g_prof.capacity) {
      slot = g_prof.used++;
      g_prof.kid[slot] = kid;
      (void)hipEventRecord(g_prof.start[slot], s);
    }
  }
  ~ProfSpan() {
    if (slot >= 0) (void)hipEventRecord(g_prof.stop[slot], s);
  }
};

void fill_u32(uint32_t* ptr, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 4096);
  ProfSpan ps(W2W_KID_FILL, s);
  hipLaunchKernelGGL(k_fill_u32, dim3(blocks), dim3(256), 0, s, ptr, v, n);
}

// Collects fills and launches them together (k_fill_ranges); a full set is
// launched before the next range is added.  Stream order: the fills run
// where launch() is called.
struct FillBatch {
  FillRanges f{};
  uint64_t longest = 0;
  void add(uint32_t* p, uint64_t n, uint32_t v, hipStream_t s) {
    if (n == 0) return;
    if (f.n == kFillRanges) launch(s);
    f.ptr[f.n] = p; f.len[f.n] = n; f.val[f.n] = v; f.n++;
    longest = std::max(longest, n);
  }
  void launch(hipStream_t s) {
    if (f.n == 0) return;
    const uint32_t blocks = (uint32_t)std::min<uint64_t>((longest + 255) / 256, 4096);
    ProfSpan ps(W2W_KID_FILL, s);
    hipLaunchKernelGGL(k_fill_ranges, dim3(blocks), dim3(256), 0, s, f);
    f = FillRanges{};
    longest = 0;
  }
};

int launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? W2W_OK : W2W_ERR_HIP;
}

template <typename T>
T* at(void* base, size_t off) { return reinterpret_cast<T*>(static_cast<char*>(base) + off); }

int tile_bits(uint64_t T) {
  int b = 0;
  while ((1ull << b) < T) b++;
  return b < 1 ? 1 : b;
}

size_t table_bytes(uint32_t scenes_cap, uint32_t frames_cap) {
  return TABLE_HEADER_BYTES + (size_t)scenes_cap * sizeof(SceneDesc) +
         (size_t)frames_cap * sizeof(FrameDesc);
}

// Workspace plan: a pure function of the capacities.
struct Plan {
  W2WLayout L;
  W2WCaps c;
  uint32_t dtiles, stiles, ctiles;
  int tpasses, tlast_bits;
  bool key16;  // tile keys of 16 bits (at most 65536 tiles)
  size_t status, table, dkey0, dkey1, dkey2, id1, id2, dhist, dstatus, slb;  // geom
  size_t tkey0, tkey1, tid0, tid1, thist, tstatus;               // binning
  // depth-sliced binning (c.depth_split in 1..999)
  bool sliced;
  size_t ocount, boff, slb_far, open_bm, hint_bm, open_units, hint_sat, big;  // geom
  size_t tid_far;                         // binning: the far sort's second id buffer
  size_t ranges_far;                      // image
  size_t tile_cost, tile_order;           // image: the backward's tile order
  size_t rtot, rlb;                       // geom: k_preprocess's tile sums, their scan's look-back words
  uint32_t rtiles;                        //   tiles of PRE_TILE items
  // the near tile sort ends in (L.sorted_tiles, L.sorted_ids); its other
  // buffer pair is free afterwards and takes the far slice's emission
  size_t scratch_keys() const { return (tpasses & 1) ? tkey0 : tkey1; }
  size_t scratch_ids() const { return (tpasses & 1) ? tid0 : tid1; }
  // the far sort runs from the scratch pair; it ends there after an even
  // number of passes, else in (the near sorted keys, tid_far).  The near
  // sorted keys are free once the near ranges are built.
  size_t far_keys() const { return (tpasses & 1) ? L.sorted_tiles : scratch_keys(); }
  size_t far_ids() const { return (tpasses & 1) ? tid_far : scratch_ids(); }
  size_t far_scratch_keys() const { return (tpasses & 1) ? scratch_keys() : L.sorted_tiles; }
  size_t far_scratch_ids() const { return (tpasses & 1) ? scratch_ids() : tid_far; }
  uint64_t units_per_tile() const { return (uint64_t)(c.tile / kUnit) * (uint64_t)(c.tile / kUnit); }
};

bool make_plan(const W2WCaps& c, Plan& p) {
  if (c.scenes < 1 || c.frames < 1 || c.items < 0 || c.gaussians < 0 || c.tiles < 1 ||
      c.pixels < 1 || c.capacity < 0)
    return false;
  if (c.tile != 8 && c.tile != 16 && c.tile != 32) return false;
  if (c.chunk != 0 && c.chunk != 64 && c.chunk != 128 && c.chunk != 256) return false;
  // the forward blend numbers 8x8 blocks (and quads of them) in 30 bits
  if ((uint64_t)c.tiles * (uint64_t)(c.tile / 8) * (uint64_t)(c.tile / 8) >= (1ull << 30)) return false;
  if (c.depth_split < 0) return false;
  if (c.items >= (int64_t)LB_VAL || c.capacity >= (int64_t)LB_VAL || c.tiles >= (int64_t)LB_VAL ||
      c.gaussians >= (int64_t)LB_VAL)
    return false;
  p.c = c;
  W2WLayout& L = p.L;
  std::memset(&L, 0, sizeof(L));
  p.dtiles = cdiv((uint64_t)c.items, SORT_TILE);
  p.stiles = cdiv((uint64_t)c.items, SCAN_TILE);
  p.ctiles = cdiv((uint64_t)c.capacity, SORT_TILE);
  const int tb = tile_bits((uint64_t)c.tiles);
  p.tpasses = (tb + 7) / 8;
  p.tlast_bits = tb - 8 * (p.tpasses - 1);
  p.key16 = tb <= 16;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t a = o; o = align_up(o + bytes); return a; };
  const size_t I = (size_t)c.items;
  p.status = take(ST_WORDS * 4);
  p.table = take(table_bytes(1, 1));
  L.splats = take(sizeof(SplatRec) * I);
  L.radii = take(4 * I);
  L.rect = take(8 * I);
  L.counts = take(4 * I);
  p.dkey0 = take(4 * I);
  p.dkey1 = take(4 * I);
  L.perm = take(4 * I);
  p.id1 = take(4 * I);
  p.dkey2 = take(4 * I);  // the depth sort's third pair (its top pass skipped)
  p.id2 = take(4 * I);
  L.offsets = take(4 * I);
  L.item_frame = take(4 * I);
  p.dhist = take(4 * 4 * RADIX);
  p.dstatus = take((size_t)4 * p.dtiles * RADIX * 4);
  p.slb = take((size_t)p.stiles * 4 + 4);
  // each item's first backward record: the scanned tile sums of
  // k_preprocess's item tiles (L.item_off) plus the word in its splat record
  p.rtiles = cdiv((uint64_t)c.items, PRE_TILE);
  p.rtot = take((size_t)p.rtiles * 4);
  L.item_off = take((size_t)p.rtiles * 4);
  p.rlb = take((size_t)cdiv(p.rtiles, SCAN_TILE) * 4 + 4);
  L.item_grad = take(20 * I);
  p.sliced = c.depth_split > 0 && c.depth_split < 1000;
  p.ocount = p.boff = p.slb_far = p.open_bm = p.hint_bm = p.open_units = p.hint_sat = p.big = 0;
  p.tid_far = p.ranges_far = 0;
  if (p.sliced) {
    p.ocount = take(4 * I);
    p.boff = take(4 * I);
    p.slb_far = take((size_t)p.stiles * 4 + 4);
    p.open_bm = take(4 * (size_t)cdiv((uint64_t)c.tiles, 32));
    p.hint_bm = take(4 * (size_t)cdiv((uint64_t)c.tiles, 32));
    p.hint_sat = take(4 * (size_t)c.tiles);
    p.big = take(4 * I);
    p.open_units = take(4 * (size_t)c.tiles * p.units_per_tile());
  }
  L.geom_bytes = o;
  o = 0;
  const size_t C = (size_t)c.capacity;
  p.tkey0 = take(4 * C);
  p.tkey1 = take(4 * C);
  p.tid0 = take(4 * C);
  p.tid1 = take(4 * C);
  p.thist = take(4 * (size_t)kHistCopies * kHistCopyStride);
  p.tstatus = take((size_t)p.tpasses * p.ctiles * RADIX * 4);
  L.grad_rec = take(sizeof(float) * GR_FIELDS * C);
  L.grad_flag = take(4 * cdiv((uint64_t)C, 32));  // one bit per record
  if (p.sliced) p.tid_far = take(4 * C);
  L.binning_bytes = o;
  L.sorted_tiles = (p.tpasses & 1) ? p.tkey1 : p.tkey0;
  L.tile_key_bytes = p.key16 ? 2 : 4;
  L.sorted_ids = (p.tpasses & 1) ? p.tid1 : p.tid0;
  o = 0;
  L.ranges = take(8 * (size_t)c.tiles);
  L.n_contrib = take(4 * (size_t)c.pixels);
  L.final_T = take(4 * (size_t)c.pixels);
  if (p.sliced) p.ranges_far = take(8 * (size_t)c.tiles);
  p.tile_cost = take(4 * (size_t)c.tiles);
  p.tile_order = take(4 * (size_t)c.tiles);
  L.image_bytes = o;
  L.table = p.table;
  return true;
}

bool single_caps(int32_t P, int32_t W, int32_t H, int32_t ts, int64_t cap, W2WCaps& c, int32_t chunk = 0) {
  if (P < 0 || W <= 0 || H <= 0 || (ts != 8 && ts != 16 && ts != 32)) return false;
  const int64_t tx = (W + ts - 1) / ts, ty = (H + ts - 1) / ts;
  if (tx > 65535 || ty > 65535) return false;
  c.scenes = 1; c.frames = 1; c.tile = ts; c.chunk = chunk; c.reserved = 0; c.depth_split = 0;
  c.items = P; c.gaussians = P; c.tiles = tx * ty; c.pixels = (int64_t)W * H; c.capacity = cap;
  return true;
}

bool ws_ok(const W2WWorkspace* ws, const Plan& p) {
  return ws && ws->geom && ws->binning && ws->image && ws->geom_bytes >= p.L.geom_bytes &&
         ws->binning_bytes >= p.L.binning_bytes && ws->image_bytes >= p.L.image_bytes &&
         ws->capacity >= p.c.capacity;
}

bool frame_ok(const W2WFrame* f) {
  if (!f) return false;
  if (f->D < 0 || f->D > 3 || f->M < (f->D + 1) * (f->D + 1)) return false;
  if (f->chunk != 0 && f->chunk != 64 && f->chunk != 128 && f->chunk != 256) return false;
  if (!(f->focal_x > 0.0f) || !(f->focal_y > 0.0f)) return false;
  if (!f->viewmatrix || !f->projmatrix || !f->campos || !f->background) return false;
  if (f->P > 0 && (!f->means || !f->scales || !f->rotations || !f->opacities || !f->sh)) return false;
  return true;
}

// one-scene, one-frame table written on the device from the caller's
// device-resident camera (no host read-back)
__global__ void k_make_single(void* table, const float* __restrict__ view,
                              const float* __restrict__ proj, const float* __restrict__ campos,
                              const float* __restrict__ bg, uint32_t P, int32_t M, int32_t D,
                              int32_t W, int32_t H, int32_t ts, float fx, float fy) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  BatchHeader* h = static_cast<BatchHeader*>(table);
  SceneDesc* sd = reinterpret_cast<SceneDesc*>(static_cast<char*>(table) + TABLE_HEADER_BYTES);
  FrameDesc* fd = reinterpret_cast<FrameDesc*>(static_cast<char*>(table) + TABLE_HEADER_BYTES +
                                               sizeof(SceneDesc));
  const int32_t tx = (W + ts - 1) / ts, ty = (H + ts - 1) / ts;
  h->n_scenes = 1; h->n_frames = 1; h->n_items = P; h->n_gaussians = P;
  h->n_tiles = (uint32_t)(tx * ty); h->pad0 = h->pad1 = h->pad2 = 0;
  h->n_pixels = (uint64_t)W * H; h->pad3 = 0;
  sd->g_off = 0; sd->n = P; sd->item_off = 0; sd->frame_off = 0; sd->n_views = 1;
  sd->M = M; sd->D = D; sd->g_base = 0; sd->sh_off = 0; sd->pad2 = 0;
  for (int k = 0; k < 16; k++) {
    fd->V[k] = view[k];
    fd->F[k] = proj[k];
  }
  for (int k = 0; k < 3; k++) {
    fd->cp[k] = campos[k];
    fd->bg[k] = bg[k];
  }
  fd->fx = fx; fd->fy = fy; fd->W = W; fd->H = H; fd->tiles_x = tx; fd->tiles_y = ty;
  fd->tile_off = 0; fd->scene = 0; fd->view = 0; fd->flags = FRAME_BACKWARD;
  fd->pix_off = 0; fd->target_off = 0;
}

// persistent blend grids: enough workgroups to fill every CU several times
// over; each loops claiming tiles from a ticket
constexpr uint32_t kBlendGrid = 256 * 8;
// Emission grids: every workgroup adds its LDS histogram of the tile keys to
// the sort's global one at its end (device-scope atomics on a few hundred
// words, executed memory-side), so the grids stay near one resident round.
constexpr uint32_t kEmitGrid = 2048;
constexpr uint32_t kEmitFarGrid = 4096;

// workgroups of four independent waves, one 8x8 pixel block per wave at a time
template <int TS>
uint32_t blend_grid(uint32_t tiles_cap) {
  const uint64_t units = (uint64_t)tiles_cap * units_per_tile<TS>();
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((units + 3) / 4, kBlendGrid));
}

template <int TS, int CW>
void launch_blend(const BlendArgs& a, int mode, uint32_t tiles_cap, hipStream_t s) {
  ProfSpan ps(mode == BLEND_FAR ? W2W_KID_BLEND_FAR : W2W_KID_BLEND, s);
  const dim3 grid(blend_grid<TS>(tiles_cap));
  if (mode == BLEND_NEAR)
    hipLaunchKernelGGL((k_blend<TS, BLEND_NEAR, CW>), grid, dim3(256), 0, s, a);
  else if (mode == BLEND_FAR)
    hipLaunchKernelGGL((k_blend<TS, BLEND_FAR, CW>), grid, dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_blend<TS, BLEND_FULL, CW>), grid, dim3(256), 0, s, a);
}

template <int TS>
void launch_blend_chunk(const BlendArgs& a, int mode, int chunk, uint32_t tiles_cap, hipStream_t s) {
  switch (chunk) {
    case 128: launch_blend<TS, 2>(a, mode, tiles_cap, s); break;
    case 256: launch_blend<TS, 4>(a, mode, tiles_cap, s); break;
    default: launch_blend<TS, 1>(a, mode, tiles_cap, s); break;
  }
}

void launch_blend_tile(const BlendArgs& a, int mode, const W2WCaps& c, uint32_t tiles_cap, hipStream_t s) {
  switch (c.tile) {
    case 8: launch_blend_chunk<8>(a, mode, c.chunk, tiles_cap, s); break;
    case 16: launch_blend_chunk<16>(a, mode, c.chunk, tiles_cap, s); break;
    default: launch_blend_chunk<32>(a, mode, c.chunk, tiles_cap, s); break;
  }
}

// tile edges 8 and 16: one wave per tile (workgroups of four independent
// waves); edge 32: one workgroup per tile, one wave per 8x8 block of it
template <int TS>
void launch_blend_bwd(const BlendBwdArgs& a, uint32_t tiles_cap, hipStream_t s) {
  ProfSpan ps(W2W_KID_BLEND_BWD, s);
  const bool aux = a.dL_dalpha || a.dL_ddepth;
  if constexpr (TS <= 16) {
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((tiles_cap + 3) / 4, kBlendGrid));
    if (aux && a.stats)
      hipLaunchKernelGGL((k_blend_bwd_tile<TS, true, true>), dim3(grid), dim3(256), 0, s, a);
    else if (aux)
      hipLaunchKernelGGL((k_blend_bwd_tile<TS, true, false>), dim3(grid), dim3(256), 0, s, a);
    else if (a.stats)
      hipLaunchKernelGGL((k_blend_bwd_tile<TS, false, true>), dim3(grid), dim3(256), 0, s, a);
    else
      hipLaunchKernelGGL((k_blend_bwd_tile<TS, false, false>), dim3(grid), dim3(256), 0, s, a);
  } else {
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(tiles_cap, kBlendGrid));
    if (aux)
      hipLaunchKernelGGL((k_blend_bwd<TS, true>), dim3(grid), dim3(64 * units_per_tile<TS>()), 0, s, a);
    else
      hipLaunchKernelGGL((k_blend_bwd<TS, false>), dim3(grid), dim3(64 * units_per_tile<TS>()), 0, s, a);
  }
}

// the sort histogram's grid: its workgroups flush their digit counts with
// global atomics, one per nonzero bin, so keys with few distinct digits
// (depth bits; the far lists' few open tiles) take fewer workgroups
constexpr uint32_t kHistGrid = 1024;
constexpr uint32_t kShortHistGrid = 256;

template <typename KT>
void sort_passes(const KT* kin, const uint32_t* vin, KT* kA, uint32_t* vA, KT* kB, uint32_t* vB,
                 const uint32_t* d_n, uint32_t n_cap, int npasses, int last_bits, bool write_last_keys,
                 uint32_t* hist, uint32_t* status, uint32_t* tickets, int kid_hist, int kid_pass,
                 hipStream_t s, bool hist_ready = false, uint32_t hist_grid = kHistGrid, KT* kC = nullptr,
                 uint32_t* vC = nullptr, uint32_t copies = 1, const uint32_t* top_differs = nullptr,
                 uint32_t top_digit = 0) {
  const uint32_t ntiles = cdiv(n_cap, SORT_TILE);
  if (ntiles == 0) return;
  const uint32_t hgrid = std::min<uint32_t>(cdiv(n_cap, 256), hist_grid);
  {
    ProfSpan ps(kid_hist, s);
    hipLaunchKernelGGL(k_sort_hist<KT>, dim3(hgrid), dim3(256), 0, s, hist_ready ? nullptr : kin, d_n,
                       n_cap, 0, npasses, last_bits, hist, status, ntiles, tickets, copies, top_differs,
                       top_digit);
  }
  const uint32_t pgrid = std::min<uint32_t>(ntiles, 1024);  // 2 workgroups of 512 per CU (LDS)
  const KT* ks = kin;
  const uint32_t* vs = vin;
  // with a third buffer pair C and four passes, a top pass found to be the
  // identity on the device is skipped: the other three go kin -> C -> B -> A
  const bool skippable = kC && vC && npasses == 4;
  const KT* const skin[3] = {kin, kC, kB};
  const uint32_t* const svin[3] = {vin, vC, vB};
  KT* const skout[3] = {kC, kB, kA};
  uint32_t* const svout[3] = {vC, vB, vA};
  for (int p = 0; p < npasses; p++) {
    const bool toA = ((npasses - 1 - p) % 2) == 0;
    KT* kd = toA ? kA : kB;
    uint32_t* vd = toA ? vA : vB;
    const int nbits = (p == npasses - 1) ? last_bits : 8;
    uint32_t* st = status + (size_t)p * ntiles * RADIX;
    TopSkip<KT> sk{nullptr, 0, nullptr, nullptr, nullptr, nullptr};
    if (skippable) {
      sk.top_hist = hist + (npasses - 1) * RADIX;
      sk.is_top = p == npasses - 1;
      if (p < npasses - 1) {
        sk.kin = skin[p];
        sk.vin = svin[p];
        sk.kout = skout[p];
        sk.vout = svout[p];
      }
    }
    ProfSpan ps(kid_pass, s);
    if (p < npasses - 1 || write_last_keys)
      hipLaunchKernelGGL((k_sort_pass<KT, true>), dim3(pgrid), dim3(SORT_THREADS), 0, s, ks, vs, kd,
                         vd, d_n, n_cap, 8 * p, nbits, hist + p * RADIX, st, tickets + p, sk);
    else
      hipLaunchKernelGGL((k_sort_pass<KT, false>), dim3(pgrid), dim3(SORT_THREADS), 0, s, ks, vs,
                         kd, vd, d_n, n_cap, 8 * p, nbits, hist + p * RADIX, st, tickets + p, sk);
    ks = kd;
    vs = vd;
  }
}

struct Params {
  const float* means;
  const float* scales;
  const float* rots;
  const float* opac;
  const float* sh;
};

// the forward launch sequence over a table already in device memory
// Near lists (or the one pass): instance emission in depth order, the
// stable tile sort and the per-tile ranges, with KT-wide tile keys.
template <typename KT>
void bin_near(const Plan& p, const void* table, void* G, void* B, void* I, uint32_t* status, uint32_t items_cap,
              uint32_t cap, uint32_t tiles_cap, const uint32_t* d_near_end, const uint32_t* d_kn, hipStream_t s) {
  const W2WCaps& c = p.c;
  {
    ProfSpan ps(W2W_KID_EMIT, s);
    hipLaunchKernelGGL(k_emit<KT>, dim3(std::min<uint32_t>(cdiv(items_cap, 256), kEmitGrid)),
                       dim3(256), 0, s, table, (uint32_t)c.scenes, at<uint32_t>(G, p.L.perm),
                       at<uint32_t>(G, p.L.offsets), at<uint32_t>(G, p.L.counts),
                       at<uint2>(G, p.L.rect), at<uint32_t>(G, p.L.item_frame), items_cap, cap,
                       d_near_end,
                       at<KT>(B, p.tkey0), at<uint32_t>(B, p.tid0), at<uint32_t>(B, p.thist),
                       p.tpasses, p.tlast_bits);
  }
  if (p.sliced) {
    ProfSpan ps(W2W_KID_EMIT_FAR, s);
    hipLaunchKernelGGL(k_emit_far<KT>, dim3(std::min<uint32_t>(cdiv(items_cap, 256), kEmitFarGrid)), dim3(256), 0, s,
                       table, (uint32_t)c.scenes, tiles_cap, items_cap, cap, status, at<uint32_t>(G, p.L.perm),
                       at<uint32_t>(G, p.L.offsets), status + ST_KNEAR_RAW, at<uint2>(G, p.L.rect),
                       at<uint32_t>(G, p.L.item_frame), at<uint32_t>(G, p.hint_bm), nullptr, nullptr,
                       at<KT>(B, p.tkey0), at<uint32_t>(B, p.tid0), at<uint32_t>(B, p.thist),
                       p.tpasses, p.tlast_bits, nullptr, at<uint32_t>(G, p.big), status + ST_BIG_H);
    hipLaunchKernelGGL(k_emit_far_big<KT>, dim3(1024), dim3(256), 0, s, table, (uint32_t)c.scenes, cap,
                       at<uint32_t>(G, p.L.perm), at<uint32_t>(G, p.L.offsets), at<uint2>(G, p.L.rect),
                       at<uint32_t>(G, p.L.item_frame), at<uint32_t>(G, p.hint_bm), nullptr, nullptr,
                       at<uint32_t>(G, p.big), status + ST_BIG_H, items_cap, at<KT>(B, p.tkey0),
                       at<uint32_t>(B, p.tid0), at<uint32_t>(B, p.thist), p.tpasses, p.tlast_bits, nullptr);
  }
  sort_passes<KT>(at<KT>(B, p.tkey0), at<uint32_t>(B, p.tid0),
                  at<KT>(B, p.L.sorted_tiles), at<uint32_t>(B, p.L.sorted_ids),
                  at<KT>(B, p.scratch_keys()), at<uint32_t>(B, p.scratch_ids()), d_kn,
                  cap, p.tpasses, p.tlast_bits, true, at<uint32_t>(B, p.thist),
                  at<uint32_t>(B, p.tstatus), status + ST_TTICKET, W2W_KID_TSORT_HIST,
                  W2W_KID_TSORT_PASS, s, /*hist_ready=*/true, kHistGrid, nullptr, nullptr, kHistCopies);
  const uint32_t rgrid = std::min<uint32_t>(cdiv(cap, 256), 4096);
  ProfSpan ps(W2W_KID_RANGES, s);
  hipLaunchKernelGGL(k_ranges<KT>, dim3(rgrid), dim3(256), 0, s, at<KT>(B, p.L.sorted_tiles),
                     d_kn, table, cap, tiles_cap, at<uint2>(I, p.L.ranges));
}

// Far lists of depth-sliced binning (the open tiles the hint missed):
// emission at the compact offsets, the tile sort and the ranges.  The sort
// counts the keys itself: they fall on a few open tiles, whose histogram
// bins every emitting workgroup would otherwise flush with contending
// atomics.
template <typename KT>
void bin_far(const Plan& p, const void* table, void* G, void* B, void* I, uint32_t* status, uint32_t items_cap,
             uint32_t cap, uint32_t tiles_cap, const uint32_t* open_bm, const uint32_t* hint_bm,
             uint2* ranges_far, hipStream_t s) {
  const W2WCaps& c = p.c;
  {
    ProfSpan ps(W2W_KID_EMIT_FAR, s);
    hipLaunchKernelGGL(k_emit_far<KT>, dim3(std::min<uint32_t>(cdiv(items_cap, 256), kEmitFarGrid)), dim3(256), 0, s,
                       table, (uint32_t)c.scenes, tiles_cap, items_cap, cap, status, at<uint32_t>(G, p.L.perm),
                       at<uint32_t>(G, p.boff), status + ST_KB, at<uint2>(G, p.L.rect),
                       at<uint32_t>(G, p.L.item_frame), open_bm, hint_bm, nullptr,
                       at<KT>(B, p.scratch_keys()), at<uint32_t>(B, p.scratch_ids()), at<uint32_t>(B, p.thist),
                       0, p.tlast_bits, status + ST_OPEN_UNITS, at<uint32_t>(G, p.big), status + ST_BIG_F);
    hipLaunchKernelGGL(k_emit_far_big<KT>, dim3(1024), dim3(256), 0, s, table, (uint32_t)c.scenes, cap,
                       at<uint32_t>(G, p.L.perm), at<uint32_t>(G, p.boff), at<uint2>(G, p.L.rect),
                       at<uint32_t>(G, p.L.item_frame), open_bm, hint_bm, nullptr, at<uint32_t>(G, p.big),
                       status + ST_BIG_F, items_cap, at<KT>(B, p.scratch_keys()),
                       at<uint32_t>(B, p.scratch_ids()), at<uint32_t>(B, p.thist), 0, p.tlast_bits,
                       status + ST_OPEN_UNITS);
  }
  sort_passes<KT>(at<KT>(B, p.scratch_keys()), at<uint32_t>(B, p.scratch_ids()),
                  at<KT>(B, p.far_keys()), at<uint32_t>(B, p.far_ids()),
                  at<KT>(B, p.far_scratch_keys()), at<uint32_t>(B, p.far_scratch_ids()),
                  status + ST_KB, cap, p.tpasses, p.tlast_bits, true, at<uint32_t>(B, p.thist),
                  at<uint32_t>(B, p.tstatus), status + ST_TTICKET, W2W_KID_TSORT_HIST,
                  W2W_KID_TSORT_PASS, s, /*hist_ready=*/false, kShortHistGrid);
  ProfSpan ps(W2W_KID_RANGES, s);
  hipLaunchKernelGGL(k_ranges<KT>, dim3(std::min<uint32_t>(cdiv(cap, 256), 4096)), dim3(256), 0, s,
                     at<KT>(B, p.far_keys()), status + ST_KB, table, cap, tiles_cap, ranges_far);
}

int run_forward(const Plan& p, const void* table, const W2WWorkspace* ws, const Params& prm,
                float* out_color, float* out_depth, float* out_alpha, int32_t* out_radii,
                hipStream_t s) {
  void* G = ws->geom;
  void* B = ws->binning;
  void* I = ws->image;
  uint32_t* status = at<uint32_t>(G, p.status);
  const W2WCaps& c = p.c;
  const uint32_t items_cap = (uint32_t)c.items;
  const uint32_t cap = (uint32_t)c.capacity;
  const uint32_t tiles_cap = (uint32_t)c.tiles;
  const uint32_t* d_items = &static_cast<const BatchHeader*>(table)->n_items;

  // depth-sliced binning: the hint is the last forward's set of tiles its
  // near lists left unfinished, grown by one tile (a guess that only steers
  // work; any content gives the same outputs), taken before the open set is
  // cleared below
  const bool hinted = p.sliced && cap > 0 && items_cap > 0;
  const size_t tile_words = cdiv((uint64_t)c.tiles, 32);
  if (hinted)
    hipLaunchKernelGGL(k_dilate_tiles, dim3((uint32_t)std::min<uint64_t>(cdiv((uint64_t)c.tiles, 256), 4096)), dim3(256),
                       0, s, table, (uint32_t)c.scenes, (uint32_t)c.tiles, at<uint32_t>(G, p.open_bm),
                       at<uint32_t>(G, p.hint_bm));
  FillBatch fb;
  fb.add(status, ST_WORDS, 0u, s);
  fb.add(at<uint32_t>(I, p.L.ranges), 2 * (uint64_t)c.tiles, 0u, s);
  if (items_cap > 0) {
    fb.add(at<uint32_t>(G, p.rlb), cdiv(p.rtiles, SCAN_TILE) + 1, 0u, s);
    fb.add(at<uint32_t>(G, p.dhist), 4 * RADIX, 0u, s);
    fb.add(at<uint32_t>(G, p.slb), p.stiles + 1, 0u, s);
    fb.add(at<uint32_t>(B, p.thist), kHistCopies * kHistCopyStride, 0u, s);
  }
  if (hinted) fb.add(at<uint32_t>(G, p.open_bm), tile_words, 0u, s);
  fb.launch(s);
  // the forward blend's block order: every tile, heaviest first by the last
  // forward's per-tile costs (any content is a valid order), then this
  // forward's costs are cleared for its blend to count
  const bool ordered = items_cap > 0 && cap > 0;
  if (ordered) {
    TileOrderArgs oa;
    oa.table = table; oa.tiles_cap = tiles_cap;
    oa.cost = at<uint32_t>(I, p.tile_cost);
    oa.hist = status + ST_COST_HIST;
    oa.cursor = status + ST_COST_CURSOR;
    oa.n_order = status + ST_ORDER_N;
    oa.order = at<uint32_t>(I, p.tile_order);
    oa.include_zero = 1u;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(cdiv((uint64_t)c.tiles, 256), 512));
    ProfSpan ps(W2W_KID_TILE_ORDER, s);
    hipLaunchKernelGGL(k_tile_cost_hist, dim3(grid), dim3(256), 0, s, oa);
    hipLaunchKernelGGL(k_tile_order, dim3(grid), dim3(256), 0, s, oa);
  }
  {
    FillBatch fc;
    fc.add(at<uint32_t>(I, p.tile_cost), (uint64_t)c.tiles, 0u, s);
    fc.launch(s);
  }
  if (items_cap > 0) {

    PreArgs pa;
    pa.table = table; pa.scenes_cap = (uint32_t)c.scenes; pa.items_cap = items_cap;
    pa.means = prm.means; pa.scales = prm.scales; pa.rots = prm.rots; pa.opac = prm.opac;
    pa.sh = prm.sh; pa.ts = c.tile;
    pa.splats = at<SplatRec>(G, p.L.splats);
    pa.radii = at<int32_t>(G, p.L.radii);
    pa.rect = at<uint2>(G, p.L.rect);
    pa.counts = at<uint32_t>(G, p.L.counts);
    pa.dkey = at<uint32_t>(G, p.dkey0);
    pa.ids = at<uint32_t>(G, p.L.perm);
    pa.item_frame = at<uint32_t>(G, p.L.item_frame);
    pa.hint_sat = nullptr;
    pa.hcount = nullptr;
    if (hinted) {
      // hinted tiles get their far items in their near lists too, behind the
      // near items, so depth order is kept
      hipLaunchKernelGGL(k_tile_sat, dim3((uint32_t)std::min<int32_t>(c.frames, 4096)), dim3(256), 0, s, table,
                         (uint32_t)c.scenes, (uint32_t)c.frames, at<uint32_t>(G, p.hint_bm),
                         at<uint32_t>(G, p.hint_sat));
      pa.hint_sat = at<uint32_t>(G, p.hint_sat);
      pa.hcount = at<uint32_t>(G, p.ocount);
    }
    pa.tile_tot = at<uint32_t>(G, p.rtot);
    pa.status = status;
    pa.split_permille = p.sliced ? (uint32_t)c.depth_split : 0u;
    {
      ProfSpan ps(W2W_KID_PREPROCESS, s);
      hipLaunchKernelGGL(k_preprocess, dim3(std::min<uint32_t>(p.rtiles, 8192)), dim3(PRE_TILE), 0, s, pa);
    }
    {
      // the tiles' first backward-record positions, K (every tile count
      // summed) and the capacity-overflow flag
      ProfSpan ps(W2W_KID_SCAN, s);
      hipLaunchKernelGGL(k_scan<false>, dim3(std::min<uint32_t>(cdiv(p.rtiles, SCAN_TILE), 2048)),
                         dim3(SCAN_THREADS), 0, s, at<uint32_t>(G, p.rtot), nullptr,
                         at<uint32_t>(G, p.L.item_off), at<uint32_t>(G, p.rlb) + 1, status + ST_REC_TICKET,
                         status + ST_K, status + ST_OVERFLOW, nullptr, p.rtiles, cap);
    }
    // depth sort: (depth bits, item) over the items, 4 passes (3 when every
    // key has the same top byte), result in perm
    sort_passes<uint32_t>(at<uint32_t>(G, p.dkey0), at<uint32_t>(G, p.L.perm),
                          at<uint32_t>(G, p.dkey0), at<uint32_t>(G, p.L.perm),
                          at<uint32_t>(G, p.dkey1), at<uint32_t>(G, p.id1), d_items, items_cap, 4,
                          8, false, at<uint32_t>(G, p.dhist), at<uint32_t>(G, p.dstatus),
                          status + ST_DTICKET, W2W_KID_DSORT_HIST, W2W_KID_DSORT_PASS, s, false,
                          kShortHistGrid, at<uint32_t>(G, p.dkey2), at<uint32_t>(G, p.id2), 1,
                          status + ST_DTOP, CULLED_DEPTH_KEY >> 24);
    if (cap > 0) {
      // emission offsets: the tile counts gathered in depth order and
      // scanned.  Depth-sliced binning: the near slice is binned here, the
      // far slice after the near blend (only on the tiles it leaves open);
      // far items are counted on the hinted tiles only, which places them in
      // the near lists behind the near items (depth order kept)
      const uint32_t* d_near_end = nullptr;
      const uint32_t* d_kn = status + ST_K;
      {
        ProfSpan ps(W2W_KID_SCAN, s);
        hipLaunchKernelGGL(k_scan<true>, dim3(std::min<uint32_t>(p.stiles, 2048)), dim3(SCAN_THREADS),
                           0, s, at<uint32_t>(G, p.L.counts), at<uint32_t>(G, p.L.perm),
                           at<uint32_t>(G, p.L.offsets), at<uint32_t>(G, p.slb) + 1,
                           status + ST_SCAN_TICKET, status + (p.sliced ? ST_KNEAR_RAW : ST_K), nullptr,
                           d_items, items_cap, cap, p.sliced ? at<uint32_t>(G, p.ocount) : nullptr,
                           p.sliced ? status + ST_SPLIT : nullptr);
      }
      if (p.sliced) {
        hipLaunchKernelGGL(k_near_total, dim3(1), dim3(1), 0, s, table, items_cap, at<uint32_t>(G, p.L.offsets),
                           status, cap);
        d_near_end = status + ST_SPLIT;
        d_kn = status + ST_KNEAR;
      }
      if (p.key16)
        bin_near<uint16_t>(p, table, G, B, I, status, items_cap, cap, tiles_cap, d_near_end, d_kn, s);
      else
        bin_near<uint32_t>(p, table, G, B, I, status, items_cap, cap, tiles_cap, d_near_end, d_kn, s);
    }
    if (out_radii)
      (void)hipMemcpyAsync(out_radii, at<int32_t>(G, p.L.radii), 4 * (size_t)items_cap,
                           hipMemcpyDeviceToDevice, s);
  }
  BlendArgs ba;
  ba.table = table; ba.scenes_cap = (uint32_t)c.scenes; ba.tiles_cap = tiles_cap;
  ba.items_cap = items_cap;
  ba.ticket = status + ST_BLEND_TICKET;
  ba.ranges = at<uint2>(I, p.L.ranges);
  ba.ids = at<uint32_t>(B, p.L.sorted_ids);
  ba.splats = at<SplatRec>(G, p.L.splats);
  ba.stats = g_stats;
  ba.color = out_color; ba.alpha = out_alpha; ba.depth = out_depth;
  ba.n_contrib = at<uint32_t>(I, p.L.n_contrib);
  ba.final_T = at<float>(I, p.L.final_T);
  ba.ranges_far = nullptr;
  ba.ids_far = nullptr;
  ba.open_bm = nullptr;
  ba.hint_bm = nullptr;
  ba.open_units = nullptr;
  ba.n_open_units = nullptr;
  ba.tile_cost = at<uint32_t>(I, p.tile_cost);
  ba.order = ordered ? at<uint32_t>(I, p.tile_order) : nullptr;
  if (!p.sliced || items_cap == 0 || cap == 0) {
    launch_blend_tile(ba, BLEND_FULL, c, tiles_cap, s);
    return launch_status();
  }
  // near lists: final outputs of finished blocks and hinted tiles, unfinished
  // pixel state otherwise; the open tiles (zeroed by the hint move)
  uint32_t* open_bm = at<uint32_t>(G, p.open_bm);
  const uint32_t* hint_bm = at<uint32_t>(G, p.hint_bm);
  ba.open_bm = open_bm;
  ba.hint_bm = hint_bm;
  ba.open_units = at<uint32_t>(G, p.open_units);
  ba.n_open_units = status + ST_OPEN_UNITS;  // zeroed with the status block
  launch_blend_tile(ba, BLEND_NEAR, c, tiles_cap, s);
  // far lists on the open tiles that were not hinted: counts, compact
  // offsets, emission, sort, ranges
  {
    ProfSpan ps(W2W_KID_COUNT_OPEN, s);
    hipLaunchKernelGGL(k_count_tiles, dim3(std::min<uint32_t>(cdiv(items_cap, 256), 8192)), dim3(256), 0, s,
                       table, (uint32_t)c.scenes, items_cap, status, at<uint32_t>(G, p.L.perm),
                       at<uint2>(G, p.L.rect), at<uint32_t>(G, p.L.item_frame), open_bm, hint_bm,
                       at<uint32_t>(G, p.ocount), status + ST_OPEN_UNITS);
  }
  uint2* ranges_far = at<uint2>(I, p.ranges_far);
  {
    FillBatch fb;
    fb.add(at<uint32_t>(G, p.slb_far), p.stiles + 1, 0u, s);
    fb.add(at<uint32_t>(B, p.thist), kHistCopies * kHistCopyStride, 0u, s);  // the near sort is done with it
    fb.add(reinterpret_cast<uint32_t*>(ranges_far), 2 * (uint64_t)c.tiles, 0u, s);
    fb.launch(s);
  }
  {
    ProfSpan ps(W2W_KID_SCAN, s);
    hipLaunchKernelGGL(k_scan<false>, dim3(std::min<uint32_t>(p.stiles, 2048)), dim3(SCAN_THREADS), 0, s,
                       at<uint32_t>(G, p.ocount), nullptr, at<uint32_t>(G, p.boff),
                       at<uint32_t>(G, p.slb_far) + 1, status + ST_SCANB_TICKET, status + ST_KB, nullptr,
                       d_items, items_cap, cap, nullptr, nullptr, status + ST_OPEN_UNITS);
  }
  if (p.key16)
    bin_far<uint16_t>(p, table, G, B, I, status, items_cap, cap, tiles_cap, open_bm, hint_bm, ranges_far, s);
  else
    bin_far<uint32_t>(p, table, G, B, I, status, items_cap, cap, tiles_cap, open_bm, hint_bm, ranges_far, s);
  ba.ticket = status + ST_BLENDB_TICKET;
  ba.ranges_far = ranges_far;
  ba.ids_far = at<uint32_t>(B, p.far_ids());
  launch_blend_tile(ba, BLEND_FAR, c, tiles_cap, s);
  return launch_status();
}

// The SH coefficients' optimizer step taken by the SH backward in place of
// writing dL/dSH (the batched trainer's step; ShBwdArgs has the meaning).
struct ShAdam {
  float* sh;
  float* m;
  float* v;
  const float* hyper;
  int hyper_dc, hyper_ac;
  const uint32_t* skip_if;
};

int run_backward(const Plan& p, const void* table, const W2WWorkspace* ws, const Params& prm,
                 const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha,
                 float* dL_dmeans2D, float* dL_dcolors, float* dL_dopacity, float* dL_dmeans3D,
                 float* dL_dcov3D, float* dL_dsh, float* dL_dscales, float* dL_drotations,
                 hipStream_t s, const ShAdam* sh_adam = nullptr) {
  void* G = ws->geom;
  void* B = ws->binning;
  void* I = ws->image;
  const W2WCaps& c = p.c;
  if (c.items == 0) return W2W_OK;
  uint32_t* status = at<uint32_t>(G, p.status);
  float* rec = at<float>(B, p.L.grad_rec);
  uint32_t* flag = at<uint32_t>(B, p.L.grad_flag);
  {
    FillBatch fb;
    fb.add(flag, cdiv((uint64_t)c.capacity, 32), 0u, s);
    fb.add(status + ST_BWD_TICKET, 1, 0u, s);
    fb.add(status + ST_ORDER_N, 1 + 2 * kCostBuckets, 0u, s);  // count, histogram, cursors
    fb.launch(s);
  }
  {
    // the tiles with work, heaviest first
    TileOrderArgs oa;
    oa.table = table; oa.tiles_cap = (uint32_t)c.tiles;
    oa.cost = at<uint32_t>(I, p.tile_cost);
    oa.hist = status + ST_COST_HIST;
    oa.cursor = status + ST_COST_CURSOR;
    oa.n_order = status + ST_ORDER_N;
    oa.order = at<uint32_t>(I, p.tile_order);
    oa.include_zero = 0u;
    const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>(cdiv((uint64_t)c.tiles, 256), 512));
    ProfSpan ps(W2W_KID_TILE_ORDER, s);
    hipLaunchKernelGGL(k_tile_cost_hist, dim3(grid), dim3(256), 0, s, oa);
    hipLaunchKernelGGL(k_tile_order, dim3(grid), dim3(256), 0, s, oa);
  }
  BlendBwdArgs ba;
  ba.table = table; ba.scenes_cap = (uint32_t)c.scenes; ba.tiles_cap = (uint32_t)c.tiles;
  ba.items_cap = (uint32_t)c.items; ba.capacity = (uint32_t)c.capacity;
  ba.ticket = status + ST_BWD_TICKET;
  ba.ranges = at<uint2>(I, p.L.ranges);
  ba.ids = at<uint32_t>(B, p.L.sorted_ids);
  ba.splats = at<SplatRec>(G, p.L.splats);
  ba.rec_base = at<uint32_t>(G, p.L.item_off);
  ba.rec = rec;
  ba.flag = flag;
  ba.stats = g_stats ? g_stats + 3 : nullptr;
  ba.n_contrib = at<uint32_t>(I, p.L.n_contrib);
  ba.final_T = at<float>(I, p.L.final_T);
  ba.dL_dcolor = dL_dcolor; ba.dL_dalpha = dL_dalpha; ba.dL_ddepth = dL_ddepth;
  ba.ranges_far = p.sliced ? at<uint2>(I, p.ranges_far) : nullptr;
  ba.ids_far = p.sliced ? at<uint32_t>(B, p.far_ids()) : nullptr;
  ba.order = at<uint32_t>(I, p.tile_order);
  ba.n_order = status + ST_ORDER_N;
  ba.tile_cost = at<uint32_t>(I, p.tile_cost);
  switch (c.tile) {
    case 8: launch_blend_bwd<8>(ba, (uint32_t)c.tiles, s); break;
    case 16: launch_blend_bwd<16>(ba, (uint32_t)c.tiles, s); break;
    default: launch_blend_bwd<32>(ba, (uint32_t)c.tiles, s); break;
  }
  PreBwdArgs pa;
  pa.table = table; pa.scenes_cap = (uint32_t)c.scenes; pa.items_cap = (uint32_t)c.items;
  pa.gaussians_cap = (uint32_t)c.gaussians;
  pa.means = prm.means; pa.scales = prm.scales; pa.rots = prm.rots; pa.sh = prm.sh;
  pa.radii = at<int32_t>(G, p.L.radii);
  pa.counts = at<uint32_t>(G, p.L.counts);
  pa.rec = rec;
  pa.flag = flag;
  pa.capacity = (uint32_t)c.capacity;
  pa.splats = at<SplatRec>(G, p.L.splats);
  pa.rec_base = at<uint32_t>(G, p.L.item_off);
  pa.item_frame = at<uint32_t>(G, p.L.item_frame);
  pa.item_grad = at<float>(G, p.L.item_grad);
  pa.dL_dmeans2D = dL_dmeans2D; pa.dL_dcolors = dL_dcolors; pa.dL_dopacity = dL_dopacity;
  pa.dL_dmeans3D = dL_dmeans3D; pa.dL_dcov3D = dL_dcov3D; pa.dL_dsh = dL_dsh;
  pa.dL_dscales = dL_dscales; pa.dL_drots = dL_drotations;
  if (c.gaussians > 0) {
    ProfSpan ps(W2W_KID_PREPROCESS_BWD, s);
    hipLaunchKernelGGL(k_item_grads, dim3(std::min<uint32_t>(cdiv((uint64_t)c.items, 256), 16384)), dim3(256), 0, s,
                       pa);
    hipLaunchKernelGGL(k_preprocess_bwd, dim3(std::min<uint32_t>(cdiv((uint64_t)c.gaussians, 256), 8192)),
                       dim3(256), 0, s, pa);
  }
  if (c.gaussians > 0) {
    ShBwdArgs sa;
    sa.table = table; sa.scenes_cap = (uint32_t)c.scenes; sa.items_cap = (uint32_t)c.items;
    sa.gaussians_cap = (uint32_t)c.gaussians;
    sa.means = prm.means; sa.sh = prm.sh; sa.radii = at<int32_t>(G, p.L.radii);
    sa.dL_dcolors = dL_dcolors; sa.dL_dsh = dL_dsh; sa.dL_dmeans3D = dL_dmeans3D;
    sa.sh_out = sh_adam ? sh_adam->sh : nullptr;
    sa.adam_m = sh_adam ? sh_adam->m : nullptr;
    sa.adam_v = sh_adam ? sh_adam->v : nullptr;
    sa.hyper = sh_adam ? sh_adam->hyper : nullptr;
    sa.hyper_dc = sh_adam ? sh_adam->hyper_dc : 0;
    sa.hyper_ac = sh_adam ? sh_adam->hyper_ac : 0;
    sa.skip_if = sh_adam ? sh_adam->skip_if : nullptr;
    ProfSpan ps(W2W_KID_SH_BWD, s);
    if (sh_adam)
      hipLaunchKernelGGL(k_sh_bwd_adam, dim3(std::min<uint32_t>(cdiv((uint64_t)c.gaussians, 256), 8192)),
                         dim3(256), 0, s, sa);
    else
      hipLaunchKernelGGL(k_sh_bwd, dim3(std::min<uint32_t>(cdiv((uint64_t)c.gaussians, 256), 8192)), dim3(256), 0,
                         s, sa);
  }
  return launch_status();
}

int single_setup(const W2WFrame* f, const W2WWorkspace* ws, Plan& p, hipStream_t s) {
  W2WCaps c;
  if (!ws || !single_caps(f->P, f->W, f->H, f->tile, ws->capacity, c, f->chunk) || !make_plan(c, p))
    return W2W_ERR_INVALID;
  if (!ws_ok(ws, p)) return W2W_ERR_CAPACITY;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_make_single, dim3(1), dim3(64), 0, s, at<void>(ws->geom, p.table),
                     f->viewmatrix, f->projmatrix, f->campos, f->background, (uint32_t)f->P, f->M,
                     f->D, f->W, f->H, f->tile, f->focal_x, f->focal_y);
  return W2W_OK;
}

bool caps_ws(const W2WCaps* caps, const W2WWorkspace* ws, Plan& p, int& err) {
  err = W2W_ERR_INVALID;
  if (!caps || !ws || !make_plan(*caps, p)) return false;
  if (!ws_ok(ws, p)) {
    err = W2W_ERR_CAPACITY;
    return false;
  }
  return true;
}

}  // namespace

extern "C" {

int w2w_version(void) { return W2W_ABI_VERSION; }

int w2w_block_touches(const float* splat, const float* rect, uint64_t n, uint8_t* out) {
  if (n && (!splat || !rect || !out)) return W2W_ERR_INVALID;
  for (uint64_t i = 0; i < n; i++) {
    const float* f = splat + 6 * i;
    const float* r = rect + 4 * i;
    out[i] = ellipse_touches(f[0], f[1], f[2], f[3], f[4], f[5], r[0], r[1], r[2], r[3]) ? 1 : 0;
  }
  return W2W_OK;
}

int w2w_splat_extents(const float* conic, const float* skip, uint64_t n, uint32_t* out) {
  if (n && (!conic || !skip || !out)) return W2W_ERR_INVALID;
  for (uint64_t i = 0; i < n; i++)
    out[i] = cull_extent(conic[3 * i], conic[3 * i + 1], conic[3 * i + 2], skip[i]);
  return W2W_OK;
}

int w2w_workspace_layout(int32_t P, int32_t W, int32_t H, int32_t tile, int64_t capacity,
                         W2WLayout* out) {
  W2WCaps c;
  Plan p;
  if (!out || !single_caps(P, W, H, tile, capacity, c) || !make_plan(c, p)) return W2W_ERR_INVALID;
  *out = p.L;
  return W2W_OK;
}

int w2w_batch_workspace_layout(const W2WCaps* caps, W2WLayout* out) {
  Plan p;
  if (!out || !caps || !make_plan(*caps, p)) return W2W_ERR_INVALID;
  *out = p.L;
  return W2W_OK;
}

size_t w2w_batch_table_bytes(int32_t scenes_cap, int32_t frames_cap) {
  if (scenes_cap < 1 || frames_cap < 1) return 0;
  return table_bytes((uint32_t)scenes_cap, (uint32_t)frames_cap);
}

int w2w_batch_encode(void* host_table, size_t table_bytes_avail, const W2WCaps* caps,
                     int32_t n_scenes, const int64_t* scene_n, const int32_t* scene_views,
                     const int32_t* scene_M, const int32_t* scene_D, const float* frame_view,
                     const float* frame_proj, const float* frame_campos, const float* frame_bg,
                     const float* frame_focal, const int32_t* frame_wh, const uint32_t* frame_flags,
                     const int64_t* scene_base, const int64_t* scene_sh_off,
                     const int64_t* frame_target_off, int64_t* totals) {
  if (!host_table || !caps) return W2W_ERR_INVALID;
  const int32_t scenes_cap = caps->scenes, frames_cap = caps->frames, tile = caps->tile;
  if (scenes_cap < 1 || frames_cap < 1 || n_scenes < 0) return W2W_ERR_INVALID;
  if (n_scenes > scenes_cap) return W2W_ERR_CAPACITY;
  if (tile != 8 && tile != 16 && tile != 32) return W2W_ERR_INVALID;
  if (table_bytes_avail < table_bytes((uint32_t)scenes_cap, (uint32_t)frames_cap))
    return W2W_ERR_CAPACITY;
  std::memset(host_table, 0, table_bytes((uint32_t)scenes_cap, (uint32_t)frames_cap));
  BatchHeader* h = static_cast<BatchHeader*>(host_table);
  SceneDesc* sd = reinterpret_cast<SceneDesc*>(static_cast<char*>(host_table) + TABLE_HEADER_BYTES);
  FrameDesc* fd = reinterpret_cast<FrameDesc*>(static_cast<char*>(host_table) + TABLE_HEADER_BYTES +
                                               (size_t)scenes_cap * sizeof(SceneDesc));
  uint64_t g = 0, items = 0, frames = 0, tiles = 0, pix = 0, shf = 0;
  for (int32_t k = 0; k < n_scenes; k++) {
    if (scene_n[k] < 0 || scene_views[k] < 1 || scene_D[k] < 0 || scene_D[k] > 3 ||
        scene_M[k] < (scene_D[k] + 1) * (scene_D[k] + 1))
      return W2W_ERR_INVALID;
    if (frames + (uint64_t)scene_views[k] > (uint64_t)frames_cap) return W2W_ERR_CAPACITY;
    sd[k].g_off = (uint32_t)g; sd[k].n = (uint32_t)scene_n[k]; sd[k].item_off = (uint32_t)items;
    sd[k].frame_off = (uint32_t)frames; sd[k].n_views = (uint32_t)scene_views[k];
    sd[k].M = scene_M[k]; sd[k].D = scene_D[k];
    sd[k].g_base = scene_base ? (uint32_t)scene_base[k] : (uint32_t)g;
    sd[k].sh_off = scene_sh_off ? (uint64_t)scene_sh_off[k] : shf;
    if (scene_base && (scene_base[k] < 0 || scene_base[k] + scene_n[k] >= (int64_t)LB_VAL))
      return W2W_ERR_INVALID;
    for (int32_t v = 0; v < scene_views[k]; v++) {
      const uint64_t f = frames + v;
      FrameDesc& F = fd[f];
      std::memcpy(F.V, frame_view + 16 * f, 64);
      std::memcpy(F.F, frame_proj + 16 * f, 64);
      std::memcpy(F.cp, frame_campos + 3 * f, 12);
      std::memcpy(F.bg, frame_bg + 3 * f, 12);
      F.fx = frame_focal[2 * f]; F.fy = frame_focal[2 * f + 1];
      F.W = frame_wh[2 * f]; F.H = frame_wh[2 * f + 1];
      if (F.W <= 0 || F.H <= 0 || !(F.fx > 0.0f) || !(F.fy > 0.0f)) return W2W_ERR_INVALID;
      F.tiles_x = (F.W + tile - 1) / tile; F.tiles_y = (F.H + tile - 1) / tile;
      if (F.tiles_x > 65535 || F.tiles_y > 65535) return W2W_ERR_INVALID;
      F.tile_off = (uint32_t)tiles; F.scene = (uint32_t)k; F.view = (uint32_t)v;
      F.flags = frame_flags ? frame_flags[f] : FRAME_BACKWARD;
      F.pix_off = pix;
      F.target_off = frame_target_off ? (uint64_t)frame_target_off[f] : 3 * pix;
      tiles += (uint64_t)F.tiles_x * F.tiles_y;
      pix += (uint64_t)F.W * F.H;
    }
    g += (uint64_t)scene_n[k];
    items += (uint64_t)scene_n[k] * scene_views[k];
    frames += (uint64_t)scene_views[k];
    shf += (uint64_t)scene_n[k] * scene_M[k] * 3;
  }
  if (items >= LB_VAL || tiles >= LB_VAL || g >= LB_VAL) return W2W_ERR_CAPACITY;
  // every grid and buffer of a workspace is sized by its caps: a batch that
  // exceeds any of them would be cut short or written past its buffers
  if (items > (uint64_t)std::max<int64_t>(caps->items, 0) || g > (uint64_t)std::max<int64_t>(caps->gaussians, 0) ||
      tiles > (uint64_t)std::max<int64_t>(caps->tiles, 0) || pix > (uint64_t)std::max<int64_t>(caps->pixels, 0))
    return W2W_ERR_CAPACITY;
  h->n_scenes = (uint32_t)n_scenes; h->n_frames = (uint32_t)frames; h->n_items = (uint32_t)items;
  h->n_gaussians = (uint32_t)g; h->n_tiles = (uint32_t)tiles; h->n_pixels = pix;
  if (totals) {
    totals[0] = (int64_t)g; totals[1] = (int64_t)items; totals[2] = (int64_t)frames;
    totals[3] = (int64_t)tiles; totals[4] = (int64_t)pix; totals[5] = (int64_t)shf;
  }
  return W2W_OK;
}

int w2w_forward_batch(const void* table, const W2WCaps* caps, const W2WWorkspace* ws,
                      const float* means, const float* scales, const float* rotations,
                      const float* opacities, const float* sh, float* out_color, float* out_depth,
                      float* out_alpha, int32_t* out_radii, void* stream) {
  Plan p;
  int err;
  if (!table || !caps_ws(caps, ws, p, err)) return table ? err : W2W_ERR_INVALID;
  if (!out_color || !out_depth || !out_alpha) return W2W_ERR_INVALID;
  if (caps->items > 0 && (!means || !scales || !rotations || !opacities || !sh)) return W2W_ERR_INVALID;
  (void)hipGetLastError();
  Params prm{means, scales, rotations, opacities, sh};
  return run_forward(p, table, ws, prm, out_color, out_depth, out_alpha, out_radii,
                     (hipStream_t)stream);
}

int w2w_backward_batch(const void* table, const W2WCaps* caps, const W2WWorkspace* ws,
                       const float* means, const float* scales, const float* rotations,
                       const float* sh, const float* dL_dcolor, const float* dL_ddepth,
                       const float* dL_dalpha, float* dL_dmeans2D, float* dL_dcolors,
                       float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D, float* dL_dsh,
                       float* dL_dscales, float* dL_drotations, void* stream) {
  Plan p;
  int err;
  if (!table || !caps_ws(caps, ws, p, err)) return table ? err : W2W_ERR_INVALID;
  if (!dL_dcolor) return W2W_ERR_INVALID;
  if (caps->items > 0 && (!means || !scales || !rotations || !sh || !dL_dmeans2D || !dL_dcolors ||
                          !dL_dopacity || !dL_dmeans3D || !dL_dcov3D || !dL_dsh || !dL_dscales ||
                          !dL_drotations))
    return W2W_ERR_INVALID;
  (void)hipGetLastError();
  Params prm{means, scales, rotations, nullptr, sh};
  return run_backward(p, table, ws, prm, dL_dcolor, dL_ddepth, dL_dalpha, dL_dmeans2D, dL_dcolors,
                      dL_dopacity, dL_dmeans3D, dL_dcov3D, dL_dsh, dL_dscales, dL_drotations,
                      (hipStream_t)stream);
}

int w2w_backward_batch_sh_adam(const void* table, const W2WCaps* caps, const W2WWorkspace* ws,
                               const float* means, const float* scales, const float* rotations, float* sh,
                               const float* dL_dcolor, const float* dL_ddepth, const float* dL_dalpha,
                               float* dL_dmeans2D, float* dL_dcolors, float* dL_dopacity, float* dL_dmeans3D,
                               float* dL_dcov3D, float* sh_exp_avg, float* sh_exp_avg_sq,
                               float* dL_dscales, float* dL_drotations, const float* hyper,
                               int32_t hyper_dc, int32_t hyper_ac, const uint32_t* skip_if, void* stream) {
  Plan p;
  int err;
  if (!table || !caps_ws(caps, ws, p, err)) return table ? err : W2W_ERR_INVALID;
  if (!dL_dcolor || !hyper || hyper_dc < 0 || hyper_dc >= 16 || hyper_ac < 0 || hyper_ac >= 16)
    return W2W_ERR_INVALID;
  if (caps->items > 0 && (!means || !scales || !rotations || !sh || !dL_dmeans2D || !dL_dcolors ||
                          !dL_dopacity || !dL_dmeans3D || !dL_dcov3D || !sh_exp_avg || !sh_exp_avg_sq ||
                          !dL_dscales || !dL_drotations))
    return W2W_ERR_INVALID;
  (void)hipGetLastError();
  Params prm{means, scales, rotations, nullptr, sh};
  const ShAdam sa{sh, sh_exp_avg, sh_exp_avg_sq, hyper, hyper_dc, hyper_ac, skip_if};
  return run_backward(p, table, ws, prm, dL_dcolor, dL_ddepth, dL_dalpha, dL_dmeans2D, dL_dcolors,
                      dL_dopacity, dL_dmeans3D, dL_dcov3D, nullptr, dL_dscales, dL_drotations,
                      (hipStream_t)stream, &sa);
}

int w2w_forward(const W2WFrame* f, const W2WWorkspace* ws, float* out_color, float* out_depth,
                float* out_alpha, int32_t* out_radii, void* stream) {
  if (!frame_ok(f) || !out_color || !out_depth || !out_alpha || (f->P > 0 && !out_radii))
    return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  Plan p;
  int st = single_setup(f, ws, p, s);
  if (st != W2W_OK) return st;
  Params prm{f->means, f->scales, f->rotations, f->opacities, f->sh};
  return run_forward(p, at<void>(ws->geom, p.table), ws, prm, out_color, out_depth, out_alpha,
                     out_radii, s);
}

int w2w_backward(const W2WFrame* f, const W2WWorkspace* ws, const float* dL_dcolor,
                 const float* dL_ddepth, const float* dL_dalpha, float* dL_dmeans2D,
                 float* dL_dcolors, float* dL_dopacity, float* dL_dmeans3D, float* dL_dcov3D,
                 float* dL_dsh, float* dL_dscales, float* dL_drotations, void* stream) {
  if (!frame_ok(f) || !dL_dcolor) return W2W_ERR_INVALID;
  if (f->P > 0 && (!dL_dmeans2D || !dL_dcolors || !dL_dopacity || !dL_dmeans3D || !dL_dcov3D ||
                   !dL_dsh || !dL_dscales || !dL_drotations))
    return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  Plan p;
  W2WCaps c;
  if (!ws || !single_caps(f->P, f->W, f->H, f->tile, ws->capacity, c) || !make_plan(c, p))
    return W2W_ERR_INVALID;
  if (!ws_ok(ws, p)) return W2W_ERR_CAPACITY;
  if (f->P == 0) return W2W_OK;
  (void)hipGetLastError();
  // the table written by this view's forward is still in the workspace
  Params prm{f->means, f->scales, f->rotations, nullptr, f->sh};
  return run_backward(p, at<void>(ws->geom, p.table), ws, prm, dL_dcolor, dL_ddepth, dL_dalpha,
                      dL_dmeans2D, dL_dcolors, dL_dopacity, dL_dmeans3D, dL_dcov3D, dL_dsh,
                      dL_dscales, dL_drotations, s);
}

extern "C++" {
template <typename KT>
__global__ void k_sorted_keys(const KT* __restrict__ tiles, const uint32_t* __restrict__ ids,
                              const SplatRec* __restrict__ splats, const uint32_t* __restrict__ status,
                              uint32_t cap, uint32_t P, uint64_t* __restrict__ out) {
  uint32_t n = status[ST_K];
  n = n < cap ? n : cap;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t id = ids[i];
    float z = id < P ? splats[id].s.b.z : 0.0f;
    out[i] = ((uint64_t)tiles[i] << 32) | (uint64_t)__float_as_uint(z);
  }
}
}  // extern "C++"

int w2w_sorted_keys(const W2WFrame* f, const W2WWorkspace* ws, uint64_t* out_keys, void* stream) {
  if (!frame_ok(f) || !out_keys) return W2W_ERR_INVALID;
  Plan p;
  W2WCaps c;
  if (!ws || !single_caps(f->P, f->W, f->H, f->tile, ws->capacity, c) || !make_plan(c, p))
    return W2W_ERR_INVALID;
  if (!ws_ok(ws, p)) return W2W_ERR_CAPACITY;
  if (c.capacity == 0) return W2W_OK;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  const uint32_t grid = std::min<uint32_t>(cdiv((uint64_t)c.capacity, 256), 4096);
  if (p.key16)
    hipLaunchKernelGGL(k_sorted_keys<uint16_t>, dim3(grid), dim3(256), 0, s,
                       at<uint16_t>(ws->binning, p.L.sorted_tiles),
                       at<uint32_t>(ws->binning, p.L.sorted_ids), at<SplatRec>(ws->geom, p.L.splats),
                       at<uint32_t>(ws->geom, p.status), (uint32_t)c.capacity, (uint32_t)f->P, out_keys);
  else
    hipLaunchKernelGGL(k_sorted_keys<uint32_t>, dim3(grid), dim3(256), 0, s,
                       at<uint32_t>(ws->binning, p.L.sorted_tiles),
                       at<uint32_t>(ws->binning, p.L.sorted_ids), at<SplatRec>(ws->geom, p.L.splats),
                       at<uint32_t>(ws->geom, p.status), (uint32_t)c.capacity, (uint32_t)f->P, out_keys);
  return launch_status();
}

size_t w2w_sort_temp_bytes(int64_t n, int32_t end_bit) {
  if (n < 0 || end_bit < 1 || end_bit > 64) return 0;
  const int npasses = (end_bit + 7) / 8;
  const size_t tiles = cdiv((uint64_t)n, SORT_TILE);
  size_t o = 0;
  o = align_up(o + 8 * 4);                                  // tickets
  o = align_up(o + 4 * (size_t)npasses * RADIX);            // histogram
  o = align_up(o + 4 * (size_t)npasses * tiles * RADIX);    // look-back words
  o = align_up(o + 8 * (size_t)n);                          // key ping-pong
  o = align_up(o + 4 * (size_t)n);                          // value ping-pong
  return o;
}

int w2w_sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                       uint32_t* vals_out, int64_t n, int32_t end_bit, void* temp,
                       size_t temp_bytes, void* stream) {
  if (n < 0 || n >= (int64_t)LB_VAL || end_bit < 1 || end_bit > 64) return W2W_ERR_INVALID;
  if (n == 0) return W2W_OK;
  if (!keys_in || !vals_in || !keys_out || !vals_out || !temp) return W2W_ERR_INVALID;
  if (temp_bytes < w2w_sort_temp_bytes(n, end_bit)) return W2W_ERR_CAPACITY;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  const int npasses = (end_bit + 7) / 8;
  const int last_bits = end_bit - 8 * (npasses - 1);
  const size_t tiles = cdiv((uint64_t)n, SORT_TILE);
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t a = o; o = align_up(o + bytes); return a; };
  uint32_t* tickets = at<uint32_t>(temp, take(8 * 4));
  uint32_t* hist = at<uint32_t>(temp, take(4 * (size_t)npasses * RADIX));
  uint32_t* status = at<uint32_t>(temp, take(4 * (size_t)npasses * tiles * RADIX));
  uint64_t* kB = at<uint64_t>(temp, take(8 * (size_t)n));
  uint32_t* vB = at<uint32_t>(temp, take(4 * (size_t)n));
  fill_u32(hist, 0u, (size_t)npasses * RADIX, s);
  sort_passes<uint64_t>(keys_in, vals_in, keys_out, vals_out, kB, vB, nullptr, (uint32_t)n, npasses,
                        last_bits, true, hist, status, tickets, W2W_KID_SORT_HIST, W2W_KID_SORT_PASS, s);
  return launch_status();
}

size_t w2w_scan_temp_bytes(int64_t n) {
  if (n < 0) return 0;
  return align_up(4 * ((size_t)cdiv((uint64_t)n, SCAN_TILE) + 2));
}

// The scan of n > 0 values with its scratch at t (tiles + 2 words: ticket,
// spare, one look-back word per 16384-value tile), cleared here by a fill
// kernel (no memset node when the caller captures the stream).
static void scan_launch(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n, uint32_t* t,
                        hipStream_t s) {
  const uint32_t tiles = cdiv((uint64_t)n, SCAN_TILE);
  fill_u32(t, 0u, (size_t)tiles + 2, s);
  ProfSpan ps(W2W_KID_SCAN, s);
  hipLaunchKernelGGL(k_scan<false>, dim3(std::min<uint32_t>(tiles, 2048)), dim3(SCAN_THREADS), 0, s,
                     in, nullptr, out, t + 2, t, total, nullptr, nullptr, (uint32_t)n, 0xffffffffu);
}

int w2w_exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n,
                           void* temp, size_t temp_bytes, void* stream) {
  if (n < 0 || n >= (int64_t)LB_VAL || !total) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  if (n == 0) {
    fill_u32(total, 0u, 1, s);
    return launch_status();
  }
  if (!in || !out || !temp) return W2W_ERR_INVALID;
  if (temp_bytes < w2w_scan_temp_bytes(n)) return W2W_ERR_CAPACITY;
  scan_launch(in, out, total, n, static_cast<uint32_t*>(temp), s);
  return launch_status();
}

// ---- the scan for callers that cannot allocate ----
// Its scratch is `scratch`, a file-scope device array: each GPU the library
// runs on holds its own copy until the process ends.  In host code the name
// is only a stand-in for the symbol, not a device address, so it is never
// passed to a kernel or a copy: the first call on a device takes the array's
// address there with hipGetSymbolAddress (that device current) and keeps it
// per device; every later use goes through that pointer.
constexpr size_t SCAN_SCRATCH_WORDS = (size_t)(W2W_SCAN_NOALLOC_MAX_N / SCAN_TILE) + 2;
static_assert(W2W_SCAN_NOALLOC_MAX_N % SCAN_TILE == 0, "scratch sized in whole scan tiles");
static_assert(W2W_SCAN_NOALLOC_MAX_N < LB_VAL, "within the scan's 30-bit look-back values");
static __device__ __align__(256) uint32_t scratch[SCAN_SCRATCH_WORDS];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kScratchDevices] = {};

int scratch_of_current_device(uint32_t** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return W2W_ERR_HIP;
  if (dev < 0 || dev >= kScratchDevices) return W2W_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return W2W_ERR_HIP;
    g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
  }
  *out = g_scratch_ptr[dev];
  return W2W_OK;
}
}  // namespace

int w2w_exclusive_scan_u32_noalloc(const uint32_t* in, uint32_t* out, uint32_t* total, int64_t n,
                                   void* stream) {
  if (n < 0 || !total) return W2W_ERR_INVALID;
  if (n > (int64_t)W2W_SCAN_NOALLOC_MAX_N) return W2W_ERR_CAPACITY;
  if (n > 0 && (!in || !out)) return W2W_ERR_INVALID;
  uint32_t* t = nullptr;
  const int err = scratch_of_current_device(&t);
  if (err != W2W_OK) return err;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  if (n == 0) {
    fill_u32(total, 0u, 1, s);
    return launch_status();
  }
  scan_launch(in, out, total, n, t, s);
  return launch_status();
}

int w2w_prof_enable(int32_t capacity, uint32_t kernel_mask) {
  for (int i = 0; i < g_prof.capacity; i++) {
    (void)hipEventDestroy(g_prof.start[i]);
    (void)hipEventDestroy(g_prof.stop[i]);
  }
  delete[] g_prof.start;
  delete[] g_prof.stop;
  delete[] g_prof.kid;
  g_prof = Prof();
  if (capacity <= 0) return W2W_OK;
  g_prof.start = new hipEvent_t[capacity];
  g_prof.stop = new hipEvent_t[capacity];
  g_prof.kid = new int[capacity];
  for (int i = 0; i < capacity; i++) {
    if (hipEventCreate(&g_prof.start[i]) != hipSuccess || hipEventCreate(&g_prof.stop[i]) != hipSuccess) {
      g_prof.capacity = i;
      w2w_prof_enable(0, 0);
      return W2W_ERR_HIP;
    }
  }
  g_prof.capacity = capacity;
  g_prof.mask = kernel_mask;
  return W2W_OK;
}

int w2w_adam_step(float* params, const float* grads, float* exp_avg, float* exp_avg_sq, int64_t n,
                  int32_t n_groups, const int64_t* group_end, const float* group_lr, int64_t step,
                  float beta1, float beta2, float eps, void* stream) {
  if (n < 0 || n_groups < 1 || n_groups > ADAM_MAX_SEGMENTS || step < 1 || !group_end || !group_lr)
    return W2W_ERR_INVALID;
  if (n == 0) return W2W_OK;
  if (!params || !grads || !exp_avg || !exp_avg_sq) return W2W_ERR_INVALID;
  AdamArgs a;
  a.p = params; a.g = grads; a.m = exp_avg; a.v = exp_avg_sq; a.n = (uint64_t)n;
  a.nseg = n_groups;
  const double bc1 = 1.0 - std::pow((double)beta1, (double)step);
  const double bc2 = 1.0 - std::pow((double)beta2, (double)step);
  int64_t prev = 0;
  for (int k = 0; k < ADAM_MAX_SEGMENTS; k++) {
    if (k < n_groups) {
      if (group_end[k] < prev || group_end[k] > n) return W2W_ERR_INVALID;
      prev = group_end[k];
      a.seg_end[k] = (uint64_t)group_end[k];
      a.neg_step[k] = (float)(-(double)group_lr[k] / bc1);
    } else {
      a.seg_end[k] = (uint64_t)n;
      a.neg_step[k] = 0.0f;
    }
  }
  if (group_end[n_groups - 1] != n) return W2W_ERR_INVALID;
  a.inv_bc2_sqrt = (float)(1.0 / std::sqrt(bc2));
  a.one_m_b1 = (float)(1.0 - (double)beta1);
  a.b2 = beta2;
  a.one_m_b2 = (float)(1.0 - (double)beta2);
  a.eps = eps;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  const uint32_t grid = (uint32_t)std::min<uint64_t>(((uint64_t)n + 255) / 256, 8192);
  ProfSpan ps(W2W_KID_ADAM, s);
  hipLaunchKernelGGL(k_adam, dim3(grid), dim3(256), 0, s, a);
  return launch_status();
}

int w2w_activate(const float* svec, const float* logit, const float* quat, int64_t G, float* scales,
                 float* opac, float* rots, void* stream) {
  if (G < 0 || G >= (int64_t)LB_VAL) return W2W_ERR_INVALID;
  if (G == 0) return W2W_OK;
  if (!svec || !logit || !quat || !scales || !opac || !rots) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ProfSpan ps(W2W_KID_ACTIVATE, s);
  hipLaunchKernelGGL(k_activate, dim3(std::min<uint32_t>(cdiv((uint64_t)G, 256), 8192)), dim3(256), 0,
                     s, svec, logit, quat, (uint64_t)G, (const RowRange*)nullptr, 0u, scales, opac,
                     rots);
  return launch_status();
}

int w2w_activate_backward(const float* scales, const float* opac, const float* quat,
                          const float* g_scales, const float* g_opac, const float* g_rots,
                          int64_t G, float* g_svec, float* g_logit, float* g_quat, void* stream) {
  if (G < 0 || G >= (int64_t)LB_VAL) return W2W_ERR_INVALID;
  if (G == 0) return W2W_OK;
  if (!scales || !opac || !quat || !g_scales || !g_opac || !g_rots || !g_svec || !g_logit || !g_quat)
    return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ProfSpan ps(W2W_KID_ACTIVATE_BWD, s);
  hipLaunchKernelGGL(k_activate_bwd, dim3(std::min<uint32_t>(cdiv((uint64_t)G, 256), 8192)), dim3(256),
                     0, s, scales, opac, quat, g_scales, g_opac, g_rots, (uint64_t)G,
                     (const RowRange*)nullptr, 0u, g_svec, g_logit, g_quat);
  return launch_status();
}

int w2w_l2_loss(const void* table, const W2WCaps* caps, const float* color, const float* target,
                float* dL_dcolor, float* loss_per_frame, float* tile_scratch, void* stream) {
  if (!table || !caps || !color || !target || !dL_dcolor || !loss_per_frame || !tile_scratch ||
      caps->frames < 1 || caps->scenes < 1 || caps->pixels < 1 || caps->tiles < 1 ||
      (caps->tile != 8 && caps->tile != 16 && caps->tile != 32))
    return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ProfSpan ps(W2W_KID_LOSS, s);
  hipLaunchKernelGGL(k_l2_loss_tiles, dim3(std::min<uint32_t>(cdiv((uint64_t)caps->tiles, 4), 8192)), dim3(256), 0,
                     s, table, (uint32_t)caps->scenes, (uint32_t)caps->tiles, (int)caps->tile, color,
                     target, dL_dcolor, tile_scratch);
  hipLaunchKernelGGL(k_loss_frames, dim3(std::min<uint32_t>((uint32_t)caps->frames, 1024)),
                     dim3(256), 0, s, table, (uint32_t)caps->scenes, (uint32_t)caps->frames,
                     tile_scratch, loss_per_frame);
  return launch_status();
}

int w2w_adam_step_hyper(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                        int64_t n, int32_t n_groups, const int64_t* group_end,
                        const int32_t* group_period, const int32_t* group_split,
                        const float* hyper, const uint32_t* skip_if, uint32_t* sticky,
                        void* stream) {
  if (n < 0 || n_groups < 1 || n_groups > ADAM_MAX_SEGMENTS || !group_end || !hyper)
    return W2W_ERR_INVALID;
  if (n == 0) return W2W_OK;
  if (!params || !grads || !exp_avg || !exp_avg_sq) return W2W_ERR_INVALID;
  AdamGraphArgs a;
  a.p = params; a.g = grads; a.m = exp_avg; a.v = exp_avg_sq; a.n = (uint64_t)n;
  a.nseg = n_groups; a.hyper = hyper; a.skip_if = skip_if; a.sticky = sticky;
  int64_t prev = 0;
  for (int k = 0; k < ADAM_MAX_SEGMENTS; k++) {
    if (k < n_groups) {
      if (group_end[k] < prev || group_end[k] > n) return W2W_ERR_INVALID;
      prev = group_end[k];
      a.seg_end[k] = (uint64_t)group_end[k];
      const int32_t per = group_period ? group_period[k] : 0;
      const int32_t spl = group_split ? group_split[k] : 0;
      if (per > 0 && spl >= 0 && spl < per) {
        a.period[k] = (uint32_t)per; a.split[k] = (uint32_t)spl; a.lr2[k] = 8 + k;
      } else {
        a.period[k] = 1; a.split[k] = 1; a.lr2[k] = -1;
      }
    } else {
      a.seg_end[k] = (uint64_t)n; a.period[k] = 1; a.split[k] = 1; a.lr2[k] = -1;
    }
  }
  if (group_end[n_groups - 1] != n) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  const bool vec = ((reinterpret_cast<uintptr_t>(params) | reinterpret_cast<uintptr_t>(grads) |
                     reinterpret_cast<uintptr_t>(exp_avg) | reinterpret_cast<uintptr_t>(exp_avg_sq)) & 15u) == 0;
  const uint64_t work = vec ? ((uint64_t)n + 3) / 4 : (uint64_t)n;
  const uint32_t grid = (uint32_t)std::min<uint64_t>((work + 255) / 256, 8192);
  ProfSpan ps(W2W_KID_ADAM, s);
  if (vec)
    hipLaunchKernelGGL(k_adam_hyper<true>, dim3(grid), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL(k_adam_hyper<false>, dim3(grid), dim3(256), 0, s, a);
  return launch_status();
}

int w2w_activate_rows(const float* svec, const float* logit, const float* quat, const void* ranges,
                      int32_t n_ranges, int64_t total_rows, float* scales, float* opac, float* rots,
                      void* stream) {
  if (n_ranges < 1 || total_rows < 0 || !ranges) return W2W_ERR_INVALID;
  if (total_rows == 0) return W2W_OK;
  if (!svec || !logit || !quat || !scales || !opac || !rots) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ProfSpan ps(W2W_KID_ACTIVATE, s);
  hipLaunchKernelGGL(k_activate, dim3(std::min<uint32_t>(cdiv((uint64_t)total_rows, 256), 8192)),
                     dim3(256), 0, s, svec, logit, quat, (uint64_t)total_rows,
                     static_cast<const RowRange*>(ranges), (uint32_t)n_ranges, scales, opac, rots);
  return launch_status();
}

int w2w_activate_backward_rows(const float* scales, const float* opac, const float* quat,
                               const float* g_scales, const float* g_opac, const float* g_rots,
                               const void* ranges, int32_t n_ranges, int64_t total_rows,
                               float* g_svec, float* g_logit, float* g_quat, void* stream) {
  if (n_ranges < 1 || total_rows < 0 || !ranges) return W2W_ERR_INVALID;
  if (total_rows == 0) return W2W_OK;
  if (!scales || !opac || !quat || !g_scales || !g_opac || !g_rots || !g_svec || !g_logit || !g_quat)
    return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ProfSpan ps(W2W_KID_ACTIVATE_BWD, s);
  hipLaunchKernelGGL(k_activate_bwd, dim3(std::min<uint32_t>(cdiv((uint64_t)total_rows, 256), 8192)),
                     dim3(256), 0, s, scales, opac, quat, g_scales, g_opac, g_rots,
                     (uint64_t)total_rows, static_cast<const RowRange*>(ranges), (uint32_t)n_ranges,
                     g_svec, g_logit, g_quat);
  return launch_status();
}

int w2w_adam_ranges(float* params, const float* grads, float* exp_avg, float* exp_avg_sq,
                    const void* ranges, int32_t n_ranges, int64_t total, const float* hyper,
                    float beta1, float beta2, float eps, void* stream) {
  if (n_ranges < 1 || total < 0 || !ranges || !hyper) return W2W_ERR_INVALID;
  if (total == 0) return W2W_OK;
  if (!params || !grads || !exp_avg || !exp_avg_sq) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ProfSpan ps(W2W_KID_ADAM, s);
  hipLaunchKernelGGL(k_adam_ranges, dim3(std::min<uint32_t>(cdiv((uint64_t)total, 1024), 8192)),
                     dim3(256), 0, s, params, grads, exp_avg, exp_avg_sq,
                     static_cast<const ParamRange*>(ranges), (uint32_t)n_ranges, (uint64_t)total,
                     hyper, (float)(1.0 - (double)beta1), beta2, (float)(1.0 - (double)beta2), eps);
  return launch_status();
}

size_t w2w_densify_temp_bytes(int64_t n) {
  if (n < 0) return 0;
  const uint64_t nb = cdiv((uint64_t)n, DENS_BLOCK);
  return align_up((size_t)n) + align_up((size_t)nb * 8 * sizeof(uint32_t)) + kAlign;
}

namespace {
struct DensTemp {
  uint8_t* flags;
  uint32_t* blockcnt;
  uint32_t* totals;
};
DensTemp dens_temp(void* temp, int64_t n) {
  uint8_t* b = static_cast<uint8_t*>(temp);
  const uint64_t nb = cdiv((uint64_t)n, DENS_BLOCK);
  DensTemp t;
  t.flags = b;
  t.blockcnt = reinterpret_cast<uint32_t*>(b + align_up((size_t)n));
  t.totals = reinterpret_cast<uint32_t*>(b + align_up((size_t)n) + align_up((size_t)nb * 8 * 4));
  return t;
}
bool dens_cfg(const W2WDensifyConfig* c, int64_t n, DensifyCfg* out) {
  if (!c || n < 0 || c->split_factor < 1 || c->split_factor > 8) return false;
  // output rows are indexed in 32 bits
  if ((uint64_t)n * (uint64_t)(2 + c->split_factor) >= 0xFFFFFFFFull) return false;
  out->grad_thresh = c->grad_thresh;
  out->min_opacity = c->min_opacity;
  out->clone_scale_max = c->clone_scale_max;
  out->split_scale_min = c->split_scale_min;
  out->prune_scale_max = c->prune_scale_max;
  out->split_div = (float)((double)c->split_factor * (double)c->split_shrink);
  out->factor = (uint32_t)c->split_factor;
  return out->split_div > 0.0f;
}
// group table -> kernel store; moments optional (plan reads parameters only)
bool dens_store(const W2WStoreGroups* g, bool moments, DensStore* s) {
  if (!g || g->n_groups < 1 || g->n_groups > DENS_MAX_GROUPS) return false;
  const int roles[4] = {g->means, g->opacity, g->svec, g->quat};
  const int widths[4] = {3, 1, 3, 4};
  for (int r = 0; r < 4; r++)
    if (roles[r] < 0 || roles[r] >= g->n_groups || g->width[roles[r]] != widths[r]) return false;
  for (int k = 0; k < DENS_MAX_GROUPS; k++) {
    const bool on = k < g->n_groups;
    if (on && (g->width[k] < 0 || g->width[k] > 1024)) return false;
    if (on && g->width[k] > 0 && (!g->param[k] || (moments && (!g->exp_avg[k] || !g->exp_avg_sq[k]))))
      return false;
    s->p[k] = on ? g->param[k] : nullptr;
    s->m[k] = on ? g->exp_avg[k] : nullptr;
    s->v[k] = on ? g->exp_avg_sq[k] : nullptr;
    s->w[k] = on ? (uint32_t)g->width[k] : 0u;
  }
  s->ng = g->n_groups;
  s->means = g->means;
  s->opacity = g->opacity;
  s->svec = g->svec;
  s->quat = g->quat;
  return true;
}
}  // namespace

int w2w_densify_plan(const W2WStoreGroups* store, int64_t n, const float* grad_accum,
                     const float* grad_count, const W2WDensifyConfig* cfg, void* temp,
                     size_t temp_bytes, int64_t* counts, void* stream) {
  DensifyCfg c;
  DensStore st;
  if (!dens_cfg(cfg, n, &c) || !temp || !counts) return W2W_ERR_INVALID;
  if (temp_bytes < w2w_densify_temp_bytes(n)) return W2W_ERR_CAPACITY;
  if (n > 0 && (!dens_store(store, false, &st) || !grad_accum || !grad_count)) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  const DensTemp t = dens_temp(temp, n);
  const uint32_t nb = cdiv((uint64_t)n, DENS_BLOCK);
  ProfSpan ps(W2W_KID_DENSIFY, s);
  if (nb) {
    DensFlagsArgs a{st, (uint64_t)n, grad_accum, grad_count, c, t.flags, t.blockcnt};
    hipLaunchKernelGGL(k_dens_flags, dim3(nb), dim3(DENS_BLOCK), 0, s, a);
  }
  hipLaunchKernelGGL(k_dens_scan, dim3(1), dim3(512), 0, s, t.blockcnt, nb, t.totals,
                     reinterpret_cast<long long*>(counts), (uint64_t)n, c.factor);
  return launch_status();
}

int w2w_densify_apply(const W2WStoreGroups* store, int64_t n, const W2WDensifyConfig* cfg,
                      const void* temp, const float* normals, int64_t normal_rows,
                      const W2WStoreGroups* out, int64_t n_out, void* stream) {
  DensifyCfg c;
  DensStore si, so;
  if (!dens_cfg(cfg, n, &c) || !temp || normal_rows < 0 || n_out < 0) return W2W_ERR_INVALID;
  if (n == 0 || n_out == 0) return W2W_OK;
  if (!dens_store(store, true, &si) || !dens_store(out, true, &so)) return W2W_ERR_INVALID;
  if (si.ng != so.ng || si.means != so.means || si.opacity != so.opacity || si.svec != so.svec ||
      si.quat != so.quat)
    return W2W_ERR_INVALID;
  for (int k = 0; k < si.ng; k++)
    if (si.w[k] != so.w[k]) return W2W_ERR_INVALID;
  if (normal_rows > 0 && !normals) return W2W_ERR_INVALID;
  if ((uint64_t)n_out > (uint64_t)n * (uint64_t)(1 + c.factor)) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  const DensTemp t = dens_temp(const_cast<void*>(temp), n);
  DensApplyArgs a{si, (uint64_t)n, t.flags, t.blockcnt, t.totals, normals, (uint64_t)normal_rows, c,
                  so, (uint64_t)n_out};
  ProfSpan ps(W2W_KID_DENSIFY, s);
  hipLaunchKernelGGL(k_dens_apply, dim3(cdiv((uint64_t)n, DENS_BLOCK)), dim3(DENS_BLOCK), 0, s, a);
  return launch_status();
}

int w2w_densify_stats(const void* table, const W2WCaps* caps, const float* means2D_grad,
                      const int32_t* radii, float* accum, float* count, void* stream) {
  if (!table || !caps || caps->scenes < 1 || caps->gaussians < 0 || caps->gaussians >= (int64_t)LB_VAL)
    return W2W_ERR_INVALID;
  if (caps->gaussians == 0) return W2W_OK;
  if (!means2D_grad || !radii || !accum || !count) return W2W_ERR_INVALID;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  ProfSpan ps(W2W_KID_DENSIFY, s);
  hipLaunchKernelGGL(k_dens_stats, dim3(std::min<uint32_t>(cdiv((uint64_t)caps->gaussians, 256), 8192)),
                     dim3(256), 0, s, table, (uint32_t)caps->scenes, means2D_grad, radii, accum, count);
  return launch_status();
}

namespace {
struct VoxTemp {
  uint32_t *keys0, *vals0, *kA, *vA, *kB, *vB, *hist, *status, *tickets;
  uint2* cells;
  int *lo, *hi;
  int npasses, last_bits;
  size_t bytes;
};
VoxTemp vox_temp(void* temp, int64_t n, int32_t nb) {
  const uint64_t nc = (uint64_t)nb + 1, ncell = nc * nc * nc;
  int bits = 1;
  while ((1ull << bits) < ncell) bits++;
  VoxTemp t;
  t.npasses = (bits + 7) / 8;
  t.last_bits = bits - 8 * (t.npasses - 1);
  const size_t tiles = std::max<size_t>(cdiv((uint64_t)n, SORT_TILE), 1);
  size_t o = 0;
  auto take = [&](size_t b) { size_t a = o; o = align_up(o + b); return a; };
  const size_t k0 = take(4 * (size_t)n), v0 = take(4 * (size_t)n), ka = take(4 * (size_t)n);
  const size_t va = take(4 * (size_t)n), kb = take(4 * (size_t)n), vb = take(4 * (size_t)n);
  const size_t hi_ = take(4 * 4 * RADIX), st = take(4 * (size_t)t.npasses * tiles * RADIX);
  const size_t tk = take(8 * 4), ce = take(8 * ncell), lo = take(4 * nc), hh = take(4 * nc);
  t.bytes = o;
  if (!temp) return t;  // sizing only
  t.keys0 = at<uint32_t>(temp, k0); t.vals0 = at<uint32_t>(temp, v0);
  t.kA = at<uint32_t>(temp, ka); t.vA = at<uint32_t>(temp, va);
  t.kB = at<uint32_t>(temp, kb); t.vB = at<uint32_t>(temp, vb);
  t.hist = at<uint32_t>(temp, hi_); t.status = at<uint32_t>(temp, st); t.tickets = at<uint32_t>(temp, tk);
  t.cells = at<uint2>(temp, ce); t.lo = at<int>(temp, lo); t.hi = at<int>(temp, hh);
  return t;
}
bool vox_ok(int64_t n, int32_t res, int32_t nb, float relax) {
  return n >= 0 && n < (int64_t)LB_VAL && res >= 2 && res <= 4096 && nb >= 1 && nb <= 255 &&
         std::isfinite(relax) && relax >= 0.0f && relax <= 64.0f;
}
}  // namespace

size_t w2w_voxel_temp_bytes(int64_t n, int32_t num_blocks) {
  if (n < 0 || num_blocks < 1 || num_blocks > 255) return 0;
  return vox_temp(nullptr, n, num_blocks).bytes;
}

int w2w_voxel_field(const float* means, const float* covs, const float* opacities, int64_t n,
                    int32_t resolution, int32_t num_blocks, float relax_ratio, float* occ, void* temp,
                    size_t temp_bytes, void* stream) {
  if (!vox_ok(n, resolution, num_blocks, relax_ratio) || !occ || !temp) return W2W_ERR_INVALID;
  if (n > 0 && (!means || !covs || !opacities)) return W2W_ERR_INVALID;
  if (temp_bytes < w2w_voxel_temp_bytes(n, num_blocks)) return W2W_ERR_CAPACITY;
  hipStream_t s = (hipStream_t)stream;
  (void)hipGetLastError();
  const VoxTemp t = vox_temp(temp, n, num_blocks);
  VoxelGeom g;
  g.res = resolution; g.nb = num_blocks;
  g.block_size = 2.0f / (float)num_blocks;  // compute_voxel_field.cu:37
  g.relax = relax_ratio;
  const uint32_t nc = (uint32_t)num_blocks + 1, ncell = nc * nc * nc;
  fill_u32(reinterpret_cast<uint32_t*>(t.cells), 0u, 2 * (size_t)ncell, s);
  fill_u32(reinterpret_cast<uint32_t*>(t.lo), (uint32_t)resolution, nc, s);
  fill_u32(reinterpret_cast<uint32_t*>(t.hi), 0u, nc, s);
  ProfSpan ps(W2W_KID_VOXEL, s);
  if (n > 0) {
    const uint32_t grid = std::min<uint32_t>(cdiv((uint64_t)n, 256), 4096);
    hipLaunchKernelGGL(k_vox_cells, dim3(grid), dim3(256), 0, s, means, (uint32_t)n, g, t.keys0, t.vals0);
    fill_u32(t.hist, 0u, 4 * RADIX, s);
    sort_passes<uint32_t>(t.keys0, t.vals0, t.kA, t.vA, t.kB, t.vB, nullptr, (uint32_t)n, t.npasses,
                          t.last_bits, true, t.hist, t.status, t.tickets, W2W_KID_VOXEL, W2W_KID_VOXEL, s);
    hipLaunchKernelGGL(k_vox_cell_ranges, dim3(grid), dim3(256), 0, s, t.kA, (uint32_t)n, t.cells);
  }
  hipLaunchKernelGGL(k_vox_axis, dim3(cdiv((uint64_t)resolution, 256)), dim3(256), 0, s, g, t.lo, t.hi);
  VoxArgs va;
  va.means = means; va.covs = covs; va.opac = opacities; va.cell_ids = t.vA; va.cell_ranges = t.cells;
  va.axis_lo = t.lo; va.axis_hi = t.hi; va.g = g;
  va.reach = (int)std::ceil(relax_ratio) + 2;  // every relaxed box lies in these cells (one cell of slack)
  va.occ = occ;
  hipLaunchKernelGGL(k_vox_field, dim3(ncell), dim3(VOX_THREADS), 0, s, va);
  return launch_status();
}

int w2w_stats_enable(uint64_t* device_counters) {
  g_stats = reinterpret_cast<unsigned long long*>(device_counters);
  return W2W_OK;
}

int w2w_prof_reset(void) {
  g_prof.used = 0;
  return W2W_OK;
}

int w2w_prof_read(int32_t* kernel_ids, float* ms, int32_t max_records, int32_t* n_records) {
  if (!n_records) return W2W_ERR_INVALID;
  const int n = std::min<int>(g_prof.used, max_records);
  for (int i = 0; i < n; i++) {
    if (hipEventSynchronize(g_prof.stop[i]) != hipSuccess) return W2W_ERR_HIP;
    float t = 0.0f;
    if (hipEventElapsedTime(&t, g_prof.start[i], g_prof.stop[i]) != hipSuccess) return W2W_ERR_HIP;
    if (kernel_ids) kernel_ids[i] = g_prof.kid[i];
    if (ms) ms[i] = t;
  }
  *n_records = n;
  return W2W_OK;
}

}  // extern "C"
