// Backward rasterizer kernels for gfx950.
//
//   k_blend_bwd_tile  per-pixel reverse traversal (the role of backward.cu:521-665)
//                     for tile edges 8 and 16: starts from the stored final T and
//                     last-contributor index, recomputes alpha with the forward's
//                     exact arithmetic and T <- T / (1 - alpha); no
//                     per-(pixel, Gaussian) state.  One wave owns a tile (every
//                     lane a pixel of each 8x8 block), tiles are taken heaviest
//                     first, and a splat's moments are summed over the wave and
//                     written once per (tile, splat) instance to that instance's
//                     slot in emission order.  No atomics: deterministic.
//   k_blend_bwd       the same for tile edge 32: a workgroup per tile, one wave
//                     per 8x8 block, the waves' sums added in LDS in a fixed order.
//   k_preprocess_bwd  per-Gaussian chain rule (the role of backward.cu:27-227 and
//                     :229-519), differentiating the forward exactly as defined
//                     in k_preprocess (radial clamp, degree-aware SH with the 0.3
//                     coefficient, NDC depth, no dilation); it first sums each
//                     item's instance records, which lie contiguously from the
//                     item's emission offset, in emission order.
#pragma once
#include "w2w_adam.hpp"
#include "w2w_device.hpp"
#include "w2w_forward.hpp"

namespace w2w {

// Per-instance backward record (floats): sums over the instance's pixels of
// q = G dL/dalpha (zero where the alpha was clamped) weighted by the
// pixel-to-mean offsets (dx, dy) = mean2D - pixel, and the colour and depth
// partials.  Every gradient of the splat's 2D quantities is a splat constant
// times one of these sums (dL/dopacity = S, dL/dconic a = -o Sxx / 2, b = -o
// Sxy, c = -o Syy / 2, dL/dmean2D = -o (a Sx + b Sy, b Sx + c Sy) (W/2,
// H/2)), so the constants are applied once per item after the records are
// summed (k_item_grads) rather than once per pixel.
enum GradField : int {
  GA_SX = 0, GA_SY = 1,                // S q dx, S q dy
  GA_SXX = 2, GA_SXY = 3, GA_SYY = 4,  // S q dx dx, S q dx dy, S q dy dy
  GA_S = 5,                            // S q
  GA_R = 6, GA_G = 7, GA_B = 8,        // dL/drgb
  GA_DEPTH = 9,                        // dL/d(NDC depth)
  // the item-level results k_item_grads forms from the summed moments
  GA_MX = 0, GA_MY = 1, GA_CA = 2, GA_CB = 3, GA_CC = 4, GA_OP = 5,
};

constexpr int GR_FIELDS = 10;


// Sum over the 64 lanes of a wave, result valid in lane 63: quad_perm
// butterflies, row rotations, then the GFX9 row broadcasts.  All lanes must
// be active.
W2W_DEV float wave_reduce_dpp(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x124, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x128, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x142, 0xA, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x143, 0xC, 0xF, false));
  return v;
}

// Sum over each 16-lane row: quad_perm butterflies then row rotations; every
// lane of a row ends with its row's sum.  All lanes must be active.
W2W_DEV float row_sum16(float v) {
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x124, 0xF, 0xF, false));
  v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x128, 0xF, 0xF, false));
  return v;
}

// gfx950 half-wave and row swaps: (x, y) -> sums that pack two fields into
// one register.  v_permlane32_swap exchanges the upper 32 lanes of x with the
// lower 32 of y; v_permlane16_swap exchanges the odd 16-lane rows of x with
// the even rows of y.  Written as inline asm with both operands read and
// written: this compiler's builtins for them return the first register for
// both results.  The s_nop pairs give the VALU -> permlane wait states.
W2W_DEV float fold32(float x, float y) {  // lanes 0-31: x_l + x_l+32; 32-63: y_l-32 + y_l
  asm volatile("s_nop 1\n\tv_permlane32_swap_b32 %0, %1\n\ts_nop 1" : "+v"(x), "+v"(y));
  return x + y;
}
W2W_DEV float fold16(float x, float y) {  // rows (x0+x1, y0+y1, x2+x3, y2+y3)
  asm volatile("s_nop 1\n\tv_permlane16_swap_b32 %0, %1\n\ts_nop 1" : "+v"(x), "+v"(y));
  return x + y;
}

// Sums of four fields over the whole wave: after the folds, row r of the
// result holds field (a, c, b, d)[r] summed over the rows; the row sum then
// leaves the totals in every lane of that row.
W2W_DEV float wave_sum4(float a, float b, float c, float d) {
  return row_sum16(fold16(fold32(a, b), fold32(c, d)));
}
// Two fields: rows (a, 0, b, 0).
W2W_DEV float wave_sum2(float a, float b) {
  return row_sum16(fold16(fold32(a, b), 0.0f));
}

// wave_sum4(a0, b0, c0, d0), wave_sum4(a1, b1, c1, d1) and wave_sum2(a2, b2)
// with the same additions, but the five half-wave swaps issued back to back
// and then the three row swaps: one pair of wait states per stage instead of
// one per swap (the swaps of a stage read and write independent registers).
W2W_DEV void wave_sum10(float a0, float b0, float c0, float d0, float a1, float b1, float c1, float d1, float a2,
                        float b2, float& s0, float& s1, float& s2) {
  asm volatile(
      "s_nop 1\n\t"
      "v_permlane32_swap_b32 %0, %1\n\t"
      "v_permlane32_swap_b32 %2, %3\n\t"
      "v_permlane32_swap_b32 %4, %5\n\t"
      "v_permlane32_swap_b32 %6, %7\n\t"
      "v_permlane32_swap_b32 %8, %9\n\t"
      "s_nop 1"
      : "+v"(a0), "+v"(b0), "+v"(c0), "+v"(d0), "+v"(a1), "+v"(b1), "+v"(c1), "+v"(d1), "+v"(a2), "+v"(b2));
  float x0 = a0 + b0, y0 = c0 + d0, x1 = a1 + b1, y1 = c1 + d1, x2 = a2 + b2, y2 = 0.0f;
  asm volatile(
      "s_nop 1\n\t"
      "v_permlane16_swap_b32 %0, %1\n\t"
      "v_permlane16_swap_b32 %2, %3\n\t"
      "v_permlane16_swap_b32 %4, %5\n\t"
      "s_nop 1"
      : "+v"(x0), "+v"(y0), "+v"(x1), "+v"(y1), "+v"(x2), "+v"(y2));
  s0 = row_sum16(x0 + y0);
  s1 = row_sum16(x1 + y1);
  s2 = row_sum16(x2 + y2);
}

W2W_DEV float lane63(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), 63));
}

struct BlendBwdArgs {
  const void* table;
  uint32_t scenes_cap, tiles_cap, items_cap, capacity;
  uint32_t* ticket;         // zeroed; tiles are claimed in order through it
  const uint2* ranges;
  const uint32_t* ids;
  const SplatRec* splats;   // per item: splat, block-test threshold, rectangle, record offset
  const uint32_t* rec_base; // per PRE_TILE items: first record position (the splat record's offset adds to it)
  unsigned long long* stats;  // optional [3]: evaluated pairs, contributing pairs, loaded instances
  const uint32_t* n_contrib;
  const float* final_T;
  const float* dL_dcolor;   // per frame [3,H,W] at 3*pix_off
  const float* dL_dalpha;   // per frame [H,W] at pix_off; may be null (zero)
  const float* dL_ddepth;   // per frame [H,W] at pix_off; may be null (zero)
  float* rec;               // [capacity][GR_FIELDS] per-instance partials
  uint32_t* flag;           // [capacity / 32] bits, zeroed: bit k set where record k was written
  // depth-sliced binning: a tile's list is its near list followed by its far
  // list (null: one list)
  const uint2* ranges_far;
  const uint32_t* ids_far;
  // the tiles with backward work, heaviest first (k_tile_order), and their count
  const uint32_t* order;
  const uint32_t* n_order;
  // per tile: contributing (pixel, splat) pairs counted by the forward blend
  // (tile edge 16: tiles from kCoopCost up are walked by a whole workgroup)
  const uint32_t* tile_cost;
};

// Backward blend (backward.cu:396-520 semantics, with our forward's
// decisions): every pixel walks its list back to front from its last
// contributor, recovering T by T <- T / (1 - alpha) from the stored final T
// (the reciprocal of 1 - alpha >= 0.01 is taken with the hardware rcp, 1 ulp).
//
// Used for tile edge 32 (k_blend_bwd_tile below covers edges 8 and 16).
// A workgroup claims a tile (heaviest first, through the order of
// k_tile_order); wave w owns 8x8 block w.  The waves walk the
// tile's list together in chunks of 64 (down from the tile's largest last
// contributor), each gathering the chunk one instance per lane, dropping by
// ballot the splats whose cull extent misses its block (every pixel of the
// block would skip them) and visiting the rest in reverse order.  For a splat
// with any contributing pixel the wave sums the 10 partials over its lanes
// (gfx950 half-wave and row swaps pack two and four fields per register, then
// four DPP row steps) and stores the sums in LDS.  After a barrier, lane j of each wave adds all
// slots of splat j in (wave, row) order and writes the sums to the
// instance's record (emission position = item offset + its tile's index in
// the item's rectangle, prefetched with the chunk) and sets its flag; a
// second barrier frees the slots for the next chunk.
//
// AUX: gradients w.r.t. the alpha and depth images are given.  Without them
// (the training step) the alpha/depth terms and their running sums drop out;
// the partials differ from the AUX variant's with zero inputs at most in the
// sign of a zero.
template <int TS, bool AUX>
__global__ __launch_bounds__(64 * units_per_tile<TS>()) __attribute__((amdgpu_waves_per_eu(4, 4)))
void k_blend_bwd(BlendBwdArgs a) {
  constexpr int UPT = (int)units_per_tile<TS>();
  constexpr int UX = TS / kUnit;
  __shared__ float s_part[UPT][GR_FIELDS][64];  // per wave, per field, per splat of the chunk
  // tile edges 8 and 16 stage each wave's chunk in LDS for broadcast reads
  // (16 waves of edge 32 would not fit; they broadcast with readlane)
  constexpr bool kStage = UPT <= 4;
  __shared__ Splat s_stage[kStage ? UPT : 1][64];
  __shared__ uint64_t s_cm[UPT];
  __shared__ uint32_t s_tile, s_max;
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t Tn = min(hdr->n_tiles, a.tiles_cap);
  const uint32_t P = min(hdr->n_items, a.items_cap);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  const int tid = (int)threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // uniform: LDS addresses in scalars
  const int row = lane >> 4;
  const bool row_tail = (lane & 15) == 15;
  // which field each row holds after wave_sum4 / wave_sum2 (and wave_sum10)
  const int fa = (row & 1) ? (row >> 1) + 2 : (row >> 1);  // rows -> fields 0, 2, 1, 3
  const float a_min = 0x1.010102p-8f;
  uint32_t n_eval = 0, n_con = 0, n_loaded = 0;
  const uint32_t n_order = min(*a.n_order, Tn);
  for (;;) {
    if (tid == 0) {
      const uint32_t t = atomicAdd(a.ticket, 1u);
      s_tile = t < n_order ? (a.order[t] & kOrderTile) : 0xffffffffu;
      s_max = 0;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= Tn) break;
    const FrameDesc& fd = frames[tile_frame(frames, hdr->n_frames, tile)];
    const bool active = (fd.flags & FRAME_BACKWARD) != 0;
    const int lt = (int)(tile - fd.tile_off);
    const int tx = lt % fd.tiles_x, ty = lt / fd.tiles_x;
    const int ux = tx * TS + (w % UX) * kUnit, uy = ty * TS + (w / UX) * kUnit;
    const int px = ux + (lane & 7), py = uy + (lane >> 3);
    const bool inside = active && px < fd.W && py < fd.H;
    const float pxf = (float)px, pyf = (float)py;
    const float x0 = (float)ux, x1 = (float)(ux + kUnit - 1);
    const float y0 = (float)uy, y1 = (float)(uy + kUnit - 1);
    const uint2 range = a.ranges[tile];
    const uint2 range_far = a.ranges_far ? a.ranges_far[tile] : make_uint2(0u, 0u);
    const int n_near = (int)(range.y - range.x);
    const int n_list = n_near + (int)(range_far.y - range_far.x);
    // instance k of the tile's list (near, then far)
    auto id_at = [&](int k) -> uint32_t {
      return k < n_near ? a.ids[range.x + k] : a.ids_far[range_far.x + (uint32_t)(k - n_near)];
    };
    const size_t lp = (size_t)py * fd.W + px;
    const size_t pid = fd.pix_off + lp;
    const size_t HW = (size_t)fd.W * fd.H;

    const int last = inside ? (int)a.n_contrib[pid] : 0;
    const float Tf = inside ? a.final_T[pid] : 0.0f;
    float gC0 = 0.0f, gC1 = 0.0f, gC2 = 0.0f, gA = 0.0f, gD = 0.0f;
    if (inside) {
      const float* gcol = a.dL_dcolor + 3 * fd.pix_off;
      gC0 = gcol[lp];
      gC1 = gcol[HW + lp];
      gC2 = gcol[2 * HW + lp];
      if constexpr (AUX) {
        gA = a.dL_dalpha ? a.dL_dalpha[pid] : 0.0f;
        gD = a.dL_ddepth ? a.dL_ddepth[pid] : 0.0f;
      }
    }
    const float bgdot = (fd.bg[0] * gC0 + fd.bg[1] * gC1) + fd.bg[2] * gC2;

    int wmax = last;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wmax = max(wmax, __shfl_xor(wmax, off, 64));
    wmax = min(__builtin_amdgcn_readfirstlane(wmax), n_list);
    if (lane == 0) atomicMax(&s_max, (uint32_t)wmax);
    __syncthreads();
    const int nproc = (int)s_max;

    float T = Tf;
    // colour (depth, alpha) accumulated behind the next contributor:
    // acc <- a_prev c_prev + (1 - a_prev) acc, formed when a contributor is
    // passed (the operands and order of backward.cu's update at the next one)
    float acc0 = 0.0f, acc1 = 0.0f, acc2 = 0.0f, accD = 0.0f, accA = 0.0f;

    // chunk c covers [nproc - 64 (c + 1), nproc - 64 c) clipped at 0; lane l
    // holds instance start + l.  Ids two chunks ahead; records and each
    // instance's record position one chunk ahead.  Lanes before the start
    // gather splat 0 (present when nproc > 0).
    uint32_t id_n = 0;
    Splat cur = {};
    float qcur = 0.0f;
    uint2 rect_cur = make_uint2(0u, 0u);
    uint32_t off_cur = 0;
    if (nproc > 0) {
      const int k = nproc - 64 + lane;
      id_n = k >= 0 ? id_at(k) : 0u;
      const uint32_t id0 = id_n < P ? id_n : 0u;
      const SplatRec& r = a.splats[id0];
      cur = r.s;
      qcur = r.q1;
      rect_cur = make_uint2(r.rect0, r.rect1);
      off_cur = r.item_off + a.rec_base[id0 / PRE_TILE];
      id_n = k >= 64 ? id_at(k - 64) : 0u;
    }
    for (int cend = nproc; cend > 0; cend -= 64) {
      const int cstart = cend - 64;  // may be negative: lanes below 0 are empty
      const Splat s = cur;
      const float kq = qcur;
      const uint2 srect = rect_cur;
      const uint32_t soff = off_cur;
      if (cstart > 0) {
        const uint32_t idc = id_n < P ? id_n : 0u;
        const SplatRec& r = a.splats[idc];
        cur = r.s;
        qcur = r.q1;
        rect_cur = make_uint2(r.rect0, r.rect1);
        off_cur = r.item_off + a.rec_base[idc / PRE_TILE];
        const int k2 = cstart - 128 + lane;
        id_n = k2 >= 0 ? id_at(k2) : 0u;
      }
      const int kk_l = cstart + lane;
      n_loaded += (uint32_t)min(64, cend);
      const bool valid = kk_l >= 0 && kk_l < wmax;
      const bool touch = valid && extent_touches(s.a.x, s.a.y, unpack_extent(s.c.w), x0, x1, y0, y1) &&
                         ellipse_touches_q(s.a.x, s.a.y, s.a.z, s.a.w, s.b.x, kq, x0, x1, y0, y1);
      uint64_t m = __ballot(touch);
      uint64_t cm = 0;
      if constexpr (kStage) {
        if (m) s_stage[w][lane] = s;  // read back by this wave only: LDS keeps its order
        __builtin_amdgcn_wave_barrier();
      }
      while (m) {
        const int j = 63 - __builtin_clzll(m);
        m &= ~(1ull << j);
        const int kk = cstart + j;
        float mx, my, ca, cb, cc, o, dz, sk, cr, cg, cbl;
        if constexpr (kStage) {
          const Splat q = s_stage[w][j];
          mx = q.a.x; my = q.a.y; ca = q.a.z; cb = q.a.w;
          cc = q.b.x; o = q.b.y; dz = q.b.z; sk = q.b.w;
          cr = q.c.x; cg = q.c.y; cbl = q.c.z;
        } else {
          mx = bcast(s.a.x, j); my = bcast(s.a.y, j);
          ca = bcast(s.a.z, j); cb = bcast(s.a.w, j); cc = bcast(s.b.x, j);
          o = bcast(s.b.y, j); dz = bcast(s.b.z, j); sk = bcast(s.b.w, j);
          cr = bcast(s.c.x, j); cg = bcast(s.c.y, j); cbl = bcast(s.c.z, j);
        }
        // straight-line per lane: every lane evaluates, selects keep the
        // state of lanes that do not contribute (no per-branch zero copies)
        const bool act = kk < last;
        if (__ballot(act) == 0ull) continue;
        const float dx = mx - pxf, dy = my - pyf;
        const float pw = power_of(ca, cb, cc, dx, dy);
        const float G = exp_neg(pw);
        const float oG = o * G;
        const float al = fminf(0.99f, oG);
        const bool contrib = act && (pw <= 0.0f && pw >= sk) && al >= a_min;
        n_eval += act ? 1u : 0u;
        n_con += contrib ? 1u : 0u;
        if (__ballot(contrib) == 0ull) continue;
        const float r1 = __builtin_amdgcn_rcpf(1.0f - al);
        const float Tn = T * r1;
        const float aT = al * Tn;
        float dLda = ((cr - acc0) * gC0 + (cg - acc1) * gC1) + (cbl - acc2) * gC2;
        if constexpr (AUX) {
          dLda = dLda + (dz - accD) * gD;
          dLda = dLda + (1.0f - accA) * gA;
        }
        dLda = dLda * Tn;
        dLda = dLda - (Tf * r1) * bgdot;
        const float aTm = contrib ? aT : 0.0f;
        const float g_r = aTm * gC0;
        const float g_g = aTm * gC1;
        const float g_b = aTm * gC2;
        const float g_d = AUX ? aTm * gD : 0.0f;
        const bool free_alpha = contrib && oG <= 0.99f;  // a clamped alpha passes no gradient
        const float g_op = free_alpha ? G * dLda : 0.0f;  // the record's moments
        const float g_mx = g_op * dx, g_my = g_op * dy;
        const float g_ca = g_mx * dx, g_cb = g_mx * dy, g_cc = g_my * dy;
        {  // selects, not branches: the state of non-contributing lanes is kept
          const float nl = 1.0f - al;
          const float n0 = al * cr + nl * acc0;
          const float n1 = al * cg + nl * acc1;
          const float n2 = al * cbl + nl * acc2;
          T = contrib ? Tn : T;
          acc0 = contrib ? n0 : acc0;
          acc1 = contrib ? n1 : acc1;
          acc2 = contrib ? n2 : acc2;
          if constexpr (AUX) {
            const float nD = al * dz + nl * accD;
            const float nA = al + nl * accA;
            accD = contrib ? nD : accD;
            accA = contrib ? nA : accA;
          }
        }
        {
          // rows of s0: Sx, Sxx, Sy, Sxy; s1: Syy, r, S, g; s2: b, 0, depth, 0
          float s0, s1, s2;
          wave_sum10(g_mx, g_my, g_ca, g_cb, g_cc, g_op, g_r, g_g, g_b, g_d, s0, s1, s2);
          if (row_tail) {
            s_part[w][fa][j] = s0;
            s_part[w][4 + fa][j] = s1;
            if ((row & 1) == 0) s_part[w][8 + (row >> 1)][j] = s2;
          }
          cm |= 1ull << j;
        }
      }
      if (lane == 0) s_cm[w] = cm;
      __syncthreads();  // also orders this chunk's staged reads before the next chunk's writes
      // record of splat j: sums of its slots in (wave, row) order, written as
      // five 8-byte pairs; thread t takes pairs t, t + 64 UPT, ... (pair p is
      // fields 2 (p mod 5), +1 of splat p / 5).  Every wave holds the chunk's
      // rectangles and offsets lane by lane, so the position comes by shuffle.
      uint64_t any = 0;
#pragma unroll
      for (int v = 0; v < UPT; v++) any |= s_cm[v];
      const uint32_t rx0 = srect.x & 0xFFFFu, ry0 = srect.x >> 16, rx1 = srect.y & 0xFFFFu;
      const uint32_t pos_l = soff + ((uint32_t)ty - ry0) * (rx1 - rx0) + ((uint32_t)tx - rx0);
      for (int pr = tid; pr < 64 * (GR_FIELDS / 2); pr += 64 * UPT) {  // uniform per wave
        const int j = pr / (GR_FIELDS / 2), fp = pr - j * (GR_FIELDS / 2);
        const uint32_t pos = (uint32_t)__shfl((int)pos_l, j, 64);
        if (((any >> j) & 1ull) && pos < a.capacity) {
          float s0 = 0.0f, s1 = 0.0f;
#pragma unroll
          for (int v = 0; v < UPT; v++) {
            if ((s_cm[v] >> j) & 1ull) {
              s0 += s_part[v][2 * fp][j];
              s1 += s_part[v][2 * fp + 1][j];
            }
          }
          *reinterpret_cast<float2*>(a.rec + (size_t)pos * GR_FIELDS + 2 * fp) = make_float2(s0, s1);
          if (fp == 0) atomicOr(a.flag + (pos >> 5), 1u << (pos & 31u));
        }
      }
      __syncthreads();  // the slots and masks are rewritten by the next chunk
    }
    __syncthreads();  // s_tile and s_max are reused by the next tile
  }
  if (a.stats) {
    unsigned long long e = n_eval, c = n_con, kf = (tid == 0) ? n_loaded : 0u;
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

// ---- the backward's tile order: heaviest tiles first ----
// The forward blend leaves per tile its number of contributing (pixel,
// splat) pairs, the measure of its backward work.  Tiles are bucketed by its
// binary logarithm and listed from the heaviest bucket down; tiles without
// work (no contributor, or a frame without the backward) are left out.
// Taken in this order the long tiles start first instead of finishing last
// (the order within a bucket is whatever the atomics give; it changes which
// wave takes a tile, never a result).  Each entry also carries an issue
// priority for the wave that takes it (bits 30-31): 3 for the heaviest
// nonempty bucket, 2 and 1 for the next two, 0 below: the heaviest tiles set
// the kernel's length, and their waves share a SIMD with lighter ones.
constexpr int kCostBuckets = 33;  // 0: no work; b >= 1: cost in [2^(b-1), 2^b)

W2W_DEV uint32_t cost_bucket(uint32_t c) { return c ? 32u - (uint32_t)__builtin_clz(c) : 0u; }

struct TileOrderArgs {
  const void* table;
  uint32_t tiles_cap;
  const uint32_t* cost;  // per tile
  uint32_t* hist;        // [kCostBuckets], zeroed
  uint32_t* cursor;      // [kCostBuckets], zeroed
  uint32_t* n_order;     // tiles listed
  uint32_t* order;       // [tiles]
  uint32_t include_zero; // list the tiles without cost too (last): the forward blend's order
};

__global__ __launch_bounds__(256) void k_tile_cost_hist(TileOrderArgs a) {
  __shared__ uint32_t h[kCostBuckets];
  if (threadIdx.x < kCostBuckets) h[threadIdx.x] = 0u;
  __syncthreads();
  const uint32_t T = min(static_cast<const BatchHeader*>(a.table)->n_tiles, a.tiles_cap);
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += gridDim.x * blockDim.x)
    atomicAdd(&h[cost_bucket(a.cost[t])], 1u);
  __syncthreads();
  if (threadIdx.x < kCostBuckets && h[threadIdx.x]) atomicAdd(a.hist + threadIdx.x, h[threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_tile_order(TileOrderArgs a) {
  __shared__ uint32_t off[kCostBuckets], cnt[kCostBuckets], base[kCostBuckets];
  __shared__ uint32_t s_top;
  const uint32_t T = min(static_cast<const BatchHeader*>(a.table)->n_tiles, a.tiles_cap);
  if (threadIdx.x == 0) {
    uint32_t o = 0, top = 0;
    for (int b = kCostBuckets - 1; b >= 1; b--) {
      off[b] = o;
      o += a.hist[b];
      if (!top && a.hist[b]) top = (uint32_t)b;
    }
    off[0] = o;
    s_top = top;
    if (blockIdx.x == 0) *a.n_order = o + (a.include_zero ? a.hist[0] : 0u);
  }
  for (uint32_t r0 = blockIdx.x * blockDim.x; r0 < T; r0 += gridDim.x * blockDim.x) {  // uniform
    if (threadIdx.x < kCostBuckets) cnt[threadIdx.x] = 0u;
    __syncthreads();
    const uint32_t t = r0 + threadIdx.x;
    const uint32_t b = t < T ? cost_bucket(a.cost[t]) : 0u;
    const bool listed = t < T && (b || a.include_zero);
    const uint32_t rank = listed ? atomicAdd(&cnt[b], 1u) : 0u;
    __syncthreads();
    if (threadIdx.x < kCostBuckets && cnt[threadIdx.x])
      base[threadIdx.x] = atomicAdd(a.cursor + threadIdx.x, cnt[threadIdx.x]);
    __syncthreads();
    if (listed) {
      const uint32_t prio = b + 3u > s_top ? 3u - (s_top - b) : 0u;
      a.order[off[b] + base[b] + rank] = t | (prio << 30);
    }
    __syncthreads();  // cnt and base are rewritten by the next round
  }
}

// Backward blend for tile edges 8 and 16: one wave per tile, lane l holding
// pixel (l mod 8, l / 8) of each of the tile's NB 8x8 blocks (NB = 1 or 4);
// waves claim tiles on their own, heaviest first, through the order of
// k_tile_order; no barriers, no LDS merge.
//
// Per chunk of 64 instances (walking the list back to front from the
// tile's largest last contributor) lane l gathers instance base + l and
// tests it against each block (cull extent, then the exact ellipse test);
// for each instance that touches a block with a pixel still behind its last
// contributor, the pixels recompute the forward's alpha bit for bit and
// T <- T / (1 - alpha) from the stored final T (the hardware rcp), each lane
// adding its pixels' moments (GradField) in block order; the colour of the
// contributors behind each pixel enters only through its dot product with
// dL/dcolour, one running sum per pixel.  One wave sum over the lanes gives
// the instance's record, stored straight from the wave.  Every sum has a
// fixed order: deterministic.
// per-pixel state of the backward walk
struct BwdPix {
  int last;                            // 1-based last contributor (0: none)
  float T;                             // transmittance in front of the current instance
  float accG;                          // colour behind the pixel . dL/dcolour
  float accD, accA;                    // depth and alpha behind (AUX)
  float gC0, gC1, gC2, gA, gD;         // dL/dcolour, dL/dalpha, dL/ddepth
  float Tfbg;                          // final T * (background . dL/dcolour)
};

template <bool AUX>
W2W_DEV void bwd_pix_load(const BlendBwdArgs& a, const FrameDesc& fd, int px, int py, int n_list, BwdPix& q) {
  const bool inside = px < fd.W && py < fd.H;
  const size_t HW = (size_t)fd.W * fd.H;
  const size_t lp = (size_t)py * fd.W + px;
  const size_t pid = fd.pix_off + lp;
  const float* gcol = a.dL_dcolor + 3 * fd.pix_off;
  q.last = inside ? min((int)a.n_contrib[pid], n_list) : 0;
  const float Tf = inside ? a.final_T[pid] : 0.0f;
  q.gC0 = inside ? gcol[lp] : 0.0f;
  q.gC1 = inside ? gcol[HW + lp] : 0.0f;
  q.gC2 = inside ? gcol[2 * HW + lp] : 0.0f;
  q.gA = q.gD = 0.0f;
  if constexpr (AUX) {
    q.gA = inside && a.dL_dalpha ? a.dL_dalpha[pid] : 0.0f;
    q.gD = inside && a.dL_ddepth ? a.dL_ddepth[pid] : 0.0f;
  }
  const float bgdot = (fd.bg[0] * q.gC0 + fd.bg[1] * q.gC1) + fd.bg[2] * q.gC2;
  q.Tfbg = Tf * bgdot;
  q.T = Tf;
  q.accG = q.accD = q.accA = 0.0f;
}

// One (instance, pixel) pair of the walk: if the instance contributes to
// the pixel, add its moments to F and step the pixel's state back over it.
// Returns whether any lane of the wave contributed.
template <bool AUX, bool STATS>
W2W_DEV bool bwd_pair(float mx, float my, float ca, float cb, float cc, float o, float dz, float sk, float cr,
                      float cg, float cbl, float pxf, float pyf, int kk, BwdPix& q, float (&F)[GR_FIELDS],
                      uint32_t& n_eval, uint32_t& n_con) {
  const float a_min = 0x1.010102p-8f;
  // straight-line per lane: selects keep the state of lanes that do not
  // contribute
  const bool act = kk < q.last;
  const float dx = mx - pxf, dy = my - pyf;
  const float pw = power_of(ca, cb, cc, dx, dy);
  const float G = exp_neg(pw);
  const float oG = o * G;
  const float al = fminf(0.99f, oG);
  const bool contrib = act && (pw <= 0.0f && pw >= sk) && al >= a_min;
  if constexpr (STATS) {
    n_eval += act ? 1u : 0u;
    n_con += contrib ? 1u : 0u;
  }
  if (__ballot(contrib) == 0ull) return false;
  const float r1 = __builtin_amdgcn_rcpf(1.0f - al);
  const float Tn = q.T * r1;
  const float aT = al * Tn;
  const float cdg = (cr * q.gC0 + cg * q.gC1) + cbl * q.gC2;
  float dLda = cdg - q.accG;
  if constexpr (AUX) {
    dLda = dLda + (dz - q.accD) * q.gD;
    dLda = dLda + (1.0f - q.accA) * q.gA;
  }
  dLda = dLda * Tn - q.Tfbg * r1;
  const bool free_alpha = contrib && oG <= 0.99f;  // a clamped alpha passes no gradient
  const float qq = free_alpha ? G * dLda : 0.0f;
  const float aTm = contrib ? aT : 0.0f;
  const float qdx = qq * dx, qdy = qq * dy;
  F[GA_SX] += qdx;
  F[GA_SY] += qdy;
  F[GA_SXX] = __builtin_fmaf(qdx, dx, F[GA_SXX]);
  F[GA_SXY] = __builtin_fmaf(qdx, dy, F[GA_SXY]);
  F[GA_SYY] = __builtin_fmaf(qdy, dy, F[GA_SYY]);
  F[GA_S] += qq;
  F[GA_R] = __builtin_fmaf(aTm, q.gC0, F[GA_R]);
  F[GA_G] = __builtin_fmaf(aTm, q.gC1, F[GA_G]);
  F[GA_B] = __builtin_fmaf(aTm, q.gC2, F[GA_B]);
  if constexpr (AUX) F[GA_DEPTH] = __builtin_fmaf(aTm, q.gD, F[GA_DEPTH]);
  const float nl = 1.0f - al;
  const float nG = __builtin_fmaf(al, cdg, nl * q.accG);
  q.T = contrib ? Tn : q.T;
  q.accG = contrib ? nG : q.accG;
  if constexpr (AUX) {
    const float nD = __builtin_fmaf(al, dz, nl * q.accD);
    const float nA = __builtin_fmaf(nl, q.accA, al);
    q.accD = contrib ? nD : q.accD;
    q.accA = contrib ? nA : q.accA;
  }
  return true;
}

// records move as native 4-element vectors (copies of HIP's float4 struct into
// LDS lower to memcpys that stay in scratch); the record's last 16 bytes are
// read as integers (threshold bits, rectangle, record offset): a bit cast of
// one element of a float vector here read element 0's bits
using v4f = float __attribute__((ext_vector_type(4)));
using v4u = uint32_t __attribute__((ext_vector_type(4)));

struct BwdShared {
  v4f stage[4][3][64];  // each wave's chunk (splat a, b, c), read back by broadcast
};

// One tile, walked by this wave: blocks b = 0 .. NB-1 of every lane.
template <int TS, bool AUX, bool STATS>
W2W_DEV void bwd_tile(const BlendBwdArgs& a, BwdShared& sh, const FrameDesc& fd, uint32_t tile, uint32_t P,
                      int lane, int wv, uint32_t& n_eval, uint32_t& n_con, uint32_t& n_loaded) {
  constexpr int BX = TS / kUnit;
  constexpr int NBL = BX * BX;  // blocks per lane
  v4f(*const stage)[64] = sh.stage[wv];
  const int row = lane >> 4;
  const bool row_tail = (lane & 15) == 15;
  const int fa = (row & 1) ? (row >> 1) + 2 : (row >> 1);  // wave_sum10 rows -> fields 0, 2, 1, 3
  const int lt = (int)(tile - fd.tile_off);
  const int tx = lt % fd.tiles_x, ty = lt / fd.tiles_x;
  const uint2 range = a.ranges[tile];
  const uint2 range_far = a.ranges_far ? a.ranges_far[tile] : make_uint2(0u, 0u);
  const int n_near = (int)(range.y - range.x);
  const int n_list = n_near + (int)(range_far.y - range_far.x);
  auto id_at = [&](int k) -> uint32_t {
    return k < n_near ? a.ids[range.x + k] : a.ids_far[range_far.x + (uint32_t)(k - n_near)];
  };
  // block b of this lane: origin (bx0[b], by0[b]) in pixels
  int bx0[NBL], by0[NBL], wmax_b[NBL];
  BwdPix q[NBL];
  int wmax = 0;
#pragma unroll
  for (int b = 0; b < NBL; b++) {
    bx0[b] = tx * TS + (b % BX) * kUnit;
    by0[b] = ty * TS + (b / BX) * kUnit;
    bwd_pix_load<AUX>(a, fd, bx0[b] + (lane & 7), by0[b] + (lane >> 3), n_list, q[b]);
    int wm = q[b].last;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) wm = max(wm, __shfl_xor(wm, off, 64));
    wmax_b[b] = __builtin_amdgcn_readfirstlane(wm);
    wmax = max(wmax, wmax_b[b]);
  }
  if (wmax == 0) return;
  const float pxf = (float)(lane & 7), pyf = (float)(lane >> 3);  // offsets inside a block

  // chunk c covers [wmax - 64 (c + 1), wmax - 64 c) clipped at 0; lane l
  // holds instance start + l.  Ids two chunks ahead; records, rectangles and
  // record offsets one chunk ahead.  Lanes before the start gather splat 0
  // (present when wmax > 0).
  // one 64-byte record per instance: splat a, b, c and (threshold,
  // rectangle, record offset)
  const v4f* const rec4 = reinterpret_cast<const v4f*>(a.splats);
  const v4u* const rec4u = reinterpret_cast<const v4u*>(a.splats);
  uint32_t id_n, base;
  v4f ra, rb, rc;
  v4u rk;
  {
    const int k = wmax - 64 + lane;
    id_n = k >= 0 ? id_at(k) : 0u;
    const uint32_t id0 = id_n < P ? id_n : 0u;
    ra = rec4[4 * (size_t)id0];
    rb = rec4[4 * (size_t)id0 + 1];
    rc = rec4[4 * (size_t)id0 + 2];
    rk = rec4u[4 * (size_t)id0 + 3];
    base = a.rec_base[id0 / PRE_TILE];
    id_n = k >= 64 ? id_at(k - 64) : 0u;
  }
  for (int cend = wmax; cend > 0; cend -= 64) {
    const int cstart = cend - 64;  // may be negative: lanes below 0 are empty
    const v4f sa = ra, sb = rb, sc = rc;
    const uint32_t kq_bits = rk.x;
    const float kq = __uint_as_float(kq_bits);
    const uint2 srect = make_uint2(rk.y, rk.z);
    const uint32_t soff = rk.w + base;
    if (cstart > 0) {
      const uint32_t idc = id_n < P ? id_n : 0u;
      ra = rec4[4 * (size_t)idc];
      rb = rec4[4 * (size_t)idc + 1];
      rc = rec4[4 * (size_t)idc + 2];
      rk = rec4u[4 * (size_t)idc + 3];
      base = a.rec_base[idc / PRE_TILE];
      const int k2 = cstart - 128 + lane;
      id_n = k2 >= 0 ? id_at(k2) : 0u;
    }
    if constexpr (STATS) n_loaded += (uint32_t)min(64, cend);
    const int kk_l = cstart + lane;
    const float2 ext = unpack_extent(sc.w);
    uint64_t mb[NBL];
    uint64_t m = 0;
#pragma unroll
    for (int b = 0; b < NBL; b++) {
      const float x0 = (float)bx0[b], y0 = (float)by0[b];
      const float x1 = x0 + (float)(kUnit - 1), y1 = y0 + (float)(kUnit - 1);
      const bool touch = kk_l >= 0 && kk_l < wmax_b[b] && extent_touches(sa.x, sa.y, ext, x0, x1, y0, y1) &&
                         ellipse_touches_q(sa.x, sa.y, sa.z, sa.w, sb.x, kq, x0, x1, y0, y1);
      mb[b] = __ballot(touch);
      m |= mb[b];
    }
    // the record position of each lane's instance: the item's first
    // instance plus the tile's index in the item's rectangle
    const uint32_t rx0 = srect.x & 0xFFFFu, ry0 = srect.x >> 16, rx1 = srect.y & 0xFFFFu;
    const uint32_t pos_l = soff + ((uint32_t)ty - ry0) * (rx1 - rx0) + ((uint32_t)tx - rx0);
    if (m) {
      stage[0][lane] = sa;  // read back by this wave only: LDS keeps its order
      stage[1][lane] = sb;
      stage[2][lane] = sc;
      __builtin_amdgcn_wave_barrier();
      // survivors back to front; the next one's LDS read is in flight while
      // the current one is evaluated
      int jn = 63 - __builtin_clzll(m);
      m &= ~(1ull << jn);
      v4f na = stage[0][jn], nb = stage[1][jn], nc = stage[2][jn];
      for (;;) {
        const int j = jn;
        const v4f qa = na, qb = nb, qc = nc;
        jn = m ? 63 - __builtin_clzll(m) : -1;
        if (jn >= 0) {
          m &= ~(1ull << jn);
          na = stage[0][jn];
          nb = stage[1][jn];
          nc = stage[2][jn];
        }
        const int kk = cstart + j;
        float F[GR_FIELDS];
#pragma unroll
        for (int f = 0; f < GR_FIELDS; f++) F[f] = 0.0f;
        bool any = false;
#pragma unroll
        for (int b = 0; b < NBL; b++) {
          if (!((mb[b] >> j) & 1ull)) continue;  // uniform
          any |= bwd_pair<AUX, STATS>(qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w, qc.x, qc.y, qc.z,
                                      pxf + (float)bx0[b], pyf + (float)by0[b], kk, q[b], F, n_eval, n_con);
        }
        if (any) {
          // rows of s0: Sx, Sxx, Sy, Sxy; s1: Syy, r, S, g; s2: b, 0, depth, 0
          float s0, s1, s2;
          wave_sum10(F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], s0, s1, s2);
          const uint32_t pos = (uint32_t)__builtin_amdgcn_readlane((int)pos_l, j);
          if (pos < a.capacity) {
            float* r = a.rec + (size_t)pos * GR_FIELDS;
            if (row_tail) {
              r[fa] = s0;
              r[4 + fa] = s1;
              if ((row & 1) == 0) r[8 + (row >> 1)] = s2;
            }
            if (lane == 0) atomicOr(a.flag + (pos >> 5), 1u << (pos & 31u));
          }
        }
        if (jn < 0) break;
      }
      __builtin_amdgcn_wave_barrier();  // this chunk's reads precede the next chunk's writes
    }
  }
}

// Heavy tiles of edge 16 (at least kCoopCost contributing pairs): a single
// wave's walk of such a tile outlasts the rest of the kernel, so the
// workgroup walks it together.  Wave w holds 8x8 block w (lane l: pixel
// (l mod 8, l / 8)); the four waves step through the same chunks of the
// tile's list (down from the tile's largest last contributor), each sums an
// instance's ten fields over its own lanes, and after every chunk the
// threads add the waves' sums in wave order and write the records.  Fixed
// order: deterministic (the sums round differently from the one-wave walk,
// which adds a lane's four blocks first; whether a tile is heavy depends
// only on the tile's own contributors).
constexpr uint32_t kCoopCost = 1u << 15;

struct BwdCoopShared {
  float part[4][GR_FIELDS][64];  // per wave, per field, per instance of the chunk
  uint64_t cm[4];                // instances with a sum, per wave
  int wmax[4];                   // each block's largest last contributor
};

template <bool AUX, bool STATS>
W2W_DEV void bwd_tile_coop(const BlendBwdArgs& a, BwdShared& sh, BwdCoopShared& cs, const FrameDesc& fd,
                           uint32_t tile, uint32_t P, int lane, int wv, uint32_t& n_eval, uint32_t& n_con,
                           uint32_t& n_loaded) {
  constexpr int TS = 16;
  v4f(*const stage)[64] = sh.stage[wv];
  const int tid = wv * 64 + lane;
  const int row = lane >> 4;
  const bool row_tail = (lane & 15) == 15;
  const int fa = (row & 1) ? (row >> 1) + 2 : (row >> 1);  // wave_sum10 rows -> fields 0, 2, 1, 3
  const int lt = (int)(tile - fd.tile_off);
  const int tx = lt % fd.tiles_x, ty = lt / fd.tiles_x;
  const uint2 range = a.ranges[tile];
  const uint2 range_far = a.ranges_far ? a.ranges_far[tile] : make_uint2(0u, 0u);
  const int n_near = (int)(range.y - range.x);
  const int n_list = n_near + (int)(range_far.y - range_far.x);
  auto id_at = [&](int k) -> uint32_t {
    return k < n_near ? a.ids[range.x + k] : a.ids_far[range_far.x + (uint32_t)(k - n_near)];
  };
  const int bx0 = tx * TS + (wv & 1) * kUnit, by0 = ty * TS + (wv >> 1) * kUnit;
  BwdPix q;
  bwd_pix_load<AUX>(a, fd, bx0 + (lane & 7), by0 + (lane >> 3), n_list, q);
  int wm = q.last;
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) wm = max(wm, __shfl_xor(wm, off, 64));
  wm = __builtin_amdgcn_readfirstlane(wm);
  if (lane == 0) cs.wmax[wv] = wm;
  __syncthreads();
  const int wmax = max(max(cs.wmax[0], cs.wmax[1]), max(cs.wmax[2], cs.wmax[3]));
  const float pxf = (float)(bx0 + (lane & 7)), pyf = (float)(by0 + (lane >> 3));
  const float x0 = (float)bx0, y0 = (float)by0;
  const float x1 = x0 + (float)(kUnit - 1), y1 = y0 + (float)(kUnit - 1);
  const v4f* const rec4 = reinterpret_cast<const v4f*>(a.splats);
  const v4u* const rec4u = reinterpret_cast<const v4u*>(a.splats);
  uint32_t id_n = 0, base = 0;
  v4f ra = {}, rb = {}, rc = {};
  v4u rk = {};
  if (wmax > 0) {
    const int k = wmax - 64 + lane;
    id_n = k >= 0 ? id_at(k) : 0u;
    const uint32_t id0 = id_n < P ? id_n : 0u;
    ra = rec4[4 * (size_t)id0];
    rb = rec4[4 * (size_t)id0 + 1];
    rc = rec4[4 * (size_t)id0 + 2];
    rk = rec4u[4 * (size_t)id0 + 3];
    base = a.rec_base[id0 / PRE_TILE];
    id_n = k >= 64 ? id_at(k - 64) : 0u;
  }
  for (int cend = wmax; cend > 0; cend -= 64) {  // the same trip count in every wave
    const int cstart = cend - 64;
    const v4f sa = ra, sb = rb, sc = rc;
    const float kq = __uint_as_float(rk.x);
    const uint2 srect = make_uint2(rk.y, rk.z);
    const uint32_t soff = rk.w + base;
    if (cstart > 0) {
      const uint32_t idc = id_n < P ? id_n : 0u;
      ra = rec4[4 * (size_t)idc];
      rb = rec4[4 * (size_t)idc + 1];
      rc = rec4[4 * (size_t)idc + 2];
      rk = rec4u[4 * (size_t)idc + 3];
      base = a.rec_base[idc / PRE_TILE];
      const int k2 = cstart - 128 + lane;
      id_n = k2 >= 0 ? id_at(k2) : 0u;
    }
    if constexpr (STATS) {
      if (wv == 0) n_loaded += (uint32_t)min(64, cend);
    }
    const int kk_l = cstart + lane;
    const bool touch = kk_l >= 0 && kk_l < wm && extent_touches(sa.x, sa.y, unpack_extent(sc.w), x0, x1, y0, y1) &&
                       ellipse_touches_q(sa.x, sa.y, sa.z, sa.w, sb.x, kq, x0, x1, y0, y1);
    uint64_t m = __ballot(touch);
    uint64_t cm = 0;
    const uint32_t rx0 = srect.x & 0xFFFFu, ry0 = srect.x >> 16, rx1 = srect.y & 0xFFFFu;
    const uint32_t pos_l = soff + ((uint32_t)ty - ry0) * (rx1 - rx0) + ((uint32_t)tx - rx0);
    if (m) {
      stage[0][lane] = sa;  // read back by this wave only: LDS keeps its order
      stage[1][lane] = sb;
      stage[2][lane] = sc;
      __builtin_amdgcn_wave_barrier();
      int jn = 63 - __builtin_clzll(m);
      m &= ~(1ull << jn);
      v4f na = stage[0][jn], nb = stage[1][jn], nc = stage[2][jn];
      for (;;) {
        const int j = jn;
        const v4f qa = na, qb = nb, qc = nc;
        jn = m ? 63 - __builtin_clzll(m) : -1;
        if (jn >= 0) {
          m &= ~(1ull << jn);
          na = stage[0][jn];
          nb = stage[1][jn];
          nc = stage[2][jn];
        }
        float F[GR_FIELDS];
#pragma unroll
        for (int f = 0; f < GR_FIELDS; f++) F[f] = 0.0f;
        if (bwd_pair<AUX, STATS>(qa.x, qa.y, qa.z, qa.w, qb.x, qb.y, qb.z, qb.w, qc.x, qc.y, qc.z, pxf, pyf,
                                 cstart + j, q, F, n_eval, n_con)) {
          float s0, s1, s2;
          wave_sum10(F[0], F[1], F[2], F[3], F[4], F[5], F[6], F[7], F[8], F[9], s0, s1, s2);
          if (row_tail) {
            cs.part[wv][fa][j] = s0;
            cs.part[wv][4 + fa][j] = s1;
            if ((row & 1) == 0) cs.part[wv][8 + (row >> 1)][j] = s2;
          }
          cm |= 1ull << j;
        }
        if (jn < 0) break;
      }
    }
    if (lane == 0) cs.cm[wv] = cm;
    __syncthreads();  // also orders this chunk's staged reads before the next chunk's writes
    // instance j's record: its waves' sums added in wave order, five 8-byte
    // pairs (pair p: fields 2 (p mod 5), +1 of instance p / 5)
    const uint64_t any = cs.cm[0] | cs.cm[1] | cs.cm[2] | cs.cm[3];
    for (int pr = tid; pr < 64 * (GR_FIELDS / 2); pr += 256) {  // uniform per wave
      const int j = pr / (GR_FIELDS / 2), fp = pr - j * (GR_FIELDS / 2);
      const uint32_t pos = (uint32_t)__shfl((int)pos_l, j, 64);
      if (((any >> j) & 1ull) && pos < a.capacity) {
        float v0 = 0.0f, v1 = 0.0f;
#pragma unroll
        for (int v = 0; v < 4; v++) {
          if ((cs.cm[v] >> j) & 1ull) {
            v0 += cs.part[v][2 * fp][j];
            v1 += cs.part[v][2 * fp + 1][j];
          }
        }
        *reinterpret_cast<float2*>(a.rec + (size_t)pos * GR_FIELDS + 2 * fp) = make_float2(v0, v1);
        if (fp == 0) atomicOr(a.flag + (pos >> 5), 1u << (pos & 31u));
      }
    }
    __syncthreads();  // the sums and masks are rewritten by the next chunk
  }
  __syncthreads();  // cs.wmax is rewritten by the next tile
}

// Without the alpha/depth terms (the training step) at least 4 waves per
// SIMD (128 VGPRs; the compiler otherwise takes 154 and 3 waves, and the
// walk's dependent chains are then hidden less well: 0.866 -> 0.808 ms per
// launch at the benchmark size, with 8 bytes per lane spilled).  The AUX
// variants would spill 116-124 bytes per lane at 4 waves and keep their own
// choice.
template <int TS, bool AUX, bool STATS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(AUX ? 1 : 4, AUX ? 8 : 4)))
void k_blend_bwd_tile(BlendBwdArgs a) {
  __shared__ BwdShared sh;
  __shared__ BwdCoopShared cs;
  __shared__ uint32_t s_claim;
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t Tn = min(hdr->n_tiles, a.tiles_cap);
  const uint32_t P = min(hdr->n_items, a.items_cap);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  const int lane = (int)(threadIdx.x & 63u);
  const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const uint32_t n_order = min(*a.n_order, Tn);
  uint32_t n_eval = 0, n_con = 0, n_loaded = 0;
  // heavy tiles lead the order: the workgroup claims entries together while
  // they are heavy; the first light one goes to wave 0, and from then on
  // every wave claims its own
  uint32_t pending = 0xffffffffu;
  if constexpr (TS == 16) {
    for (;;) {
      if (threadIdx.x == 0) s_claim = atomicAdd(a.ticket, 1u);
      __syncthreads();
      const uint32_t t = s_claim;
      __syncthreads();  // s_claim is rewritten by the next claim
      if (t >= n_order) break;
      const uint32_t e = a.order[t];
      const uint32_t tile = e & kOrderTile;
      if (tile >= Tn) continue;
      if (a.tile_cost[tile] < kCoopCost) {
        pending = t;
        break;
      }
      __builtin_amdgcn_s_setprio(3);
      const FrameDesc& fd = frames[tile_frame(frames, hdr->n_frames, tile)];
      bwd_tile_coop<AUX, STATS>(a, sh, cs, fd, tile, P, lane, wv, n_eval, n_con, n_loaded);
    }
  }
  for (;;) {
    uint32_t t = 0;
    if (wv == 0 && pending != 0xffffffffu) {
      t = pending;
      pending = 0xffffffffu;
    } else {
      if (lane == 0) t = atomicAdd(a.ticket, 1u);
      t = (uint32_t)__builtin_amdgcn_readfirstlane(__shfl((int)t, 0, 64));
    }
    if (t >= n_order) break;
    const uint32_t e = a.order[t];
    const uint32_t tile = e & kOrderTile;
    if (tile >= Tn) continue;
    // the heaviest tiles set the kernel's length: their waves take issue
    // priority over the waves of lighter tiles sharing the SIMD
    switch (e >> 30) {
      case 3: __builtin_amdgcn_s_setprio(3); break;
      case 2: __builtin_amdgcn_s_setprio(2); break;
      case 1: __builtin_amdgcn_s_setprio(1); break;
      default: __builtin_amdgcn_s_setprio(0); break;
    }
    const FrameDesc& fd = frames[tile_frame(frames, hdr->n_frames, tile)];
    bwd_tile<TS, AUX, STATS>(a, sh, fd, tile, P, lane, wv, n_eval, n_con, n_loaded);
  }
  __builtin_amdgcn_s_setprio(0);
  if (STATS && a.stats) {
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

struct PreBwdArgs {
  const void* table;
  uint32_t scenes_cap, items_cap, gaussians_cap;
  const float* means;    // packed [G,3]
  const float* scales;   // [G,3] activated
  const float* rots;     // [G,4] normalised
  const float* sh;       // packed SH blocks
  const int32_t* radii;  // per item
  const uint32_t* counts;    // per item: instances emitted
  const float* rec;          // [capacity][GR_FIELDS] per-instance partials
  const uint32_t* flag;      // [capacity / 32] bits
  uint32_t capacity;
  const SplatRec* splats;    // per item: mean2D, conic and opacity of the forward
  const uint32_t* rec_base;  // per PRE_TILE items: first record position (plus the splat record's offset)
  const uint32_t* item_frame;  // per item: its frame
  float* item_grad;          // five planes of items_cap floats: dL/dconic a, b, c, dL/dopacity, dL/ddepth
  float* dL_dmeans2D;    // per item [3]  (NDC units, z = 0)
  float* dL_dcolors;     // per item [3]
  float* dL_dopacity;    // per Gaussian
  float* dL_dmeans3D;    // per Gaussian [3]
  float* dL_dcov3D;      // per Gaussian [6]
  float* dL_dsh;         // packed like sh
  float* dL_dscales;     // per Gaussian [3]
  float* dL_drots;       // per Gaussian [4]
};

// One thread per item: the sum of its instance records in emission order
// (the flags are scanned eight at a time; the flag array is padded to 256
// bytes), then the splat constants applied to the summed moments (GradField).
// Writes the item's dL/dmean2D and dL/dcolor outputs and the other
// gradients to item_grad for k_preprocess_bwd.  A light kernel with four
// times the threads of the per-Gaussian one: the dependent flag -> record
// loads are hidden by occupancy rather than serialised over a Gaussian's
// views.
__global__ __launch_bounds__(256) void k_item_grads(PreBwdArgs a) {
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t n_items = min(hdr->n_items, a.items_cap);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += gridDim.x * blockDim.x) {
    const FrameDesc& fd = frames[a.item_frame[i]];
    const bool vis = a.radii[i] > 0 && (fd.flags & FRAME_BACKWARD);
    float ga[GR_FIELDS];
#pragma unroll
    for (int f = 0; f < GR_FIELDS; f++) ga[f] = 0.0f;
    if (vis) {
      const uint32_t k0 = a.rec_base[i / PRE_TILE] + a.splats[i].item_off;  // the item's first record
      const uint32_t k1 = min(k0 + a.counts[i], a.capacity);
      for (uint32_t kw = k0 & ~31u; kw < k1; kw += 32) {
        uint32_t f32 = a.flag[kw >> 5];
        if (kw < k0) f32 &= ~0u << (k0 - kw);
        if (k1 - kw < 32) f32 &= (1u << (k1 - kw)) - 1u;
        while (f32) {
          const uint32_t b = (uint32_t)__builtin_ctz(f32);
          f32 &= f32 - 1u;
          const float2* r = reinterpret_cast<const float2*>(a.rec + (size_t)(kw + b) * GR_FIELDS);
#pragma unroll
          for (int f = 0; f < GR_FIELDS / 2; f++) {
            const float2 rv = r[f];
            ga[2 * f] += rv.x;
            ga[2 * f + 1] += rv.y;
          }
        }
      }
      // moments -> gradients of the 2D splat (dL/dmean2D in NDC units:
      // d pixel / d ndc = W/2, H/2)
      const Splat sp = a.splats[i].s;
      const float ca = sp.a.z, cb = sp.a.w, cc = sp.b.x, o = sp.b.y;
      const float sx = ga[GA_SX], sy = ga[GA_SY];
      const float sxx = ga[GA_SXX], sxy = ga[GA_SXY], syy = ga[GA_SYY];
      ga[GA_MX] = (-(ca * sx + cb * sy) * o) * (0.5f * (float)fd.W);
      ga[GA_MY] = (-(cb * sx + cc * sy) * o) * (0.5f * (float)fd.H);
      ga[GA_CA] = (-0.5f * o) * sxx;
      ga[GA_CB] = -o * sxy;
      ga[GA_CC] = (-0.5f * o) * syy;
      // ga[GA_OP] = S is dL/dopacity as it stands
    }
    a.dL_dmeans2D[3 * (size_t)i] = ga[GA_MX];
    a.dL_dmeans2D[3 * (size_t)i + 1] = ga[GA_MY];
    a.dL_dmeans2D[3 * (size_t)i + 2] = 0.0f;
    a.dL_dcolors[3 * (size_t)i] = ga[GA_R];
    a.dL_dcolors[3 * (size_t)i + 1] = ga[GA_G];
    a.dL_dcolors[3 * (size_t)i + 2] = ga[GA_B];
    // one plane per field: the per-Gaussian kernel's lanes (a Gaussian's
    // views are consecutive items) then read 16 bytes apart instead of 128
    const size_t pl = a.items_cap;
    a.item_grad[i] = ga[GA_CA];
    a.item_grad[pl + i] = ga[GA_CB];
    a.item_grad[2 * pl + i] = ga[GA_CC];
    a.item_grad[3 * pl + i] = ga[GA_OP];
    a.item_grad[4 * pl + i] = ga[GA_DEPTH];
  }
}

// One thread per Gaussian: loops over the views of its scene in this batch,
// recomputes each view's projection and sums the chain rule in registers, so
// every per-Gaussian gradient is written once (no atomics).
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_preprocess_bwd(PreBwdArgs a) {
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t G = min(hdr->n_gaussians, a.gaussians_cap);
  const uint32_t n_scenes = min(hdr->n_scenes, a.scenes_cap);
  const SceneDesc* scenes = table_scenes(a.table);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t bg = blockIdx.x * blockDim.x + threadIdx.x; bg < G; bg += stride) {
    const uint32_t si = upper_index(n_scenes, bg, [&](uint32_t k) { return scenes[k].g_off; });
    const SceneDesc& sd = scenes[si];
    const uint32_t g = bg - sd.g_off;
    const uint32_t gg = sd.g_base + g;  // row in the parameter / gradient arrays
    float gm[3] = {0.0f, 0.0f, 0.0f};
    float gs[6] = {0, 0, 0, 0, 0, 0};  // dL/dSigma summed over views (xx, xy, xz, yy, yz, zz)
    float gop = 0.0f;
    const float p[3] = {a.means[3 * gg], a.means[3 * gg + 1], a.means[3 * gg + 2]};
    const float s[3] = {a.scales[3 * gg], a.scales[3 * gg + 1], a.scales[3 * gg + 2]};
    const float q[4] = {a.rots[4 * gg], a.rots[4 * gg + 1], a.rots[4 * gg + 2], a.rots[4 * gg + 3]};
    float sig[6];  // Sigma does not depend on the view: once per Gaussian
    {
      float R0[9], M0[9];
      sigma_of(q, s, R0, M0, sig);
    }
    for (uint32_t v = 0; v < sd.n_views; v++) {
      const uint32_t i = sd.item_off + g * sd.n_views + v;
      if (i >= a.items_cap) break;
      const FrameDesc& fd = frames[sd.frame_off + v];
      const bool vis = a.radii[i] > 0 && (fd.flags & FRAME_BACKWARD);
      // the item's record sums (k_item_grads)
      float ga[GR_FIELDS];
      for (int f = 0; f < GR_FIELDS; f++) ga[f] = 0.0f;
      if (vis) {
        ga[GA_MX] = a.dL_dmeans2D[3 * (size_t)i];
        ga[GA_MY] = a.dL_dmeans2D[3 * (size_t)i + 1];
        const size_t pl = a.items_cap;
        ga[GA_CA] = a.item_grad[i];
        ga[GA_CB] = a.item_grad[pl + i];
        ga[GA_CC] = a.item_grad[2 * pl + i];
        ga[GA_OP] = a.item_grad[3 * pl + i];
        ga[GA_DEPTH] = a.item_grad[4 * pl + i];
      }
      if (!vis) continue;
      gop += ga[GA_OP];
      Cam c;
      load_cam(c, fd);
      Geom geo;
      geometry_sig(c, p, sig, geo);
      if (!geo.visible) continue;
      const float* V = c.V;
      const float* F = c.F;
      // NDC mean and depth through the full projection
      {
        float gnx = ga[GA_MX], gny = ga[GA_MY], gnz = ga[GA_DEPTH];
        float gcx = gnx * geo.invW, gcy = gny * geo.invW, gcz = gnz * geo.invW;
        float ginvW = (gnx * geo.cx + gny * geo.cy) + gnz * geo.cz;
        float gcw = -ginvW * geo.invW * geo.invW;
#pragma unroll
        for (int j = 0; j < 3; j++)
          gm[j] += ((F[4 * j + 0] * gcx + F[4 * j + 1] * gcy) + F[4 * j + 2] * gcz) + F[4 * j + 3] * gcw;
      }
      // conic -> cov2D
      const float gca = ga[GA_CA], gcb = ga[GA_CB], gcc = ga[GA_CC];
      const float ca = geo.ca, cb = geo.cb, cc = geo.cc, h = 0.5f * gcb;
      const float q00 = ca * gca + cb * h, q01 = ca * h + cb * gcc;
      const float q10 = cb * gca + cc * h, q11 = cb * h + cc * gcc;
      const float g_cxx = -(q00 * ca + q01 * cb);
      const float g_cxy = -2.0f * (q00 * cb + q01 * cc);
      const float g_cyy = -(q10 * cb + q11 * cc);
      // cov2D -> Sigma and J
      const float* S6 = geo.Sig;
      float vv[3], ww[3];
      vv[0] = (S6[0] * geo.jx[0] + S6[1] * geo.jx[1]) + S6[2] * geo.jx[2];
      vv[1] = (S6[1] * geo.jx[0] + S6[3] * geo.jx[1]) + S6[4] * geo.jx[2];
      vv[2] = (S6[2] * geo.jx[0] + S6[4] * geo.jx[1]) + S6[5] * geo.jx[2];
      ww[0] = (S6[0] * geo.jy[0] + S6[1] * geo.jy[1]) + S6[2] * geo.jy[2];
      ww[1] = (S6[1] * geo.jy[0] + S6[3] * geo.jy[1]) + S6[4] * geo.jy[2];
      ww[2] = (S6[2] * geo.jy[0] + S6[4] * geo.jy[1]) + S6[5] * geo.jy[2];
      const float hxy = 0.5f * g_cxy;
      float GS[9];
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int l = 0; l < 3; l++)
          GS[3 * k + l] = (g_cxx * geo.jx[k] * geo.jx[l] + hxy * (geo.jx[k] * geo.jy[l] + geo.jy[k] * geo.jx[l])) +
                          g_cyy * geo.jy[k] * geo.jy[l];
      gs[0] += GS[0]; gs[1] += GS[1]; gs[2] += GS[2];
      gs[3] += GS[4]; gs[4] += GS[5]; gs[5] += GS[8];
      float gjx[3], gjy[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        gjx[k] = 2.0f * g_cxx * vv[k] + g_cxy * ww[k];
        gjy[k] = g_cxy * vv[k] + 2.0f * g_cyy * ww[k];
      }
      // J -> view-space point
      const float zi = 1.0f / geo.vz, zi2 = zi * zi;
      float g_zi = 0.0f, g_zi2 = 0.0f, g_txc = 0.0f, g_tyc = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float right = V[4 * k + 0], up = V[4 * k + 1], fwd = V[4 * k + 2];
        g_zi += gjx[k] * c.fx * right + gjy[k] * c.fy * up;
        g_zi2 -= gjx[k] * c.fx * geo.txc * fwd + gjy[k] * c.fy * geo.tyc * fwd;
        g_txc -= gjx[k] * c.fx * fwd * zi2;
        g_tyc -= gjy[k] * c.fy * fwd * zi2;
      }
      float g_vz = g_zi * (-zi * zi) + g_zi2 * (-2.0f * zi2 * zi);
      const float g_rxc = g_txc * geo.vz, g_ryc = g_tyc * geo.vz;
      g_vz += g_txc * geo.rxc + g_tyc * geo.ryc;
      float g_rx, g_ry;
      if (geo.clamped) {
        const float s3 = geo.s * geo.s * geo.s;
        const float lx2 = c.limx * c.limx, ly2 = c.limy * c.limy;
        g_rx = g_rxc * (geo.s - s3 * geo.rx * geo.rx / lx2) + g_ryc * (-s3 * geo.rx * geo.ry / lx2);
        g_ry = g_rxc * (-s3 * geo.rx * geo.ry / ly2) + g_ryc * (geo.s - s3 * geo.ry * geo.ry / ly2);
      } else {
        g_rx = g_rxc;
        g_ry = g_ryc;
      }
      const float g_vx = g_rx / geo.vz, g_vy = g_ry / geo.vz;
      g_vz -= (g_rx * geo.rx + g_ry * geo.ry) / geo.vz;
      const float g_tx = g_vx / geo.tw, g_ty = g_vy / geo.tw, g_tz = g_vz / geo.tw;
      const float g_tw = -((g_vx * geo.vx + g_vy * geo.vy) + g_vz * geo.vz) / geo.tw;
#pragma unroll
      for (int j = 0; j < 3; j++)
        gm[j] += ((V[4 * j + 0] * g_tx + V[4 * j + 1] * g_ty) + V[4 * j + 2] * g_tz) + V[4 * j + 3] * g_tw;
    }
    // Sigma = M M^T with M = R diag(s): dL/dSigma is linear, so the scale and
    // quaternion gradients are taken once from its sum over views
    {
      float R[9], Mm[9];
      rot_scale(q, s, R, Mm);
      const float GS[9] = {gs[0], gs[1], gs[2], gs[1], gs[3], gs[4], gs[2], gs[4], gs[5]};
      float gM[9], gR[9];
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int col = 0; col < 3; col++)
          gM[3 * r + col] = 2.0f * ((GS[3 * r + 0] * Mm[col] + GS[3 * r + 1] * Mm[3 + col]) + GS[3 * r + 2] * Mm[6 + col]);
#pragma unroll
      for (int r = 0; r < 3; r++)
#pragma unroll
        for (int col = 0; col < 3; col++) gR[3 * r + col] = gM[3 * r + col] * s[col];
#pragma unroll
      for (int col = 0; col < 3; col++)
        a.dL_dscales[3 * (size_t)gg + col] = (gM[col] * R[col] + gM[3 + col] * R[3 + col]) + gM[6 + col] * R[6 + col];
      const float qw = q[0], qx = q[1], qy = q[2], qz = q[3];
      float* dq = a.dL_drots + 4 * (size_t)gg;
      dq[0] = 2.0f * (((-qz * gR[1] + qy * gR[2]) + (qz * gR[3] - qx * gR[5])) + (-qy * gR[6] + qx * gR[7]));
      dq[1] = 2.0f * ((((qy * gR[1] + qz * gR[2]) + (qy * gR[3] - 2.0f * qx * gR[4])) + (-qw * gR[5] + qz * gR[6])) + (qw * gR[7] - 2.0f * qx * gR[8]));
      dq[2] = 2.0f * ((((-2.0f * qy * gR[0] + qx * gR[1]) + (qw * gR[2] + qx * gR[3])) + (qz * gR[5] - qw * gR[6])) + (qz * gR[7] - 2.0f * qy * gR[8]));
      dq[3] = 2.0f * ((((-2.0f * qz * gR[0] - qw * gR[1]) + (qx * gR[2] + qw * gR[3])) + (-2.0f * qz * gR[4] + qy * gR[5])) + (qx * gR[6] + qy * gR[7]));
    }
    a.dL_dopacity[gg] = gop;
    // the SH kernel that runs next adds the view-direction term
    a.dL_dmeans3D[3 * (size_t)gg] = gm[0];
    a.dL_dmeans3D[3 * (size_t)gg + 1] = gm[1];
    a.dL_dmeans3D[3 * (size_t)gg + 2] = gm[2];
    float* dcov = a.dL_dcov3D + 6 * (size_t)gg;
    dcov[0] = gs[0]; dcov[1] = 2.0f * gs[1]; dcov[2] = 2.0f * gs[2];
    dcov[3] = gs[3]; dcov[4] = 2.0f * gs[4]; dcov[5] = gs[5];
  }
}

struct ShBwdArgs {
  const void* table;
  uint32_t scenes_cap, items_cap, gaussians_cap;
  const float* means;       // packed [G,3]
  const float* sh;          // packed SH blocks
  const int32_t* radii;     // per item
  const float* dL_dcolors;  // per item [3], written by k_preprocess_bwd
  float* dL_dsh;            // packed like sh
  float* dL_dmeans3D;       // per Gaussian [3]: the view-direction term is added
  // optional (adam_m non-null): Adam on the SH coefficients in place of
  // writing dL_dsh -- sh_out (= sh, the parameter store's SH block), adam_m
  // and adam_v packed like sh; step sizes hyper[hyper_dc] for the first three
  // coefficients of a row (the DC term), hyper[hyper_ac] for the rest,
  // hyper[16..20] = 1/sqrt(1 - b2^t), 1 - b1, b2, 1 - b2, eps; no update
  // while *skip_if != 0
  float* sh_out;
  float* adam_m;
  float* adam_v;
  const float* hyper;
  int hyper_dc, hyper_ac;
  const uint32_t* skip_if;
};

// SH part of the colour backward (backward.cu:291-444 semantics, with our
// forward's degree-aware basis): one thread per Gaussian loops over its views;
// dL/dSH accumulates in registers (the degree is a template parameter, so the
// 3 (D+1)^2 accumulators and the basis are straight-line code) and is written
// once; the gradient through the view direction is added to dL/dmean3D.
// Coefficients per row staged in LDS by k_sh_bwd_adam (3 x 16: degree 3).
constexpr int SHA_W = 48;

template <int D>
W2W_DEV void sh_bwd_one(const ShBwdArgs& a, const SceneDesc& sd, const FrameDesc* frames, uint32_t g,
                        uint32_t gg, float* stage = nullptr) {
  constexpr int NB = (D + 1) * (D + 1);
  float acc[NB][3];
#pragma unroll
  for (int k = 0; k < NB; k++) acc[k][0] = acc[k][1] = acc[k][2] = 0.0f;
  const float p[3] = {a.means[3 * gg], a.means[3 * gg + 1], a.means[3 * gg + 2]};
  // the coefficients used, read once (16-byte loads when the row allows)
  constexpr int NF = 3 * NB;
  float cf[NF];
  {
    const float* row = a.sh + sd.sh_off + (size_t)g * sd.M * 3;
    if ((reinterpret_cast<uintptr_t>(row) & 15u) == 0) {
#pragma unroll
      for (int q = 0; q < NF / 4; q++) {
        const float4 v = reinterpret_cast<const float4*>(row)[q];
        cf[4 * q] = v.x; cf[4 * q + 1] = v.y; cf[4 * q + 2] = v.z; cf[4 * q + 3] = v.w;
      }
#pragma unroll
      for (int k = NF / 4 * 4; k < NF; k++) cf[k] = row[k];
    } else {
#pragma unroll
      for (int k = 0; k < NF; k++) cf[k] = row[k];
    }
  }
  float gm0 = 0.0f, gm1 = 0.0f, gm2 = 0.0f;
  for (uint32_t v = 0; v < sd.n_views; v++) {
    const uint32_t i = sd.item_off + g * sd.n_views + v;
    if (i >= a.items_cap) break;
    const FrameDesc& fd = frames[sd.frame_off + v];
    if (!(a.radii[i] > 0 && (fd.flags & FRAME_BACKWARD))) continue;
    const float gr0 = a.dL_dcolors[3 * (size_t)i], gr1 = a.dL_dcolors[3 * (size_t)i + 1];
    const float gr2 = a.dL_dcolors[3 * (size_t)i + 2];
    const float dx = p[0] - fd.cp[0], dy = p[1] - fd.cp[1], dz = p[2] - fd.cp[2];
    const float len = __builtin_sqrtf((dx * dx + dy * dy) + dz * dz);
    const float x = dx / len, y = dy / len, z = dz / len;
    float gdx = 0.0f, gdy = 0.0f, gdz = 0.0f;
    sh_visit<D>(x, y, z, [&](int k, float Y, float bx, float by, float bz) {
      acc[k][0] += Y * gr0;
      acc[k][1] += Y * gr1;
      acc[k][2] += Y * gr2;
      const float dot = (cf[3 * k] * gr0 + cf[3 * k + 1] * gr1) + cf[3 * k + 2] * gr2;
      gdx = __builtin_fmaf(bx, dot, gdx);
      gdy = __builtin_fmaf(by, dot, gdy);
      gdz = __builtin_fmaf(bz, dot, gdz);
    });
    const float pd = (gdx * x + gdy * y) + gdz * z;
    gm0 += (gdx - pd * x) / len;
    gm1 += (gdy - pd * y) / len;
    gm2 += (gdz - pd * z) / len;
  }
  a.dL_dmeans3D[3 * (size_t)gg] += gm0;
  a.dL_dmeans3D[3 * (size_t)gg + 1] += gm1;
  a.dL_dmeans3D[3 * (size_t)gg + 2] += gm2;
  if (a.adam_m && stage && 3 * sd.M <= SHA_W) {
    // the row's gradient to the workgroup's LDS block (zero above the
    // degree): k_sh_bwd_adam applies the step with coalesced accesses
#pragma unroll
    for (int f = 0; f < NF; f++) stage[f] = acc[f / 3][f % 3];
    for (int f = NF; f < 3 * sd.M; f++) stage[f] = 0.0f;
    return;
  }
  if (a.adam_m) {
    // the optimizer step on this row (every coefficient of the row, those
    // above the degree with a zero gradient), from the coefficients held in
    // cf: the same arithmetic as the store-wide Adam kernel.  Rows too wide
    // for the LDS block only: a thread walking its own row touches a
    // different cache line per lane and access.
    if (a.skip_if && *a.skip_if) return;
    const size_t row = sd.sh_off + (size_t)g * sd.M * 3;
    const float inv_bc2_sqrt = a.hyper[16], one_m_b1 = a.hyper[17], b2 = a.hyper[18];
    const float one_m_b2 = a.hyper[19], eps = a.hyper[20];
    const float ns_dc = a.hyper[a.hyper_dc], ns_ac = a.hyper[a.hyper_ac];
    const int nrow = sd.M * 3;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      float pv = cf[f], m = a.adam_m[row + f], v = a.adam_v[row + f];
      adam_one(pv, acc[f / 3][f % 3], m, v, f < 3 ? ns_dc : ns_ac, inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
      a.sh_out[row + f] = pv;
      a.adam_m[row + f] = m;
      a.adam_v[row + f] = v;
    }
    for (int f = NF; f < nrow; f++) {
      float pv = a.sh_out[row + f], m = a.adam_m[row + f], v = a.adam_v[row + f];
      adam_one(pv, 0.0f, m, v, f < 3 ? ns_dc : ns_ac, inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
      a.sh_out[row + f] = pv;
      a.adam_m[row + f] = m;
      a.adam_v[row + f] = v;
    }
    return;
  }
  float* dsh = a.dL_dsh + sd.sh_off + (size_t)g * sd.M * 3;
  if ((reinterpret_cast<uintptr_t>(dsh) & 15u) == 0) {
#pragma unroll
    for (int q = 0; q < NF / 4; q++) {
      const int f = 4 * q;
      reinterpret_cast<float4*>(dsh)[q] = make_float4(acc[f / 3][f % 3], acc[(f + 1) / 3][(f + 1) % 3],
                                                      acc[(f + 2) / 3][(f + 2) % 3],
                                                      acc[(f + 3) / 3][(f + 3) % 3]);
    }
#pragma unroll
    for (int f = NF / 4 * 4; f < NF; f++) dsh[f] = acc[f / 3][f % 3];
  } else {
#pragma unroll
    for (int k = 0; k < NB; k++) {
      dsh[3 * k] = acc[k][0];
      dsh[3 * k + 1] = acc[k][1];
      dsh[3 * k + 2] = acc[k][2];
    }
  }
  for (int k = NF; k < 3 * (int)sd.M; k++) dsh[k] = 0.0f;
}

__global__ __launch_bounds__(256) void k_sh_bwd(ShBwdArgs a) {
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t G = min(hdr->n_gaussians, a.gaussians_cap);
  const uint32_t n_scenes = min(hdr->n_scenes, a.scenes_cap);
  const SceneDesc* scenes = table_scenes(a.table);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t bg = blockIdx.x * blockDim.x + threadIdx.x; bg < G; bg += stride) {
    const uint32_t si = upper_index(n_scenes, bg, [&](uint32_t k) { return scenes[k].g_off; });
    const SceneDesc& sd = scenes[si];
    const uint32_t g = bg - sd.g_off;
    const uint32_t gg = sd.g_base + g;
    switch (sd.D) {
      case 0: sh_bwd_one<0>(a, sd, frames, g, gg); break;
      case 1: sh_bwd_one<1>(a, sd, frames, g, gg); break;
      case 2: sh_bwd_one<2>(a, sd, frames, g, gg); break;
      default: sh_bwd_one<3>(a, sd, frames, g, gg); break;
    }
  }
}

// k_sh_bwd with the SH coefficients' Adam step (a.adam_m set): a workgroup
// takes 256 consecutive Gaussians, each thread stages its row's gradient in
// LDS, then the threads walk the 256 rows coefficient by coefficient (the
// rows of a store's consecutive Gaussians are contiguous: coalesced) and
// apply the step to the parameter and its two moments.
__global__ __launch_bounds__(256) void k_sh_bwd_adam(ShBwdArgs a) {
  __shared__ float s_g[256][SHA_W + 1];
  __shared__ uint64_t s_row[256];  // first float of each thread's row
  __shared__ int s_w[256];         // staged coefficients of the row (0: none)
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t G = min(hdr->n_gaussians, a.gaussians_cap);
  const uint32_t n_scenes = min(hdr->n_scenes, a.scenes_cap);
  const SceneDesc* scenes = table_scenes(a.table);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  const int tid = (int)threadIdx.x;
  const bool skip = a.skip_if && *a.skip_if;
  const float inv_bc2_sqrt = a.hyper[16], one_m_b1 = a.hyper[17], b2 = a.hyper[18];
  const float one_m_b2 = a.hyper[19], eps = a.hyper[20];
  const float ns_dc = a.hyper[a.hyper_dc], ns_ac = a.hyper[a.hyper_ac];
  for (uint32_t base = blockIdx.x * 256u; base < G; base += gridDim.x * 256u) {  // uniform
    const uint32_t bg = base + (uint32_t)tid;
    s_w[tid] = 0;
    if (bg < G) {
      const uint32_t si = upper_index(n_scenes, bg, [&](uint32_t k) { return scenes[k].g_off; });
      const SceneDesc& sd = scenes[si];
      const uint32_t g = bg - sd.g_off;
      const uint32_t gg = sd.g_base + g;
      switch (sd.D) {
        case 0: sh_bwd_one<0>(a, sd, frames, g, gg, s_g[tid]); break;
        case 1: sh_bwd_one<1>(a, sd, frames, g, gg, s_g[tid]); break;
        case 2: sh_bwd_one<2>(a, sd, frames, g, gg, s_g[tid]); break;
        default: sh_bwd_one<3>(a, sd, frames, g, gg, s_g[tid]); break;
      }
      if (3 * sd.M <= SHA_W) {
        s_row[tid] = sd.sh_off + (uint64_t)g * sd.M * 3;
        s_w[tid] = 3 * sd.M;
      }
    }
    __syncthreads();
    if (!skip) {
      for (int e = tid; e < 256 * SHA_W; e += 256) {
        const int r = e / SHA_W, f = e - r * SHA_W;
        if (f < s_w[r]) {
          const uint64_t i = s_row[r] + (uint64_t)f;
          float pv = a.sh_out[i], m = a.adam_m[i], v = a.adam_v[i];
          adam_one(pv, s_g[r][f], m, v, f < 3 ? ns_dc : ns_ac, inv_bc2_sqrt, one_m_b1, b2, one_m_b2, eps);
          a.sh_out[i] = pv;
          a.adam_m[i] = m;
          a.adam_v[i] = v;
        }
      }
    }
    __syncthreads();  // the block's rows are rewritten by the next round
  }
}

}  // namespace w2w
