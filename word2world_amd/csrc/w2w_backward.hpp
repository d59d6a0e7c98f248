// Backward rasterizer kernels for gfx950.
//
//   k_blend_bwd       per-pixel reverse traversal (the role of backward.cu:521-665):
//                     starts from the stored final T and last-contributor index,
//                     recomputes alpha with the forward's exact arithmetic and
//                     T <- T / (1 - alpha); no per-(pixel, Gaussian) state.  A
//                     workgroup owns a tile, one wave per 8x8 block; each wave
//                     reduces a splat's partials over its pixels with DPP, the
//                     waves' sums are added in LDS in a fixed order and written
//                     once per (tile, splat) instance to that instance's slot in
//                     emission order.  No atomics: the result is deterministic.
//   k_preprocess_bwd  per-Gaussian chain rule (the role of backward.cu:27-227 and
//                     :229-519), differentiating the forward exactly as defined
//                     in k_preprocess (radial clamp, degree-aware SH with the 0.3
//                     coefficient, NDC depth, no dilation); it first sums each
//                     item's instance records, which lie contiguously from the
//                     item's emission offset, in emission order.
#pragma once
#include "w2w_device.hpp"
#include "w2w_forward.hpp"

namespace w2w {

// per-instance gradient record (floats)
enum GradField : int {
  GA_MX = 0, GA_MY = 1,            // dL/dmean2D in NDC units
  GA_CA = 2, GA_CB = 3, GA_CC = 4, // dL/dconic (a, b, c of the power form)
  GA_OP = 5,                       // dL/dopacity
  GA_R = 6, GA_G = 7, GA_B = 8,    // dL/drgb
  GA_DEPTH = 9,                    // dL/d(NDC depth)
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
  const SplatRec* splats;   // per item: splat and block-test constants
  const uint2* rect;        // per item: tile rectangle (x0 | y0 << 16, x1 | y1 << 16)
  const uint32_t* item_off; // per item: emission position of its first instance
  unsigned long long* stats;  // optional [3]: evaluated pairs, contributing pairs, loaded instances
  const uint32_t* n_contrib;
  const float* final_T;
  const float* dL_dcolor;   // per frame [3,H,W] at 3*pix_off
  const float* dL_dalpha;   // per frame [H,W] at pix_off; may be null (zero)
  const float* dL_ddepth;   // per frame [H,W] at pix_off; may be null (zero)
  float* rec;               // [capacity][GR_FIELDS] per-instance partials
  uint8_t* flag;            // [capacity], zeroed: 1 where rec was written
  // depth-sliced binning: a tile's list is its near list followed by its far
  // list (null: one list)
  const uint2* ranges_far;
  const uint32_t* ids_far;
};

// Backward blend (backward.cu:396-520 semantics, with our forward's
// decisions): every pixel walks its list back to front from its last
// contributor, recovering T by T <- T / (1 - alpha) from the stored final T
// (the reciprocal of 1 - alpha >= 0.01 is taken with the hardware rcp, 1 ulp).
//
// A workgroup claims a tile; wave w owns 8x8 block w.  The waves walk the
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
  for (;;) {
    if (tid == 0) {
      s_tile = atomicAdd(a.ticket, 1u);
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
    const float hw = 0.5f * (float)fd.W, hh = 0.5f * (float)fd.H;

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
    float4 kcur = {};
    uint2 rect_cur = make_uint2(0u, 0u);
    uint32_t off_cur = 0;
    if (nproc > 0) {
      const int k = nproc - 64 + lane;
      id_n = k >= 0 ? id_at(k) : 0u;
      const uint32_t id0 = id_n < P ? id_n : 0u;
      const SplatRec r = a.splats[id0];
      cur = r.s;
      kcur = r.k;
      rect_cur = a.rect[id0];
      off_cur = a.item_off[id0];
      id_n = k >= 64 ? id_at(k - 64) : 0u;
    }
    for (int cend = nproc; cend > 0; cend -= 64) {
      const int cstart = cend - 64;  // may be negative: lanes below 0 are empty
      const Splat s = cur;
      const float4 kq = kcur;
      const uint2 srect = rect_cur;
      const uint32_t soff = off_cur;
      if (cstart > 0) {
        const uint32_t idc = id_n < P ? id_n : 0u;
        const SplatRec r = a.splats[idc];
        cur = r.s;
        kcur = r.k;
        rect_cur = a.rect[idc];
        off_cur = a.item_off[idc];
        const int k2 = cstart - 128 + lane;
        id_n = k2 >= 0 ? id_at(k2) : 0u;
      }
      const int kk_l = cstart + lane;
      n_loaded += (uint32_t)min(64, cend);
      const bool valid = kk_l >= 0 && kk_l < wmax;
      const bool touch = valid && extent_touches(s.a.x, s.a.y, unpack_extent(s.c.w), x0, x1, y0, y1) &&
                         ellipse_touches_k(s.a.x, s.a.y, s.a.z, s.a.w, s.b.x, kq, x0, x1, y0, y1);
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
        const float g_op = free_alpha ? G * dLda : 0.0f;
        const float dLdp = free_alpha ? (o * dLda) * G : 0.0f;
        const float g_ca = (-0.5f * dx * dx) * dLdp;
        const float g_cb = (-dx * dy) * dLdp;
        const float g_cc = (-0.5f * dy * dy) * dLdp;
        const float g_mx = (-(ca * dx + cb * dy) * dLdp) * hw;
        const float g_my = (-(cb * dx + cc * dy) * dLdp) * hh;
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
          // rows of s0: mx, ca, my, cb; s1: cc, r, op, g; s2: b, 0, depth, 0
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
          if (fp == 0) a.flag[pos] = 1;
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

struct PreBwdArgs {
  const void* table;
  uint32_t scenes_cap, items_cap, gaussians_cap;
  const float* means;    // packed [G,3]
  const float* scales;   // [G,3] activated
  const float* rots;     // [G,4] normalised
  const float* sh;       // packed SH blocks
  const int32_t* radii;  // per item
  const uint32_t* counts;    // per item: instances emitted
  const uint32_t* item_off;  // per item: emission position of the first
  const float* rec;          // [capacity][GR_FIELDS] per-instance partials
  const uint8_t* flag;       // [capacity]
  uint32_t capacity;
  const uint32_t* item_frame;  // per item: its frame
  float* item_grad;          // per item [8]: dL/dconic a, b, c, dL/dopacity, dL/ddepth, 3 unused
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
// bytes).  Writes the item's dL/dmean2D and dL/dcolor outputs and the other
// sums to item_grad for k_preprocess_bwd.  A light kernel with four times
// the threads of the per-Gaussian one: the dependent flag -> record loads
// are hidden by occupancy rather than serialised over a Gaussian's views.
__global__ __launch_bounds__(256) void k_item_grads(PreBwdArgs a) {
  const BatchHeader* hdr = static_cast<const BatchHeader*>(a.table);
  const uint32_t n_items = min(hdr->n_items, a.items_cap);
  const FrameDesc* frames = table_frames(a.table, a.scenes_cap);
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_items; i += gridDim.x * blockDim.x) {
    const bool vis = a.radii[i] > 0 && (frames[a.item_frame[i]].flags & FRAME_BACKWARD);
    float ga[GR_FIELDS];
#pragma unroll
    for (int f = 0; f < GR_FIELDS; f++) ga[f] = 0.0f;
    if (vis) {
      const uint32_t k0 = a.item_off[i];
      const uint32_t k1 = min(k0 + a.counts[i], a.capacity);
      for (uint32_t kw = k0 & ~7u; kw < k1; kw += 8) {
        uint64_t f8 = *reinterpret_cast<const uint64_t*>(a.flag + kw);
        if (kw < k0) f8 &= ~0ull << (8 * (k0 - kw));
        if (k1 - kw < 8) f8 &= ~(~0ull << (8 * (k1 - kw)));
        while (f8) {
          const uint32_t b = (uint32_t)__builtin_ctzll(f8) >> 3;
          f8 &= ~(0xFFull << (8 * b));
          const float2* r = reinterpret_cast<const float2*>(a.rec + (size_t)(kw + b) * GR_FIELDS);
#pragma unroll
          for (int f = 0; f < GR_FIELDS / 2; f++) {
            const float2 rv = r[f];
            ga[2 * f] += rv.x;
            ga[2 * f + 1] += rv.y;
          }
        }
      }
    }
    a.dL_dmeans2D[3 * (size_t)i] = ga[GA_MX];
    a.dL_dmeans2D[3 * (size_t)i + 1] = ga[GA_MY];
    a.dL_dmeans2D[3 * (size_t)i + 2] = 0.0f;
    a.dL_dcolors[3 * (size_t)i] = ga[GA_R];
    a.dL_dcolors[3 * (size_t)i + 1] = ga[GA_G];
    a.dL_dcolors[3 * (size_t)i + 2] = ga[GA_B];
    float4* ig = reinterpret_cast<float4*>(a.item_grad + 8 * (size_t)i);
    ig[0] = make_float4(ga[GA_CA], ga[GA_CB], ga[GA_CC], ga[GA_OP]);
    ig[1] = make_float4(ga[GA_DEPTH], 0.0f, 0.0f, 0.0f);
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
        const float4* ig = reinterpret_cast<const float4*>(a.item_grad + 8 * (size_t)i);
        const float4 t0 = ig[0], t1 = ig[1];
        ga[GA_CA] = t0.x;
        ga[GA_CB] = t0.y;
        ga[GA_CC] = t0.z;
        ga[GA_OP] = t0.w;
        ga[GA_DEPTH] = t1.x;
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
};

// SH part of the colour backward (backward.cu:291-444 semantics, with our
// forward's degree-aware basis): one thread per Gaussian loops over its views;
// dL/dSH accumulates in registers (the degree is a template parameter, so the
// 3 (D+1)^2 accumulators and the basis are straight-line code) and is written
// once; the gradient through the view direction is added to dL/dmean3D.
template <int D>
W2W_DEV void sh_bwd_one(const ShBwdArgs& a, const SceneDesc& sd, const FrameDesc* frames, uint32_t g,
                        uint32_t gg) {
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
  a.dL_dmeans3D[3 * (size_t)gg] += gm0;
  a.dL_dmeans3D[3 * (size_t)gg + 1] += gm1;
  a.dL_dmeans3D[3 * (size_t)gg + 2] += gm2;
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

}  // namespace w2w
