// Stable LSD radix sort of (key, u32 value) pairs, 8-bit digits, one pass per
// digit ("onesweep"): a histogram kernel counts every digit of every pass in
// one read of the keys; each pass kernel ranks a 6144-element tile with wave64
// ballot multisplit, resolves its global offsets by decoupled look-back over
// the preceding tiles, reorders the tile in LDS and writes coalesced runs.
//
// Replaces the 32-pass, host-synchronised 2-bit sort of prepareSort.cu:303-360
// (block sort, D2H histogram, host prefix, H2D, merge per pass).  The result
// is the same stable order.  Tiles are claimed through an atomic ticket, so a
// tile only waits on tiles claimed earlier by resident workgroups (forward
// progress without co-residency assumptions).  Every count and the element
// total must stay below 2^30 (look-back words keep 2 flag bits).
#pragma once
#include "w2w_device.hpp"

namespace w2w {

// 512 threads x 12 items: measured faster than 256 x 16 on 98M 14-bit keys
// (1.15 vs 1.31 ms for two passes): the wait on the look-back chain, not the
// ranking, bounds a pass, and fewer, larger tiles shorten it.
constexpr int SORT_THREADS = 512;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ITEMS = 12;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // 6144
constexpr int RADIX = 256;
constexpr int kLookAhead = 8;  // look-back words loaded per step
// A histogram accumulated by the producer of the keys may come in several
// copies, hist[copy][4][256]: each producing workgroup flushes its counts
// into copy blockIdx.x % copies, so a word takes 1/copies of the flush
// atomics (same-address atomics from many workgroups serialise).
// k_sort_hist sums the copies into the first before the passes read it.
constexpr uint32_t kHistCopyStride = 4 * RADIX;
constexpr uint32_t kHistCopies = 16;  // the tile sort's, from the emission
constexpr uint32_t LB_AGG = 1u << 30;
constexpr uint32_t LB_INC = 2u << 30;
constexpr uint32_t LB_VAL = (1u << 30) - 1u;

W2W_DEV uint32_t sort_count(const uint32_t* d_n, uint32_t n_cap) {
  if (!d_n) return n_cap;
  uint32_t n = *d_n;
  return n < n_cap ? n : n_cap;
}

// Histogram of every digit of every pass (hist[pass][256], pre-zeroed; with
// null keys the producer's copies, at most kHistCopies, summed; the top pass
// filled in without counting when the producer found every key's top digit
// equal to top_digit), and
// zeroing of the look-back words status[pass][tile][256] (pass stride
// tiles_stride tiles) and the per-pass tickets.
template <typename KT>
__global__ __launch_bounds__(256) void k_sort_hist(const KT* __restrict__ keys,
                                                   const uint32_t* __restrict__ d_n, uint32_t n_cap,
                                                   int begin_bit, int npasses, int last_bits,
                                                   uint32_t* __restrict__ hist,
                                                   uint32_t* __restrict__ status,
                                                   uint32_t tiles_stride,
                                                   uint32_t* __restrict__ tickets, uint32_t copies,
                                                   const uint32_t* __restrict__ top_differs,
                                                   uint32_t top_digit) {
  __shared__ uint32_t sh[8][RADIX];
  const uint32_t n = sort_count(d_n, n_cap);
  for (int i = threadIdx.x; i < 8 * RADIX; i += blockDim.x) (&sh[0][0])[i] = 0;
  __syncthreads();
  const uint32_t stride = gridDim.x * blockDim.x;
  if (keys) {  // null: the histogram was accumulated by the producer of the keys
    // top_differs (null: not known) == 0: every key's top digit is
    // top_digit, so its bin holds n and the top digits are not counted
    int cpasses = npasses;
    if (top_differs && *top_differs == 0u) {
      cpasses = npasses - 1;
      if (blockIdx.x == 0 && threadIdx.x == 0) hist[(npasses - 1) * RADIX + top_digit] = n;
    }
    auto count = [&](KT k) {
      for (int p = 0; p < cpasses; p++) {
        const int bits = (p == npasses - 1) ? last_bits : 8;
        const uint32_t d = (uint32_t)(k >> (begin_bit + 8 * p)) & ((1u << bits) - 1u);
        atomicAdd(&sh[p][d], 1u);
      }
    };
    // 16-byte loads, two per thread in flight, for keys of up to 32 bits at
    // a 16-byte aligned address: the grid is small (its workgroups flush
    // with global atomics), so a load per key at a time would leave the
    // kernel waiting on memory latency
    uint32_t head = 0;
    if constexpr (sizeof(KT) <= 4) {
      if ((reinterpret_cast<uintptr_t>(keys) & 15u) == 0) {
        constexpr uint32_t VEC = 16 / sizeof(KT);
        const uint32_t nv = n / VEC;
        head = nv * VEC;
        const uint4* kv = reinterpret_cast<const uint4*>(keys);
        auto count_word = [&](uint32_t w) {
          if constexpr (sizeof(KT) == 2) {
            count((KT)(w & 0xFFFFu));
            count((KT)(w >> 16));
          } else {
            count((KT)w);
          }
        };
        for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < nv; j += 2 * stride) {
          const bool two = j + stride < nv;
          const uint4 va = kv[j];
          const uint4 vb = two ? kv[j + stride] : make_uint4(0u, 0u, 0u, 0u);
          count_word(va.x); count_word(va.y); count_word(va.z); count_word(va.w);
          if (two) {
            count_word(vb.x); count_word(vb.y); count_word(vb.z); count_word(vb.w);
          }
        }
      }
    }
    for (uint32_t i = head + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) count(keys[i]);
    __syncthreads();
    for (int i = threadIdx.x; i < npasses * RADIX; i += blockDim.x) {
      uint32_t v = (&sh[0][0])[i];
      if (v) atomicAdd(&hist[i], v);
    }
  } else if (copies > 1 && blockIdx.x == 0) {
    for (int i = threadIdx.x; i < npasses * RADIX; i += blockDim.x) {
      uint32_t v[kHistCopies];
#pragma unroll
      for (uint32_t c = 0; c < kHistCopies; c++) v[c] = c < copies ? hist[c * kHistCopyStride + i] : 0u;
      uint32_t sum = 0;
#pragma unroll
      for (uint32_t c = 0; c < kHistCopies; c++) sum += v[c];
      hist[i] = sum;
    }
  }
  // pass p owns status[p * tiles_stride * RADIX ...]; zero the tiles this n uses
  const uint32_t ntiles = (n + SORT_TILE - 1) / SORT_TILE;
  const size_t words = (size_t)ntiles * RADIX;
  for (int p = 0; p < npasses; p++) {
    uint32_t* sp = status + (size_t)p * tiles_stride * RADIX;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) sp[i] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x < (unsigned)npasses) tickets[threadIdx.x] = 0;
}

// exclusive scan of one value per thread over a 256-thread block
W2W_DEV uint32_t block_excl_scan_256(uint32_t v, uint32_t* s_tmp /*[4]*/, uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inc = wave_incl_scan_u32(v, lane);
  if (lane == 63) s_tmp[wave] = inc;
  __syncthreads();
  uint32_t wpre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    uint32_t t = s_tmp[w];
    if (w < wave) wpre += t;
    tot += t;
  }
  __syncthreads();
  if (total) *total = tot;
  return wpre + inc - v;
}

// exclusive scan over the block of one value per thread (T threads, T/64 <= 16
// waves); every thread must call it
template <int T>
W2W_DEV uint32_t block_excl_scan(uint32_t v, uint32_t* s_tmp /*[T/64]*/) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = wave_incl_scan_u32(v, lane);
  if (lane == 63) s_tmp[wave] = inc;
  __syncthreads();
  uint32_t wpre = 0;
#pragma unroll
  for (int w = 0; w < T / 64; w++)
    if (w < wave) wpre += s_tmp[w];
  __syncthreads();
  return wpre + inc - v;
}

// Skipping the top pass of a sort whose keys share their top digit (null
// top_hist: never): the pass's buffers when it is skipped.
template <typename KT>
struct TopSkip {
  const uint32_t* top_hist;  // hist[npasses - 1][256]
  int is_top;
  const KT* kin;
  const uint32_t* vin;
  KT* kout;
  uint32_t* vout;
};

template <typename KT, bool WRITE_KEYS>
__global__ __launch_bounds__(SORT_THREADS) void k_sort_pass(
    const KT* __restrict__ kin, const uint32_t* __restrict__ vin, KT* __restrict__ kout,
    uint32_t* __restrict__ vout, const uint32_t* __restrict__ d_n, uint32_t n_cap, int shift,
    int nbits, const uint32_t* __restrict__ hist, uint32_t* __restrict__ status,
    uint32_t* __restrict__ ticket, TopSkip<KT> skip) {
  __shared__ KT s_k[SORT_TILE];
  __shared__ uint32_t s_v[SORT_TILE];
  __shared__ uint32_t s_wcnt[SORT_WAVES][RADIX];
  __shared__ uint32_t s_dstart[RADIX];
  __shared__ uint32_t s_gbase[RADIX];
  __shared__ uint32_t s_tmp[SORT_WAVES];
  __shared__ uint32_t s_tile;
  static_assert(SORT_THREADS >= RADIX, "one thread per digit");

  const uint32_t n = sort_count(d_n, n_cap);
  if (n == 0) return;  // e.g. the far lists of a step with no open tile: no ticket, no histogram scan
  const uint32_t ntiles = (n + SORT_TILE - 1) / SORT_TILE;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const uint32_t dmask = (1u << nbits) - 1u;
  if (skip.top_hist) {
    // every key has the same top digit (one bin holds all n): the top pass
    // would keep the order as it is, so it returns at once and the other
    // passes take the chain that ends in the output without it
    if (__syncthreads_or(tid < RADIX && skip.top_hist[tid] == n)) {
      if (skip.is_top) return;
      kin = skip.kin;
      vin = skip.vin;
      kout = skip.kout;
      vout = skip.vout;
    }
  }
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));

  // exclusive prefix of the global digit histogram (thread d < RADIX: digit d)
  const uint32_t hexcl = block_excl_scan<SORT_THREADS>(tid < RADIX ? hist[tid] : 0u, s_tmp);

  for (;;) {
    if (tid == 0) s_tile = atomicAdd(ticket, 1u);
#pragma unroll
    for (int q = 0; q < RADIX / 64; q++) s_wcnt[wave][lane + 64 * q] = 0;
    __syncthreads();
    const uint32_t tile = s_tile;
    if (tile >= ntiles) return;
    const uint32_t base = tile * SORT_TILE;
    const uint32_t wbase = base + wave * (SORT_ITEMS * 64);

    KT keys[SORT_ITEMS];
    uint32_t vals[SORT_ITEMS];
    uint32_t woff[SORT_ITEMS];
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
      uint32_t idx = wbase + j * 64 + lane;
      bool valid = idx < n;
      keys[j] = valid ? kin[idx] : (KT)0;
      vals[j] = valid ? vin[idx] : 0u;
    }
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
      uint32_t idx = wbase + j * 64 + lane;
      bool valid = idx < n;
      uint32_t d = (uint32_t)(keys[j] >> shift) & dmask;
      uint64_t peers = __ballot(valid);
      for (int b = 0; b < nbits; b++) {
        bool bit = (d >> b) & 1u;
        uint64_t bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
      }
      uint32_t rank = (uint32_t)__popcll(peers & lt_mask);
      uint32_t cnt = (uint32_t)__popcll(peers);
      uint32_t pre = valid ? s_wcnt[wave][d] : 0u;
      woff[j] = pre + rank;
      if (valid && rank == 0) s_wcnt[wave][d] = pre + cnt;
    }
    __syncthreads();
    // digit totals of this tile and the per-wave exclusive offsets; the
    // totals are published for the tiles behind at once
    uint32_t total = 0;
    uint32_t* st = status + (size_t)tile * RADIX;
    if (tid < RADIX) {
#pragma unroll
      for (int w = 0; w < SORT_WAVES; w++) {
        const uint32_t c = s_wcnt[w][tid];
        s_wcnt[w][tid] = total;
        total += c;
      }
      st_release_dev(st + tid, (tile == 0 ? LB_INC : LB_AGG) | total);
    }
    const uint32_t dstart = block_excl_scan<SORT_THREADS>(total, s_tmp);
    if (tid < RADIX) s_dstart[tid] = dstart;
    __syncthreads();
    // the tile reordered by digit in LDS before the look-back, which then
    // waits with the keys and values out of registers
#pragma unroll
    for (int j = 0; j < SORT_ITEMS; j++) {
      uint32_t idx = wbase + j * 64 + lane;
      if (idx < n) {
        uint32_t d = (uint32_t)(keys[j] >> shift) & dmask;
        uint32_t pos = s_dstart[d] + s_wcnt[wave][d] + woff[j];
        if (pos < (uint32_t)SORT_TILE) {
          s_k[pos] = keys[j];
          s_v[pos] = vals[j];
        }
      }
    }
    if (tid < RADIX) {
      // decoupled look-back for digit tid: the preceding tiles' words,
      // kLookAhead at a time (independent loads), consumed in order up to
      // the first inclusive one; a word not yet published is read again
      // from there
      uint32_t excl = 0;
      if (tile > 0) {
        int p = (int)tile - 1;
        for (;;) {
          uint32_t w[kLookAhead];
#pragma unroll
          for (int k = 0; k < kLookAhead; k++)
            w[k] = p - k >= 0 ? ld_acquire_dev(status + (size_t)(p - k) * RADIX + tid) : LB_INC;
          bool found = false;
          int k = 0;
          for (; k < kLookAhead; k++) {
            const uint32_t flag = w[k] & ~LB_VAL;
            if (flag == 0) break;
            excl += w[k] & LB_VAL;
            if (flag == LB_INC) {
              found = true;
              break;
            }
          }
          if (found) break;
          p -= k;
        }
        st_release_dev(st + tid, LB_INC | (excl + total));
      }
      s_gbase[tid] = hexcl + excl - dstart;
    }
    __syncthreads();
    const uint32_t tile_n = (n - base) < (uint32_t)SORT_TILE ? (n - base) : (uint32_t)SORT_TILE;
    for (uint32_t i = tid; i < tile_n; i += SORT_THREADS) {
      KT k = s_k[i];
      uint32_t d = (uint32_t)(k >> shift) & dmask;
      uint32_t o = s_gbase[d] + i;
      if (o < n) {
        if (WRITE_KEYS) kout[o] = k;
        vout[o] = s_v[i];
      }
    }
    __syncthreads();
  }
}

}  // namespace w2w
