// sort.hip -- one LSD digit pass of a onesweep radix sort over (key, value)
// pairs, replacing radixSortB/radixSortC (shaders.wgsl:338-498).
//
// Differences from the reference's pass, all for gfx950:
//  - ranking is warp-level multi-split (WLMS) on wave64 ballots: BITS
//    ballots give each lane the mask of lanes holding the same digit, mbcnt
//    its rank among them, and a per-wave LDS counter (read by all peers,
//    bumped by the lowest peer) carries the rank across the wave's items.
//    Items are visited in input order, so the pass is stable -- the
//    reference ranks with LDS atomicAdd (shaders.wgsl:428-429), which is not.
//  - tiles are TILE = THREADS*ITEMS keys (4096, 8192 or 16384) taken in order from an
//    atomic counter by a persistent grid (reference: 1024-key tiles, one
//    workgroup each, shaders.wgsl:406-420); full tiles skip every per-key
//    bounds test.
//  - the digit width is per pass (5..8 or 10 bits): a last pass over the
//    top bits of a narrower key takes only those bits.
//  - the global digit offsets come from the upsweep histogram and are
//    scanned by every block on entry (radixSortB's serial scan disappears).
//  - the views of a batch sort side by side (view = blockIdx.y), each with
//    its own tile counter and status words.
//  - decoupled look-back per digit with agent-scope status words
//    (aggregate / inclusive flags + 30-bit count, as shaders.wgsl:438-457).
//  - keys, then values, are reordered through LDS so that global stores of
//    one digit run are contiguous (shaders.wgsl:463-497 does the same).
#include "spl_device.h"
#include "spl_internal.h"

namespace spl {

template <bool B>
struct Flag {
  static constexpr bool value = B;
};

template <int BITS, int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_onesweep_pass(SortPassLaunch a) {
  constexpr int RADIX = 1 << BITS;
  constexpr int WAVES = THREADS / kWave;
  constexpr int TILE = THREADS * ITEMS;
  constexpr int DPT = (RADIX + THREADS - 1) / THREADS;  // digits per thread
  constexpr uint32_t MASK = RADIX - 1;
  __shared__ uint32_t s_buf[TILE];
  __shared__ uint32_t s_wave_hist[WAVES * RADIX];
  __shared__ uint32_t s_digit_start[RADIX];
  __shared__ uint32_t s_local[RADIX];
  __shared__ uint32_t s_base[RADIX];
  __shared__ uint32_t s_tmp[WAVES];
  __shared__ uint32_t s_tile;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (kWave - 1), w = tid / kWave;
  const uint32_t v = blockIdx.y;
  if (a.clear_a) clear_words(a.clear_a + v * a.clear_a_stride, a.clear_a_words);
  if (a.clear_b) clear_words(a.clear_b + v * a.clear_b_stride, a.clear_b_words);
  uint32_t* counters = a.counters + (size_t)v * kCounterWords;
  const uint32_t* keys_in = a.keys_in + v * a.entry_stride;
  const uint32_t* vals_in = a.vals_in ? a.vals_in + v * a.entry_stride : nullptr;
  uint32_t* keys_out = a.keys_out + v * a.entry_stride;
  uint32_t* vals_out = a.vals_out + v * a.entry_stride;
  const uint32_t* hist = a.hist + v * a.hist_stride;
  uint32_t* status = a.status + v * a.status_stride;
  uint32_t* tile_counter = counters + a.tile_slot;

  uint32_t n = counters[a.count_slot];
  if (n > a.count_cap) n = a.count_cap;
  const uint32_t num_tiles = (n + TILE - 1) / TILE;
  const int shift = a.shift;

  {  // global exclusive digit starts
    uint32_t hv[DPT], sum = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
      const uint32_t d = tid * DPT + k;
      hv[k] = d < (uint32_t)RADIX ? hist[d] : 0u;
      sum += hv[k];
    }
    uint32_t total;
    uint32_t excl = block_exclusive_scan<THREADS>(sum, s_tmp, total);
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
      const uint32_t d = tid * DPT + k;
      if (d < (uint32_t)RADIX) s_digit_start[d] = excl;
      excl += hv[k];
    }
  }

  // the key bit that holds bit b of the digit, per b (uniform): a run-time
  // offset keeps the extraction a single v_bfe_i32.  Its offset field has 5
  // bits, so every pass needs shift + BITS <= 32: the depth keys' passes end
  // at bit 31 (8-bit digits) or 29 (10-bit), the tile keys' at bit 23.
  static_assert(BITS <= 16, "digit width");
  uint32_t bit_at[BITS];
#pragma unroll
  for (int b = 0; b < BITS; ++b) bit_at[b] = (uint32_t)shift + (uint32_t)b;

  // One tile.  FULL tiles (all but the last of a view) hold TILE keys and
  // skip every per-key bounds test.
  auto tile_pass = [&](auto full_tag, const uint32_t t, const uint32_t tile_n) {
    constexpr bool FULL = decltype(full_tag)::value;
    const uint32_t tile_base = t * TILE;

    // Register budget: keys die once they are in LDS, values are fetched
    // while the look-back waits, and the scatter keeps digits packed (four
    // 8-bit or two 16-bit to a register) instead of 32-bit positions.
    uint32_t key[ITEMS], pos[ITEMS];
    const uint32_t wave_base = tile_base + w * (kWave * ITEMS) + lane;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t idx = wave_base + j * kWave;
      key[j] = FULL || idx < n ? keys_in[idx] : 0xFFFFFFFFu;
    }

    // WLMS ranking, items in input order.  The wave's digit counters are
    // read and bumped through an LDS-qualified volatile pointer: the peers'
    // reads must see the leader's store of the previous item, and a generic
    // pointer would turn them into flat accesses that wait on global loads.
    volatile __attribute__((address_space(3))) uint32_t* wh =
        (volatile __attribute__((address_space(3))) uint32_t*)(s_wave_hist + w * RADIX);
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const bool valid = FULL || (wave_base + j * kWave) < n;
      const uint32_t d = (key[j] >> shift) & MASK;
      const uint64_t vm = FULL ? ~0ull : __ballot(valid);
      uint32_t plo = (uint32_t)vm, phi = (uint32_t)(vm >> 32);
#pragma unroll
      for (int b = 0; b < BITS; ++b) {
        // lanes whose bit b equals mine: the ballot where the bit is set,
        // its complement where clear (m is the bit sign-extended to a word;
        // one v_bitop3 per half computes acc & ~(bal ^ m), truth table 0x90)
        const uint32_t m = (uint32_t)__builtin_amdgcn_sbfe((int32_t)key[j], bit_at[b], 1);
        const uint64_t bal = __ballot(m != 0u);
        plo = __builtin_amdgcn_bitop3_b32(plo, (uint32_t)bal, m, 0x90);
        phi = __builtin_amdgcn_bitop3_b32(phi, (uint32_t)(bal >> 32), m, 0x90);
      }
      const uint64_t peers = ((uint64_t)phi << 32) | plo;
      const uint32_t below = lanes_below(peers);
      const uint32_t cnt = (uint32_t)__popc(plo) + (uint32_t)__popc(phi);
      uint32_t base = 0;
      if (valid) base = wh[d];
      __builtin_amdgcn_wave_barrier();
      if (valid && below == 0) wh[d] = base + cnt;
      __builtin_amdgcn_wave_barrier();
      pos[j] = base + below;
    }
    __syncthreads();

    // per digit: exclusive prefix over waves, tile count
    uint32_t cnt_d[DPT], sum = 0;
#pragma unroll
    for (int k = 0; k < DPT; ++k) {
      const uint32_t d = tid * DPT + k;
      uint32_t run = 0;
      if (d < (uint32_t)RADIX) {
#pragma unroll
        for (int ww = 0; ww < WAVES; ++ww) {
          const uint32_t x = s_wave_hist[ww * RADIX + d];
          s_wave_hist[ww * RADIX + d] = run;
          run += x;
        }
      }
      cnt_d[k] = run;
      sum += run;
    }
    // publish aggregates early so later tiles can move on
    if (t > 0) {
#pragma unroll
      for (int k = 0; k < DPT; ++k) {
        const uint32_t d = tid * DPT + k;
        if (d < (uint32_t)RADIX) store_status(status + (size_t)t * RADIX + d, kFlagAggregate | cnt_d[k]);
      }
    }
    uint32_t total;
    const uint32_t excl0 = block_exclusive_scan<THREADS>(sum, s_tmp, total);
    {
      uint32_t excl = excl0;
#pragma unroll
      for (int k = 0; k < DPT; ++k) {
        const uint32_t d = tid * DPT + k;
        if (d < (uint32_t)RADIX) s_local[d] = excl;
        excl += cnt_d[k];
      }
    }
    __syncthreads();

    // keys to their place in the tile, sorted by digit
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if (FULL || wave_base + j * kWave < n) {
        const uint32_t d = (key[j] >> shift) & MASK;
        pos[j] += s_local[d] + s_wave_hist[w * RADIX + d];
        if (pos[j] < (uint32_t)TILE) s_buf[pos[j]] = key[j];
      }
    }
    uint32_t val[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t idx = wave_base + j * kWave;
      val[j] = FULL || idx < n ? (vals_in ? vals_in[idx] : idx) : 0u;
    }
    {
      uint32_t excl = excl0;
#pragma unroll
      for (int k = 0; k < DPT; ++k) {
        const uint32_t d = tid * DPT + k;
        if (d < (uint32_t)RADIX) {
          const uint32_t prefix = t == 0 ? 0u : lookback_window<16>(status + d, t, RADIX);
          store_status(status + (size_t)t * RADIX + d, kFlagInclusive | (prefix + cnt_d[k]));
          s_base[d] = s_digit_start[d] + prefix - excl;
        }
        excl += cnt_d[k];
      }
    }
    __syncthreads();

    constexpr int PW = BITS <= 8 ? 8 : 16;  // bits per packed digit
    constexpr int PER = 32 / PW;
    uint32_t dig[(ITEMS + PER - 1) / PER];
#pragma unroll
    for (int k = 0; k < (ITEMS + PER - 1) / PER; ++k) dig[k] = 0u;
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const uint32_t i = k * THREADS + tid;
      if (FULL || i < tile_n) {
        const uint32_t kk = s_buf[i];
        const uint32_t d = (kk >> shift) & MASK;
        dig[k / PER] |= d << (PW * (k % PER));
        const uint32_t o = s_base[d] + i;
        // a position past n means the histogram and the keys disagree:
        // report it instead of storing out of bounds
        if (o < n) keys_out[o] = kk;
        else counters[kScatterError] = 1u;
      }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; ++j)
      if ((FULL || wave_base + j * kWave < n) && pos[j] < (uint32_t)TILE) s_buf[pos[j]] = val[j];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const uint32_t i = k * THREADS + tid;
      if (FULL || i < tile_n) {
        const uint32_t d = (dig[k / PER] >> (PW * (k % PER))) & ((1u << PW) - 1u);
        const uint32_t o = s_base[d] + i;
        if (o < n) vals_out[o] = s_buf[i];
      }
    }
    __syncthreads();
  };

  while (true) {
    if (tid == 0) s_tile = atomicAdd(tile_counter, 1u);
    for (int i = tid; i < WAVES * RADIX; i += THREADS) s_wave_hist[i] = 0u;
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= num_tiles) break;
    const uint32_t left = n - t * TILE;
    if (left >= (uint32_t)TILE) tile_pass(Flag<true>{}, t, (uint32_t)TILE);
    else tile_pass(Flag<false>{}, t, left);
  }
  finish_persistent(tile_counter, counters + kDone + a.tile_slot);
}

template <int BITS>
void launch_sort_pass_as(const SortPassLaunch& a, hipStream_t s) {
  const dim3 grid(a.grid_x, a.views);
  if (a.threads == kSortThreadsXL)
    hipLaunchKernelGGL((k_onesweep_pass<BITS, kSortThreadsXL, kSortItems>), grid, dim3(kSortThreadsXL), 0, s, a);
  else if (a.threads == kSortThreadsWide)
    hipLaunchKernelGGL((k_onesweep_pass<BITS, kSortThreadsWide, kSortItems>), grid, dim3(kSortThreadsWide), 0, s, a);
  else
    hipLaunchKernelGGL((k_onesweep_pass<BITS, kSortThreads, kSortItems>), grid, dim3(kSortThreads), 0, s, a);
}

// bits: the digit width of this pass, 5..8 or 10.  A pass over the top
// bits of a key narrower than the other passes' digits takes a narrower
// digit (fewer ballots, a shorter look-back); its histogram is the first
// 2^bits bins of the full-width one, since the bits above are zero.
void launch_sort_pass(int bits, const SortPassLaunch& a, hipStream_t s) {
  switch (bits) {
    case 5: launch_sort_pass_as<5>(a, s); break;
    case 6: launch_sort_pass_as<6>(a, s); break;
    case 7: launch_sort_pass_as<7>(a, s); break;
    case 10: launch_sort_pass_as<10>(a, s); break;
    default: launch_sort_pass_as<8>(a, s); break;
  }
}

}  // namespace spl
