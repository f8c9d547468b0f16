// spl_device.h -- device helpers shared by the splat pipeline kernels
// (gfx950: wave64, 160 KB LDS per CU, 8 XCDs with private L2s).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spl {

constexpr int kWave = 64;

// Lookback status words: 2 flag bits + 30-bit count (shaders.wgsl:438-457
// uses the same split: 0x4... aggregate, 0x8... inclusive prefix).
constexpr uint32_t kFlagAggregate = 0x40000000u;
constexpr uint32_t kFlagInclusive = 0x80000000u;
constexpr uint32_t kFlagMask = 0xC0000000u;
constexpr uint32_t kCountMask = 0x3FFFFFFFu;

// Number of set bits of `mask` held by lanes below this one.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ uint32_t lane_id() { return lanes_below(~0ull); }

// Status words are read and written at agent scope so that the per-XCD L2s
// do not hide one tile's publication from a tile running on another XCD.
__device__ __forceinline__ uint32_t load_status(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_status(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Look-back for one digit counter that inspects WIN predecessors per step
// (their loads are issued together), so the wait for an inclusive prefix
// that is t tiles away costs t/WIN memory round trips instead of t.
// Publishes nothing: the caller publishes the aggregate beforehand and the
// inclusive prefix (returned + count) afterwards.
template <int WIN>
__device__ __forceinline__ uint32_t lookback_window(const uint32_t* status, uint32_t tile, size_t stride) {
  static_assert(WIN <= 32, "window masks are 32 bits");
  uint32_t prefix = 0;
  int64_t p = (int64_t)tile - 1;
  while (true) {
    uint32_t s[WIN];
    uint32_t inv = 0, inc = 0;
#pragma unroll
    for (int i = 0; i < WIN; ++i) {
      s[i] = (p - i >= 0) ? load_status(status + (size_t)(p - i) * stride) : kFlagInclusive;
      inv |= ((s[i] & kFlagMask) == 0 ? 1u : 0u) << i;
      inc |= ((s[i] & kFlagInclusive) != 0 ? 1u : 0u) << i;
    }
    const uint32_t first_inv = inv ? (uint32_t)__builtin_ctz(inv) : (uint32_t)WIN;
    const uint32_t first_inc = inc ? (uint32_t)__builtin_ctz(inc) : (uint32_t)WIN;
    const bool done = first_inc < first_inv;
    const uint32_t take = done ? first_inc + 1 : first_inv;  // words summed this round
#pragma unroll
    for (int i = 0; i < WIN; ++i) prefix += (uint32_t)i < take ? (s[i] & kCountMask) : 0u;
    if (done) break;
    p -= take;
    if (take < (uint32_t)WIN) __builtin_amdgcn_s_sleep(1);
  }
  return prefix;
}

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

// Look-back for a single counter per tile run by one whole wave: the 64
// lanes read the 64 nearest predecessors at once and a ballot finds the
// nearest inclusive prefix (or the first not yet published).  Returns the
// exclusive prefix in every lane; lane 0 publishes aggregate and inclusive.
__device__ __forceinline__ uint32_t wave_lookback(uint32_t* status, uint32_t tile, uint32_t count) {
  const uint32_t lane = lane_id();
  if (tile == 0) {
    if (lane == 0) store_status(status, kFlagInclusive | count);
    return 0;
  }
  if (lane == 0) store_status(status + tile, kFlagAggregate | count);
  uint32_t prefix = 0;
  int64_t base = (int64_t)tile - 1;
  while (true) {
    const int64_t i = base - (int64_t)lane;
    const uint32_t s = i >= 0 ? load_status(status + i) : kFlagInclusive;
    const uint64_t inv = __ballot((s & kFlagMask) == 0);
    const uint64_t inc = __ballot((s & kFlagInclusive) != 0);
    const uint32_t first_inc = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;
    const uint32_t first_inv = inv ? (uint32_t)__builtin_ctzll(inv) : 64u;
    if (first_inc < first_inv) {
      prefix += wave_sum(lane <= first_inc ? (s & kCountMask) : 0u);
      break;
    }
    prefix += wave_sum(lane < first_inv ? (s & kCountMask) : 0u);
    base -= first_inv;
    if (first_inv < 64u) __builtin_amdgcn_s_sleep(1);
  }
  if (lane == 0) store_status(status + tile, kFlagInclusive | (prefix + count));
  return prefix;
}

// 64-bit variant for counts that may pass 2^30 (screen-tile pairs): flags in
// the top two bits, 62-bit count.
constexpr uint64_t kFlagAggregate64 = 1ull << 62;
constexpr uint64_t kFlagInclusive64 = 1ull << 63;
constexpr uint64_t kCountMask64 = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t wave_lookback64(uint64_t* status, uint32_t tile, uint64_t count) {
  const uint32_t lane = lane_id();
  auto ld = [](const uint64_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
  auto st = [](uint64_t* p, uint64_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); };
  if (tile == 0) {
    if (lane == 0) st(status, kFlagInclusive64 | count);
    return 0;
  }
  if (lane == 0) st(status + tile, kFlagAggregate64 | count);
  uint64_t prefix = 0;
  int64_t base = (int64_t)tile - 1;
  while (true) {
    const int64_t i = base - (int64_t)lane;
    const uint64_t s = i >= 0 ? ld(status + i) : kFlagInclusive64;
    const uint64_t inv = __ballot((s >> 62) == 0);
    const uint64_t inc = __ballot((s & kFlagInclusive64) != 0);
    const uint32_t first_inc = inc ? (uint32_t)__builtin_ctzll(inc) : 64u;
    const uint32_t first_inv = inv ? (uint32_t)__builtin_ctzll(inv) : 64u;
    if (first_inc < first_inv) {
      prefix += wave_sum64(lane <= first_inc ? (s & kCountMask64) : 0ull);
      break;
    }
    prefix += wave_sum64(lane < first_inv ? (s & kCountMask64) : 0ull);
    base -= first_inv;
    if (first_inv < 64u) __builtin_amdgcn_s_sleep(1);
  }
  if (lane == 0) st(status + tile, kFlagInclusive64 | (prefix + count));
  return prefix;
}

// Exclusive scan of one value per thread over a block of THREADS threads.
// s_tmp holds THREADS/64 words.  Contains two barriers.
template <int THREADS>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_tmp, uint32_t& total) {
  constexpr int WAVES = THREADS / kWave;
  const uint32_t lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  uint32_t x = v;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    uint32_t y = __shfl_up(x, o, kWave);
    if (lane >= (uint32_t)o) x += y;
  }
  if (lane == kWave - 1) s_tmp[w] = x;
  __syncthreads();
  uint32_t before = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < WAVES; ++i) {
    uint32_t t = s_tmp[i];
    before += (i < (int)w) ? t : 0u;
    tot += t;
  }
  total = tot;
  __syncthreads();
  return before + x - v;
}

// Persistent-kernel bookkeeping: the last block to leave resets the tile
// counter (and itself) for the next launch of the same stage.
__device__ __forceinline__ void finish_persistent(uint32_t* tile_counter, uint32_t* done_counter) {
  if (threadIdx.x == 0) {
    __threadfence();
    uint32_t prev = atomicAdd(done_counter, 1u);
    if (prev == gridDim.x - 1) {
      atomicExch(tile_counter, 0u);
      atomicExch(done_counter, 0u);
    }
  }
}

// Grid-stride zeroing of another stage's finished bookkeeping.
__device__ __forceinline__ void clear_words(uint32_t* p, size_t n) {
  if (!p) return;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0u;
}

}  // namespace spl
