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

// Decoupled look-back for one counter of tile `tile`: publishes the
// aggregate, walks back to the nearest inclusive prefix, publishes the
// inclusive prefix and returns the exclusive prefix.  `stride` is the
// distance between consecutive tiles' words.  Tiles are taken in order from
// an atomic counter by resident blocks, so every predecessor is running or
// done and the wait ends.
__device__ __forceinline__ uint32_t lookback(uint32_t* status, uint32_t tile, size_t stride, uint32_t count,
                                             bool publish_aggregate = true) {
  if (tile == 0) {
    store_status(status, kFlagInclusive | count);
    return 0;
  }
  uint32_t* mine = status + (size_t)tile * stride;
  if (publish_aggregate) store_status(mine, kFlagAggregate | count);
  uint32_t prefix = 0;
  int64_t prev = (int64_t)tile - 1;
  while (true) {
    uint32_t s = load_status(status + (size_t)prev * stride);
    if ((s & kFlagMask) == 0) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    prefix += s & kCountMask;
    if (s & kFlagInclusive) break;
    --prev;
  }
  store_status(mine, kFlagInclusive | (prefix + count));
  return prefix;
}

// 64-bit variant for counts that may pass 2^30 (screen-tile pairs): flags in
// the top two bits, 62-bit count.
constexpr uint64_t kFlagAggregate64 = 1ull << 62;
constexpr uint64_t kFlagInclusive64 = 1ull << 63;
constexpr uint64_t kCountMask64 = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t lookback64(uint64_t* status, uint32_t tile, uint64_t count) {
  if (tile == 0) {
    __hip_atomic_store(status, kFlagInclusive64 | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
  }
  __hip_atomic_store(status + tile, kFlagAggregate64 | count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t prefix = 0;
  int64_t prev = (int64_t)tile - 1;
  while (true) {
    uint64_t s = __hip_atomic_load(status + prev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if ((s >> 62) == 0) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    prefix += s & kCountMask64;
    if (s & kFlagInclusive64) break;
    --prev;
  }
  __hip_atomic_store(status + tile, kFlagInclusive64 | (prefix + count), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
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
