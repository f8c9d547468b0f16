// cull.hip -- frustum cull + depth key + order-preserving stream compaction.
//
// One pass over the splat centres replaces the CPU fallback's filter_map
// (renderer.rs:372-392) and the key half of radixSortA (shaders.wgsl:303-336):
//  - key: h = VP*(x,y,z,1) summed in mat4_transform's order (utils.rs:68-70),
//    clip = h * (1/h.w) (renderer.rs:382), visible iff |x|<tol, |y|<tol,
//    |z-0.5|<0.5 (renderer.rs:383-385), key = bits(clip.z) (renderer.rs:387);
//    built with -ffp-contract=off so every product and sum rounds as in Rust.
//  - compaction keeps splat-index order: per 64-splat wave a ballot gives the
//    visible mask, mbcnt its intra-wave offsets, a scan over the tile's waves
//    the intra-tile offsets and a decoupled look-back across tiles the tile
//    offset, so entry c of the output is the c-th visible splat by index.
//  - the digit histograms of every radix pass over the visible keys are
//    gathered on the way (the onesweep upsweep), LDS first, then global.
#include "spl_device.h"
#include "spl_internal.h"

namespace spl {

template <int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_cull_key_compact(CullLaunch a) {
  constexpr int WAVES = THREADS / kWave;
  constexpr int TILE = THREADS * ITEMS;
  static_assert(ITEMS * WAVES <= kWave, "one wave scans the tile's wave counts");
  __shared__ uint32_t s_hist[3 * 1024];
  __shared__ uint32_t s_cnt[ITEMS * WAVES];
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_prefix;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (kWave - 1), w = tid / kWave;
  const int radix = 1 << a.bits;
  const int hist_words = a.passes * radix;
  const uint32_t digit_mask = (uint32_t)radix - 1u;
  for (int i = tid; i < hist_words; i += THREADS) s_hist[i] = 0u;

  const uint32_t n = a.n;
  const uint32_t num_tiles = (n + TILE - 1) / TILE;
  if (num_tiles == 0 && blockIdx.x == 0 && tid == 0) a.counters[kVisible] = 0u;

  const float* vp = a.vp;
  while (true) {
    if (tid == 0) s_tile = atomicAdd(&a.counters[kCtrCull], 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= num_tiles) break;

    uint64_t vis[ITEMS];
    uint32_t key[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t idx = t * TILE + j * THREADS + tid;
      bool visible = false;
      key[j] = 0u;
      if (idx < n) {
        const float x = a.centers[3 * (size_t)idx + 0];
        const float y = a.centers[3 * (size_t)idx + 1];
        const float z = a.centers[3 * (size_t)idx + 2];
        float h[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) h[r] = ((vp[r] * x + vp[4 + r] * y) + vp[8 + r] * z) + vp[12 + r] * 1.0f;
        float c0, c1, c2;
        if (a.gpu_predicate) {
          const float wv = h[3] + 0.0000001f;
          c0 = h[0] / wv; c1 = h[1] / wv; c2 = h[2] / wv;
        } else {
          const float rcp = 1.0f / h[3];
          c0 = h[0] * rcp; c1 = h[1] * rcp; c2 = h[2] * rcp;
        }
        visible = fabsf(c0) < a.tol && fabsf(c1) < a.tol && fabsf(c2 - 0.5f) < 0.5f;
        key[j] = __float_as_uint(c2);
      }
      vis[j] = __ballot(visible);
      if (lane == 0) s_cnt[j * WAVES + w] = (uint32_t)__popcll(vis[j]);
    }
    __syncthreads();
    if (w == 0) {
      // exclusive scan of the (item, wave) counts in splat-index order
      const uint32_t v = lane < ITEMS * WAVES ? s_cnt[lane] : 0u;
      uint32_t x = v;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, kWave);
        if (lane >= (uint32_t)o) x += y;
      }
      if (lane < ITEMS * WAVES) s_cnt[lane] = x - v;
      const uint32_t total = __shfl(x, kWave - 1, kWave);
      const uint32_t prefix = wave_lookback(a.status, t, total);
      if (lane == 0) {
        s_prefix = prefix;
        if (t == num_tiles - 1) a.counters[kVisible] = prefix + total;
      }
    }
    __syncthreads();
    const uint32_t prefix = s_prefix;
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      if ((vis[j] >> lane) & 1ull) {
        const uint32_t pos = prefix + s_cnt[j * WAVES + w] + lanes_below(vis[j]);
        if (a.keys_out) a.keys_out[pos] = key[j];
        a.idx_out[pos] = t * TILE + j * THREADS + tid;
        for (int p = 0; p < a.passes; ++p)
          atomicAdd(&s_hist[p * radix + ((key[j] >> (p * a.bits)) & digit_mask)], 1u);
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < hist_words; i += THREADS)
    if (s_hist[i]) atomicAdd(&a.hist[i], s_hist[i]);
  finish_persistent(&a.counters[kCtrCull], &a.counters[kDone + kCtrCull]);
}

void launch_cull(const CullLaunch& a, hipStream_t s) {
  hipLaunchKernelGGL((k_cull_key_compact<kCullThreads, kCullItems>), dim3(a.grid), dim3(kCullThreads), 0, s, a);
}

}  // namespace spl
