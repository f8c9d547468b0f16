// cull.hip -- frustum cull + depth key + order-preserving stream compaction
// for every view of a batch in one pass over the splat centres, and the
// digit histograms that the radix passes start from.
//
// Per view the cull replaces the CPU fallback's filter_map
// (renderer.rs:372-392) and the key half of radixSortA (shaders.wgsl:303-336):
//  - key: h = VP*(x,y,z,1) summed in mat4_transform's order (utils.rs:68-70),
//    clip = h * (1/h.w) (renderer.rs:382), visible iff |x|<tol, |y|<tol,
//    |z-0.5|<0.5 (renderer.rs:383-385), key = bits(clip.z) (renderer.rs:387);
//    built with -ffp-contract=off so every product and sum rounds as in Rust.
//  - compaction keeps splat-index order: per 64-splat wave a ballot gives the
//    visible mask, mbcnt its intra-wave offsets, a scan over the tile's waves
//    the intra-tile offsets and a decoupled look-back across tiles the tile
//    offset, so entry c of a view's list is its c-th visible splat by index.
//  - each centre is read once for all views; the splat's visibility over the
//    views (one bit per view) and each tile's per-view prefix are kept for
//    the projection, which reads the splat's other attributes once for all
//    the views that see it.
#include "spl_device.h"
#include "spl_internal.h"

namespace spl {

struct __attribute__((aligned(4))) Float3 {
  float x, y, z;
};

template <int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_cull_views(CullLaunch a) {
  constexpr int WAVES = THREADS / kWave;
  constexpr int TILE = THREADS * ITEMS;
  static_assert(ITEMS * WAVES == kWave, "one wave scans a view's (item, wave) counts");
  __shared__ float s_vp[kMaxViews][16];
  __shared__ float s_tol[kMaxViews];
  __shared__ uint32_t s_cnt[kMaxViews][ITEMS * WAVES];
  __shared__ uint32_t s_prefix[kMaxViews];
  __shared__ uint32_t s_tile;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & (kWave - 1), w = tid / kWave;
  const uint32_t nv = a.views;
  for (uint32_t i = tid; i < nv * 16; i += THREADS) s_vp[i / 16][i % 16] = a.uniforms[i / 16].view_projection_matrix[i % 16];
  if (tid < nv) s_tol[tid] = a.uniforms[tid].frustum_culling_tolerance;
  __syncthreads();

  const uint32_t n = a.n;
  const uint32_t num_tiles = (n + TILE - 1) / TILE;
  if (num_tiles == 0 && blockIdx.x == 0 && tid < nv) a.counters[(size_t)tid * kCounterWords + kVisible] = 0u;
  // The consistency flag that the sort's scatter and the binning raise
  // describes one batch: the cull, first in the sequence, lowers it.
  if (blockIdx.x == 0 && tid < nv) a.counters[(size_t)tid * kCounterWords + kScatterError] = 0u;

  // visibility of one centre in view v, and its depth key
  auto classify = [&](uint32_t v, float x, float y, float z, uint32_t& key) -> bool {
    const float* vp = s_vp[v];
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
    key = __float_as_uint(c2);
    const float tol = s_tol[v];
    return fabsf(c0) < tol && fabsf(c1) < tol && fabsf(c2 - 0.5f) < 0.5f;
  };

  while (true) {
    if (tid == 0) s_tile = atomicAdd(&a.counters[kCtrCull], 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= num_tiles) break;

    float x[ITEMS], y[ITEMS], z[ITEMS];
    uint32_t vm[ITEMS];
    // All sixteen centres are requested before any is used: past the end a
    // lane reads centre 0 (n > 0 here) and drops it, so that no load sits in
    // a branch of its own, which would make each wait for the one before.
    // One 12-byte load per centre (a single dwordx3 request per lane).
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t idx = t * TILE + j * THREADS + tid;
      const bool in = idx < n;
      const Float3 c = *reinterpret_cast<const Float3*>(a.centers + 3 * (size_t)(in ? idx : 0u));
      vm[j] = 0u;
      x[j] = in ? c.x : 0.0f;
      y[j] = in ? c.y : 0.0f;
      z[j] = in ? c.z : 0.0f;
    }
    // A: visibility per (view, item), and per view the (item, wave) counts
    for (uint32_t v = 0; v < nv; ++v) {
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const uint32_t idx = t * TILE + j * THREADS + tid;
        uint32_t key;
        const bool vis = idx < n && classify(v, x[j], y[j], z[j], key);
        vm[j] |= (vis ? 1u : 0u) << v;
        const uint64_t b = __ballot(vis);
        if (lane == 0) s_cnt[v][j * WAVES + w] = (uint32_t)__popcll(b);
      }
    }
#pragma unroll
    for (int j = 0; j < ITEMS; ++j) {
      const uint32_t idx = t * TILE + j * THREADS + tid;
      if (idx < n) a.vis[idx] = vm[j];
    }
    __syncthreads();
    // B: per view, exclusive scan of the counts in splat-index order and the
    // tile's look-back; wave w serves views w, w + WAVES, ...
    for (uint32_t v = w; v < nv; v += WAVES) {
      const uint32_t c = s_cnt[v][lane];
      uint32_t incl = c;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y2 = __shfl_up(incl, o, kWave);
        if (lane >= (uint32_t)o) incl += y2;
      }
      s_cnt[v][lane] = incl - c;
      const uint32_t total = __shfl(incl, kWave - 1, kWave);
      const uint32_t prefix = wave_lookback(a.status + v * a.status_stride, t, total);
      if (lane == 0) {
        s_prefix[v] = prefix;
        a.tile_prefix[v * a.tp_stride + t] = prefix;
        if (t == num_tiles - 1) a.counters[(size_t)v * kCounterWords + kVisible] = prefix + total;
      }
    }
    __syncthreads();
    // C: scatter each view's entries
    for (uint32_t v = 0; v < nv; ++v) {
      uint32_t* keys = a.keys + v * a.keys_stride;
      uint32_t* idxs = a.idx + v * a.idx_stride;
      const uint32_t pv = s_prefix[v];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const bool vis = (vm[j] >> v) & 1u;
        const uint64_t b = __ballot(vis);
        if (vis) {
          const uint32_t pos = pv + s_cnt[v][j * WAVES + w] + lanes_below(b);
          idxs[pos] = t * TILE + j * THREADS + tid;
          if (a.write_keys) {
            uint32_t key;
            (void)classify(v, x[j], y[j], z[j], key);
            keys[pos] = key;
          }
        }
      }
    }
    __syncthreads();
  }
  finish_persistent(&a.counters[kCtrCull], &a.counters[kDone + kCtrCull]);
}

// Onesweep upsweep: the digit histograms of every radix pass over a view's
// keys, LDS first, then one global atomic per bin and block.  Depth keys pile
// into few top digits (their float exponents), so each bin is replicated
// REP times across lanes to keep same-address LDS atomics apart.
template <int BITS>
__global__ __launch_bounds__(256) void k_digit_histograms(HistLaunch a) {
  constexpr int RADIX = 1 << BITS;
  constexpr int REP = BITS <= 8 ? 8 : 2;
  __shared__ uint32_t s_hist[kMaxDepthPasses * RADIX * REP];
  const uint32_t v = blockIdx.y;
  if (a.clear) clear_words(a.clear + v * a.clear_stride, a.clear_words);
  const uint32_t* keys = a.keys + v * a.keys_stride;
  uint32_t* hist = a.hist + v * a.hist_stride;
  const uint32_t n = a.counters[(size_t)v * kCounterWords + kVisible];
  const int words = a.passes * RADIX;
  for (int i = threadIdx.x; i < words * REP; i += blockDim.x) s_hist[i] = 0u;
  __syncthreads();
  const uint32_t rep = threadIdx.x % REP;
  // eight keys per lane are requested before any is counted (past the end
  // a lane reads key 0 and drops it), so the loads do not wait in turn
  constexpr uint32_t U = 8;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i0 = blockIdx.x * blockDim.x + threadIdx.x; i0 < n; i0 += U * stride) {
    uint32_t k[U];
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      const uint32_t i = i0 + u * stride;
      k[u] = keys[i < n ? i : 0u];
    }
#pragma unroll
    for (uint32_t u = 0; u < U; ++u) {
      if (i0 + u * stride >= n) break;
      for (int p = 0; p < a.passes; ++p)
        atomicAdd(&s_hist[(p * RADIX + ((k[u] >> (p * BITS)) & (RADIX - 1))) * REP + rep], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < words; i += blockDim.x) {
    uint32_t sum = 0;
#pragma unroll
    for (int r = 0; r < REP; ++r) sum += s_hist[i * REP + r];
    if (sum) atomicAdd(&hist[i], sum);
  }
}

void launch_cull(const CullLaunch& a, hipStream_t s) {
  hipLaunchKernelGGL((k_cull_views<kCullThreads, kCullItems>), dim3(a.grid), dim3(kCullThreads), 0, s, a);
}

void launch_histograms(const HistLaunch& a, hipStream_t s) {
  if (a.bits == 10)
    hipLaunchKernelGGL((k_digit_histograms<10>), dim3(a.grid_x, a.views), dim3(256), 0, s, a);
  else
    hipLaunchKernelGGL((k_digit_histograms<8>), dim3(a.grid_x, a.views), dim3(256), 0, s, a);
}

}  // namespace spl
