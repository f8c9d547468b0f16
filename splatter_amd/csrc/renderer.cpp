// renderer.cpp -- Renderer (renderer.rs:64-467): workspace, the per-frame host
// math of renderer.rs:364-409, and the launch sequence of the splat pipeline
// for a batch of views, replayed from a hipGraph per batch shape.  Built with
// -ffp-contract=off so the host-side f32 camera math rounds as the
// reference's Rust does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "spl_internal.h"
#include "spl_objects.h"
#include "splatter.h"

using namespace spl;

namespace {

constexpr int kStageSlots = 4;  // pinned uniform uploads in flight
constexpr uint32_t kSingleHeavyMult = 2;
constexpr int kStageCount = 8;  // cull, histograms, sort, project, bin, pair sort, ranges, raster

struct UploadSlot {
  Uniforms* uniforms = nullptr;  // pinned host
  uint32_t** pixels = nullptr;   // pinned host
  hipEvent_t done = nullptr;
};

// Everything a captured launch sequence bakes into its kernel arguments.
struct BatchShape {
  const spl_scene* scene = nullptr;
  uint64_t scene_uid = 0;  // the scene's process-unique id: its address may be reused after it is destroyed
  uint64_t n = 0;
  uint32_t views = 0;
  uint32_t max_tiles = 0;
  int pair_passes = 0;
  bool conics = false;
  bool tile_stats = false;
  bool operator==(const BatchShape& o) const {
    return scene_uid == o.scene_uid && n == o.n && views == o.views && max_tiles == o.max_tiles &&
           pair_passes == o.pair_passes && conics == o.conics && tile_stats == o.tile_stats;
  }
};

// Instantiated launch sequences, one per batch shape seen recently.
struct GraphEntry {
  BatchShape shape;
  hipGraphExec_t exec = nullptr;
  uint64_t used = 0;
};
constexpr size_t kGraphCache = 16;

int pair_passes_for(uint32_t tiles) {
  if (tiles <= 256) return 1;
  if (tiles <= 65536) return 2;
  return 3;
}

}  // namespace

struct spl_renderer {
  spl_device* dev = nullptr;
  spl_configuration cfg{};
  uint64_t cap = 0;
  uint32_t max_views = 1;
  int depth_bits = 8, depth_passes = 4;
  int sort_threads = kSortThreads;  // lanes per radix-sort tile
  uint64_t pair_cap = 0;
  // per-view sizes (words)
  size_t cull_tiles = 0, sort_words = 0, bin_words = 0, pair_words = 0;
  // per-view buffers, view v at base + v * stride
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* compact_idx = nullptr;
  float4* records = nullptr;
  uint2* boxes = nullptr;
  uint32_t* pair_keys[2] = {nullptr, nullptr};
  uint32_t* pair_vals[2] = {nullptr, nullptr};
  uint32_t* cull_status = nullptr;
  uint32_t* tile_prefix = nullptr;
  uint32_t* sort_status = nullptr;
  uint64_t* bin_status = nullptr;
  uint32_t* pair_status = nullptr;
  uint32_t* counters = nullptr;
  uint32_t* depth_hist = nullptr;
  uint32_t* pair_hist = nullptr;
  uint2* ranges = nullptr;
  uint32_t* tile_order = nullptr;  // per view, ranges_cap tiles
  uint32_t ranges_cap = 0;
  // shared
  uint32_t* vis = nullptr;
  Uniforms* d_uniforms = nullptr;
  uint32_t** d_pixels = nullptr;
  UploadSlot slots[kStageSlots];
  int next_slot = 0;
  double* conics = nullptr;
  bool conics_on = false;
  uint2* tile_work = nullptr;
  uint32_t tile_work_cap = 0;
  bool tile_stats_on = false;
  // last batch
  uint32_t last_views = 0;
  bool last_sorted = false;
  int last_final = 0;
  bool have_frame = false;
  uint32_t debug_view = 0;
  uint32_t last_tiles[kMaxViews] = {};
  // stage timing: one set of events per batch, read out in bulk
  bool timing = false;
  std::vector<std::array<hipEvent_t, kStageCount + 1>> ev_pool;
  std::vector<uint32_t> ev_views;
  size_t ev_used = 0;
  double stage_sum_ms[kStageCount] = {};
  uint64_t stage_frames = 0;
  // replay of the launch sequence
  std::vector<GraphEntry> graphs;
  uint64_t graph_clock = 0;
  // the stream every launch and copy of this renderer goes to: the device's,
  // or one of its own (a server slot)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  // the projection's branch of a sorted batch: it needs only the cull, so
  // it runs beside the depth sort and the binning waits for both
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  uint64_t bytes = 0;  // device memory held
};

namespace spl {
spl_device* renderer_device(const spl_renderer* r) { return r->dev; }
uint64_t renderer_capacity(const spl_renderer* r) { return r->cap; }
}  // namespace spl

namespace {

void drop_graph(spl_renderer* r) {
  for (auto& g : r->graphs) (void)hipGraphExecDestroy(g.exec);
  r->graphs.clear();
}

void renderer_free(spl_renderer* r) {
  drop_graph(r);
  for (int i = 0; i < 2; ++i) {
    dev_free(r->keys[i]);
    dev_free(r->vals[i]);
    dev_free(r->pair_keys[i]);
    dev_free(r->pair_vals[i]);
  }
  dev_free(r->compact_idx);
  dev_free(r->records);
  dev_free(r->boxes);
  dev_free(r->cull_status);
  dev_free(r->tile_prefix);
  dev_free(r->sort_status);
  dev_free(r->bin_status);
  dev_free(r->pair_status);
  dev_free(r->counters);
  dev_free(r->depth_hist);
  dev_free(r->pair_hist);
  dev_free(r->ranges);
  dev_free(r->tile_order);
  dev_free(r->vis);
  dev_free(r->d_uniforms);
  dev_free(r->d_pixels);
  dev_free(r->conics);
  dev_free(r->tile_work);
  for (auto& s : r->slots) {
    if (s.uniforms) (void)hipHostFree(s.uniforms);
    if (s.pixels) (void)hipHostFree(s.pixels);
    if (s.done) (void)hipEventDestroy(s.done);
    s = UploadSlot{};
  }
  for (auto& set : r->ev_pool)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e), e = nullptr;
  r->ev_pool.clear();
  r->ev_views.clear();
  r->ev_used = 0;
  if (r->fork) (void)hipEventDestroy(r->fork);
  if (r->join) (void)hipEventDestroy(r->join);
  if (r->side) {
    (void)hipStreamSynchronize(r->side);
    release_stream(r->dev, r->side);
  }
  r->fork = r->join = nullptr;
  r->side = nullptr;
  if (r->own_stream && r->stream) {
    (void)hipStreamSynchronize(r->stream);
    release_stream(r->dev, r->stream);
  }
  r->stream = nullptr;
  r->own_stream = false;
}

// utils.rs:58-65
void mat4_multiplication(const float a[16], const float b[16], float out[16]) {
  float t[16];
  for (int j = 0; j < 4; ++j)
    for (int r = 0; r < 4; ++r)
      t[4 * j + r] = ((a[r] * b[4 * j + 0] + a[4 + r] * b[4 * j + 1]) + a[8 + r] * b[4 * j + 2]) + a[12 + r] * b[4 * j + 3];
  std::memcpy(out, t, sizeof(t));
}

// renderer.rs:364-370,397-409 and utils.rs:45-55.  field_of_view_y replaces
// the hard-coded PI/2 (renderer.rs:366) and view_width = aspect*view_height
// replaces aspect/view_height (renderer.rs:368): at 90 degrees tanf gives
// exactly 1.0f and both are bit-identical.
void frame_uniforms(const spl_configuration& cfg, const float cam[16], const float view[16], uint32_t W, uint32_t H,
                    Uniforms& u) {
  const float view_height = tanf(cfg.field_of_view_y * 0.5f);
  const float view_width = ((float)W / (float)H) * view_height;
  float P[16] = {};
  const float near_ = 1.0f, far_ = 1000.0f;
  const float denominator = 1.0f / (near_ - far_);
  P[0] = 1.0f / view_width;
  P[5] = 1.0f / view_height;
  P[10] = -far_ * denominator;
  P[11] = 1.0f;
  P[14] = (near_ * far_) * denominator;
  std::memcpy(u.camera_matrix, cam, 64);
  std::memcpy(u.view_matrix, view, 64);
  mat4_multiplication(P, view, u.view_projection_matrix);
  u.view_size[0] = view_width;
  u.view_size[1] = view_height;
  u.image_size[0] = W;
  u.image_size[1] = H;
  u.frustum_culling_tolerance = cfg.frustum_culling_tolerance;
  u.ellipse_size_bias = (0.2f * view_width) / (float)W;
  u.ellipse_margin = cfg.ellipse_margin;
  u.splat_scale = cfg.splat_scale;
  u.tiles_x = div_up(W, kTileSize);
  u.tiles_y = div_up(H, kTileSize);
}

int ensure_ranges(spl_renderer* r, uint32_t tiles) {
  if (tiles <= r->ranges_cap) return SPL_OK;
  HIP_TRY(hipStreamSynchronize(r->stream));
  drop_graph(r);
  dev_free(r->ranges);
  dev_free(r->tile_order);
  r->bytes -= (uint64_t)r->ranges_cap * r->max_views * (sizeof(uint2) + sizeof(uint32_t));
  r->ranges_cap = 0;
  const size_t words = (size_t)tiles * r->max_views;
  HIP_TRY(dev_alloc(&r->ranges, words));
  HIP_TRY(dev_alloc(&r->tile_order, words));
  HIP_TRY(hipMemsetAsync(r->ranges, 0, words * sizeof(uint2), r->stream));
  HIP_TRY(hipStreamSynchronize(r->stream));
  r->ranges_cap = tiles;
  r->bytes += (uint64_t)words * (sizeof(uint2) + sizeof(uint32_t));
  return SPL_OK;
}

int ensure_tile_work(spl_renderer* r, uint32_t tiles) {
  if (!r->tile_stats_on || r->tile_work_cap >= tiles) return SPL_OK;
  HIP_TRY(hipStreamSynchronize(r->stream));
  drop_graph(r);
  dev_free(r->tile_work);
  r->tile_work_cap = 0;
  HIP_TRY(dev_alloc(&r->tile_work, tiles));
  r->tile_work_cap = tiles;
  return SPL_OK;
}

// The whole pipeline for one batch, as a sequence of launches on `q`.
// Every stage clears the bookkeeping of an earlier stage once that stage is
// done (status words, histograms, tile ranges), so the sequence needs no
// memset and replays as a graph.
void launch_sequence(spl_renderer* r, const BatchShape& b, hipStream_t q, hipEvent_t* ev) {
  const spl_configuration& cfg = r->cfg;
  const spl_scene* s = b.scene;
  const uint32_t n = (uint32_t)b.n, nv = b.views;
  const bool sorted = cfg.depth_sorting != SPL_DEPTH_SORTING_NONE;
  const int P = sorted ? r->depth_passes : 0;
  const int bits = r->depth_bits;
  const int radix = 1 << bits;
  const int Q = b.pair_passes;
  int tile_bits = 1;  // bits of the largest tile key, max_tiles - 1
  while (tile_bits < 31 && (b.max_tiles - 1u) >> tile_bits) ++tile_bits;
  const uint32_t cus = (uint32_t)r->dev->cu_count;
  const uint32_t per_view = std::max<uint32_t>(1, cus * 4 / nv);
  // radix-sort tiles: one 1024-lane, two 512-lane or four 256-lane blocks per CU, shared
  // by the views (which sort side by side, so their look-back chains overlap)
  // A single view sorts in 16384-key tiles whatever the batches use: with
  // one view on the device the look-back chains set the sort's time, and
  // larger tiles shorten them (the autotuner picks the batches' tile).
  const int sort_threads = nv == 1 ? kSortThreadsXL : r->sort_threads;
  const uint32_t sort_tile = (uint32_t)sort_threads * kSortItems;
  const uint32_t sort_per_view = std::max<uint32_t>(1, cus * 4 * kSortThreads / sort_threads / nv);
  const size_t cap = r->cap;
  const size_t sort_stride = r->sort_words * kMaxDepthPasses;
  const size_t pair_stride_status = r->pair_words * kMaxPairPasses;
  const size_t sort_used = (size_t)div_up(n, kSortThreads * kSortItems) * radix;
  const size_t cull_used = div_up(n, kCullTile);
  auto mark = [&](int i) { if (ev) (void)hipEventRecord(ev[i], q); };
  // A sorted batch projects on the side stream while the depth sort runs
  // (the timing pass keeps one stream, so each stage's events bracket it
  // alone).
  const bool fork = sorted && !ev && r->side != nullptr;
  auto clampg = [](uint64_t want, uint32_t hi) { return (int)std::max<uint64_t>(1, std::min<uint64_t>(want, hi)); };

  mark(0);
  {
    CullLaunch a{};
    a.centers = s->centers;
    a.n = n;
    a.views = nv;
    a.uniforms = r->d_uniforms;
    a.gpu_predicate = sorted ? 0 : 1;
    a.write_keys = sorted ? 1 : 0;
    a.keys = r->keys[0]; a.keys_stride = cap;
    a.idx = r->compact_idx; a.idx_stride = cap;
    a.vis = r->vis;
    a.tile_prefix = r->tile_prefix; a.tp_stride = r->cull_tiles;
    a.status = r->cull_status; a.status_stride = r->cull_tiles;
    a.counters = r->counters;
    a.grid = clampg(div_up(n, kCullTile), cus * 4);
    launch_cull(a, q);
  }
  auto launch_proj = [&](hipStream_t qq) {
    ProjectLaunch a{};
    a.scene = SceneView{s->centers, s->rotations, s->scale_alpha, s->sh, n};
    a.views = nv;
    a.uniforms = r->d_uniforms;
    a.vis = r->vis;
    a.tile_prefix = r->tile_prefix; a.tp_stride = r->cull_tiles;
    a.sh_order = (int)cfg.spherical_harmonics_order;
    a.use_covariance_for_scale = cfg.use_covariance_for_scale ? 1 : 0;
    a.use_unaligned_rectangles = cfg.use_unaligned_rectangles ? 1 : 0;
    a.records = r->records; a.rec_stride = cap * kRecordFloat4;
    a.boxes = r->boxes; a.box_stride = cap;
    a.conics = b.conics ? r->conics : nullptr;
    if (P == 0) {  // (sorted, the histograms clear the cull's words and the binning the sort's)
      a.clear_a = r->cull_status;
      a.clear_a_words = cull_used;
      a.clear_a_stride = r->cull_tiles;
    }
    a.grid = clampg(div_up(n, kCullTile), cus * 3);  // three resident blocks per CU
    launch_project(a, qq);
  };
  mark(1);
  if (fork) {
    (void)hipEventRecord(r->fork, q);
    (void)hipStreamWaitEvent(r->side, r->fork, 0);
    launch_proj(r->side);
    (void)hipEventRecord(r->join, r->side);
  }
  if (sorted) {
    HistLaunch a{};
    a.keys = r->keys[0]; a.keys_stride = cap;
    a.hist = r->depth_hist; a.hist_stride = (size_t)kMaxDepthPasses * 1024;
    a.counters = r->counters;
    a.passes = P;
    a.bits = bits;
    a.clear = r->cull_status; a.clear_words = cull_used; a.clear_stride = r->cull_tiles;
    a.grid_x = clampg(div_up(n, 256 * 64), per_view);  // 64+ keys a lane: fewer blocks flushing bins
    a.views = (int)nv;
    launch_histograms(a, q);
  }
  mark(2);
  for (int p = 0; p < P; ++p) {
    SortPassLaunch a{};
    a.keys_in = r->keys[p & 1];
    a.vals_in = p == 0 ? nullptr : r->vals[p & 1];
    a.keys_out = r->keys[(p + 1) & 1];
    a.vals_out = r->vals[(p + 1) & 1];
    a.entry_stride = cap;
    a.counters = r->counters;
    a.count_slot = kVisible;
    a.count_cap = n;
    a.shift = p * bits;
    a.hist = r->depth_hist + (size_t)p * radix; a.hist_stride = (size_t)kMaxDepthPasses * 1024;
    a.status = r->sort_status + (size_t)p * r->sort_words; a.status_stride = sort_stride;
    a.tile_slot = kCtrSort0 + p;
    if (p > 0) {
      a.clear_a = r->sort_status + (size_t)(p - 1) * r->sort_words;
      a.clear_a_words = sort_used;
      a.clear_a_stride = sort_stride;
    }
    a.threads = sort_threads;
    a.grid_x = clampg(div_up(n, sort_tile), sort_per_view);
    a.views = (int)nv;
    // keys are below 2^30 (0 < clip.z < 1): the last 8-bit pass needs 6 bits
    launch_sort_pass(std::min(bits, 30 - p * bits), a, q);
  }
  const int fin = P & 1;
  mark(3);
  if (!fork) launch_proj(q);
  mark(4);
  if (fork) (void)hipStreamWaitEvent(q, r->join, 0);
  {
    BinLaunch a{};
    if (P > 0) {  // the depth sort's last status words and its histograms
      a.clear_a = r->sort_status + (size_t)(P - 1) * r->sort_words;
      a.clear_a_words = sort_used;
      a.clear_a_stride = sort_stride;
      a.clear_b = r->depth_hist;
      a.clear_b_words = (size_t)kMaxDepthPasses * 1024;
      a.clear_b_stride = (size_t)kMaxDepthPasses * 1024;
    }
    a.order = sorted ? r->vals[fin] : nullptr; a.order_stride = cap;
    a.boxes = r->boxes; a.box_stride = cap;
    a.uniforms = r->d_uniforms;
    a.pair_keys = r->pair_keys[0];
    a.pair_vals = r->pair_vals[0]; a.pair_stride = r->pair_cap;
    a.pair_cap = (uint32_t)r->pair_cap;
    a.hist = r->pair_hist; a.hist_stride = (size_t)kMaxPairPasses * 256;
    a.pair_passes = Q;
    a.status = r->bin_status; a.status_stride = r->bin_words;
    a.counters = r->counters;
    a.grid_x = clampg(div_up(n, kBinThreads * kBinItems), per_view);
    a.views = (int)nv;
    launch_bin(a, q);
  }
  mark(5);
  for (int p = 0; p < Q; ++p) {
    SortPassLaunch a{};
    a.keys_in = r->pair_keys[p & 1];
    a.vals_in = r->pair_vals[p & 1];
    a.keys_out = r->pair_keys[(p + 1) & 1];
    a.vals_out = r->pair_vals[(p + 1) & 1];
    a.entry_stride = r->pair_cap;
    a.counters = r->counters;
    a.count_slot = kPairs;
    a.count_cap = (uint32_t)r->pair_cap;
    a.shift = 8 * p;
    a.hist = r->pair_hist + (size_t)p * 256; a.hist_stride = (size_t)kMaxPairPasses * 256;
    a.status = r->pair_status + (size_t)p * r->pair_words; a.status_stride = pair_stride_status;
    a.tile_slot = kCtrPair0 + p;
    if (p == 0) {
      a.clear_a = reinterpret_cast<uint32_t*>(r->bin_status);
      a.clear_a_words = (size_t)div_up(n, kBinThreads * kBinItems) * 2;
      a.clear_a_stride = r->bin_words * 2;
      // the tile ranges the previous frame's raster read: tiles without
      // pairs in this frame must read as empty
      a.clear_b = reinterpret_cast<uint32_t*>(r->ranges);
      a.clear_b_words = (size_t)b.max_tiles * 2;
      a.clear_b_stride = (size_t)r->ranges_cap * 2;
    } else {
      a.clear_a = r->pair_status + (size_t)(p - 1) * r->pair_words;
      a.clear_a_words = r->pair_words;
      a.clear_a_stride = pair_stride_status;
    }
    a.threads = sort_threads;
    a.grid_x = clampg(div_up(r->pair_cap, sort_tile), sort_per_view);
    a.views = (int)nv;
    // tile keys are below max_tiles: the last pass takes only their top bits
    launch_sort_pass(std::max(5, std::min(8, tile_bits - 8 * p)), a, q);
  }
  const int pfin = Q & 1;
  mark(6);
  {
    RangesLaunch a{};
    a.pair_keys = r->pair_keys[pfin]; a.pair_stride = r->pair_cap;
    a.counters = r->counters;
    a.pair_cap = (uint32_t)r->pair_cap;
    a.uniforms = r->d_uniforms;
    a.ranges = r->ranges; a.ranges_stride = r->ranges_cap;
    a.clear_a = r->pair_status + (size_t)(Q - 1) * r->pair_words;
    a.clear_a_words = r->pair_words;
    a.clear_a_stride = pair_stride_status;
    a.clear_b = r->pair_hist;
    a.clear_b_words = (size_t)kMaxPairPasses * 256;
    a.clear_b_stride = (size_t)kMaxPairPasses * 256;
    a.grid_x = clampg(div_up(r->pair_cap, 1024), std::max<uint32_t>(1, cus * 32 / nv));
    a.views = (int)nv;
    launch_ranges(a, q);
  }
  const uint32_t slots = ((b.max_tiles + 7) / 8) * 8;
  const uint32_t heavy_cap = std::max<uint32_t>(8, slots / 8);
  // A single view rasters in 256-entry batches, its longest tiles first:
  // with only one view's tiles on the device the longest ones set the
  // frame's time (660.6 against 614.8 plain frames/s, 6M splats at 1080p on
  // MI355X).  Batches of views keep the configured form, which suits
  // throughput (the autotuner picks it).  Every form gives the same image.
  const bool single = nv == 1;
  const uint32_t raster_batch = single ? 256u : cfg.raster_batch;
  const bool longest_first = single || cfg.raster_longest_first != 0;
  if (longest_first) {
    OrderLaunch a{};
    a.ranges = r->ranges; a.ranges_stride = r->ranges_cap;
    a.uniforms = r->d_uniforms;
    a.counters = r->counters;
    a.pair_cap = (uint32_t)r->pair_cap;
    a.order = r->tile_order; a.order_stride = r->ranges_cap;
    a.heavy_cap = heavy_cap;
    // one view: tiles from twice the mean list go first (its time is set by
    // its longest tiles), those from 8x the mean on four blocks each
    a.heavy_mult = single ? kSingleHeavyMult : 8u;
    a.heavy_min = single ? 512u : 2048u;
    a.split_mult = single && cfg.raster_mode == 3 ? 8u : 0u;
    a.views = (int)nv;
    launch_tile_order(a, q);
  }
  mark(7);
  {
    RasterLaunch a{};
    a.pair_vals = r->pair_vals[pfin]; a.pair_stride = r->pair_cap;
    a.records = r->records; a.rec_stride = cap * kRecordFloat4;
    a.ranges = r->ranges; a.ranges_stride = r->ranges_cap;
    a.uniforms = r->d_uniforms;
    a.pixels = r->d_pixels;
    a.counters = r->counters;
    a.use_unaligned_rectangles = cfg.use_unaligned_rectangles ? 1 : 0;
    a.blend_f32 = cfg.blend_mode == SPL_BLEND_F32 ? 1 : 0;
    a.walk = (int)cfg.raster_mode;
    a.batch_entries = (int)raster_batch;
    a.longest_first = longest_first ? 1 : 0;
    a.bgra = cfg.frame_format == SPL_FRAME_FORMAT_BGRA8_UNORM ? 1 : 0;
    a.tile_work = b.tile_stats ? r->tile_work : nullptr;
    a.order = r->tile_order; a.order_stride = r->ranges_cap;
    a.slots_per_view = slots;
    a.heavy_cap = heavy_cap;
    // one view in mode 3: each heavy tile's four quadrants go to four blocks
    a.split_heavy = single && cfg.raster_mode == 3 ? 1 : 0;
    a.grid_x = (int)(nv * slots + (longest_first ? ((nv * heavy_cap + 7) / 8) * 8 * (a.split_heavy ? 5 : 1) : 0));
    a.views = (int)nv;
    launch_raster(a, q);
  }
  mark(8);
}

// One batch of at most max_views views.
int enqueue_batch(spl_renderer* r, spl_scene* s, uint32_t nv, spl_frame* const* frames, const float* cams,
                  const float* views) {
  hipStream_t q = r->stream;
  Uniforms u[kMaxViews];
  uint32_t max_tiles = 1;
  for (uint32_t v = 0; v < nv; ++v) {
    spl_frame* f = frames[v];
    if (!f) return fail(SPL_ERROR_INVALID_ARGUMENT, "null frame");
    if (f->dev != r->dev) return fail(SPL_ERROR_INVALID_ARGUMENT, "frame and renderer must share a device");
    frame_uniforms(r->cfg, cams + 16 * (size_t)v, views + 16 * (size_t)v, f->width, f->height, u[v]);
    max_tiles = std::max(max_tiles, u[v].tiles_x * u[v].tiles_y);
    r->last_tiles[v] = u[v].tiles_x * u[v].tiles_y;
  }
  int rc = ensure_ranges(r, max_tiles);
  if (rc) return rc;
  rc = ensure_tile_work(r, max_tiles);
  if (rc) return rc;

  // stage this batch's uniforms and frame pointers through pinned memory
  UploadSlot& slot = r->slots[r->next_slot];
  r->next_slot = (r->next_slot + 1) % kStageSlots;
  HIP_TRY(hipEventSynchronize(slot.done));
  std::memcpy(slot.uniforms, u, nv * sizeof(Uniforms));
  for (uint32_t v = 0; v < nv; ++v) slot.pixels[v] = frames[v]->pixels;
  HIP_TRY(hipMemcpyAsync(r->d_uniforms, slot.uniforms, nv * sizeof(Uniforms), hipMemcpyHostToDevice, q));
  HIP_TRY(hipMemcpyAsync(r->d_pixels, slot.pixels, nv * sizeof(uint32_t*), hipMemcpyHostToDevice, q));
  HIP_TRY(hipEventRecord(slot.done, q));

  BatchShape b;
  b.scene = s;
  b.scene_uid = s->uid;
  b.n = s->count;
  b.views = nv;
  b.max_tiles = max_tiles;
  b.pair_passes = pair_passes_for(max_tiles);
  b.conics = r->conics_on;
  b.tile_stats = r->tile_stats_on;

  if (r->timing) {
    if (r->ev_used == r->ev_pool.size()) {
      std::array<hipEvent_t, kStageCount + 1> set{};
      for (auto& e : set) HIP_TRY(hipEventCreate(&e));
      r->ev_pool.push_back(set);
      r->ev_views.push_back(0);
    }
    r->ev_views[r->ev_used] = nv;
    launch_sequence(r, b, q, r->ev_pool[r->ev_used++].data());
  } else if (r->cfg.use_hip_graphs) {
    GraphEntry* hit = nullptr;
    for (auto& g : r->graphs)
      if (g.shape == b) hit = &g;
    if (!hit) {
      if (r->graphs.size() >= kGraphCache) {
        auto lru = std::min_element(r->graphs.begin(), r->graphs.end(),
                                    [](const GraphEntry& x, const GraphEntry& y) { return x.used < y.used; });
        (void)hipGraphExecDestroy(lru->exec);
        r->graphs.erase(lru);
      }
      hipGraph_t g = nullptr;
      HIP_TRY(hipStreamBeginCapture(q, hipStreamCaptureModeThreadLocal));
      launch_sequence(r, b, q, nullptr);
      hipError_t e = hipStreamEndCapture(q, &g);
      if (e != hipSuccess) return fail(SPL_ERROR_DEVICE, std::string("graph capture: ") + hipGetErrorString(e));
      GraphEntry entry;
      entry.shape = b;
      e = hipGraphInstantiate(&entry.exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess) return fail(SPL_ERROR_DEVICE, std::string("graph instantiate: ") + hipGetErrorString(e));
      r->graphs.push_back(entry);
      hit = &r->graphs.back();
    }
    hit->used = ++r->graph_clock;
    HIP_TRY(hipGraphLaunch(hit->exec, q));
  } else {
    launch_sequence(r, b, q, nullptr);
  }
  HIP_TRY(hipGetLastError());
  r->last_views = nv;
  r->last_sorted = r->cfg.depth_sorting != SPL_DEPTH_SORTING_NONE;
  r->last_final = r->last_sorted ? (r->depth_passes & 1) : 0;
  r->have_frame = true;
  if (r->debug_view >= nv) r->debug_view = 0;
  return SPL_OK;
}

int read_counter(spl_renderer* r, uint32_t view, int slot, uint32_t* v) {
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->stream));
  HIP_TRY(hipMemcpy(v, r->counters + (size_t)view * kCounterWords + slot, 4, hipMemcpyDeviceToHost));
  return SPL_OK;
}

// draw order of the selected view: the compacted position of each drawn splat
int draw_order(spl_renderer* r, std::vector<uint32_t>& order, std::vector<uint32_t>* keys) {
  const uint32_t view = r->debug_view;
  uint32_t v = 0, err = 0;
  int rc = read_counter(r, view, kScatterError, &err);
  if (rc) return rc;
  if (err) return fail(SPL_ERROR_DEVICE, "radix sort scatter left its range (inconsistent histogram)");
  rc = read_counter(r, view, kVisible, &v);
  if (rc) return rc;
  order.resize(v);
  if (keys) keys->assign(v, 0u);
  if (v == 0) return SPL_OK;
  const size_t off = (size_t)view * r->cap;
  if (r->last_sorted) {
    HIP_TRY(hipMemcpy(order.data(), r->vals[r->last_final] + off, (size_t)v * 4, hipMemcpyDeviceToHost));
    if (keys) HIP_TRY(hipMemcpy(keys->data(), r->keys[r->last_final] + off, (size_t)v * 4, hipMemcpyDeviceToHost));
  } else {
    for (uint32_t i = 0; i < v; ++i) order[i] = i;
  }
  return SPL_OK;
}

// The checks of spl_renderer_create on a configuration.
int validate_config(const spl_configuration* cfg) {
  if (cfg->spherical_harmonics_order > 3)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spherical_harmonics_order must be 0..=3");
  if (cfg->radix_bits_per_digit != 8 && cfg->radix_bits_per_digit != 10)
    return fail(SPL_ERROR_UNSUPPORTED, "radix_bits_per_digit must be 8 or 10");
  if (cfg->max_splat_count == 0 || cfg->max_splat_count > (1ull << 30))
    return fail(SPL_ERROR_INVALID_ARGUMENT, "max_splat_count must be in 1..=2^30");
  if ((unsigned)cfg->depth_sorting > 3u || (unsigned)cfg->frame_format > 1u || (unsigned)cfg->blend_mode > 1u)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_renderer_create: enum out of range");
  if (!(cfg->field_of_view_y > 0.0f && cfg->field_of_view_y < 3.14159f))
    return fail(SPL_ERROR_INVALID_ARGUMENT, "field_of_view_y must be in (0, pi)");
  if (cfg->max_views > (uint32_t)kMaxViews)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "max_views must be in 1..=32");
  if (cfg->sort_tile_keys != 0 && cfg->sort_tile_keys != 4096 && cfg->sort_tile_keys != 8192 &&
      cfg->sort_tile_keys != 16384)
    return fail(SPL_ERROR_UNSUPPORTED, "sort_tile_keys must be 4096, 8192 or 16384");
  if (cfg->raster_mode > 3) return fail(SPL_ERROR_INVALID_ARGUMENT, "raster_mode must be 0, 1, 2 or 3");
  if (cfg->raster_batch != 0 && cfg->raster_batch != 128 && cfg->raster_batch != 256)
    return fail(SPL_ERROR_UNSUPPORTED, "raster_batch must be 128 or 256");
  if (cfg->raster_longest_first > 1) return fail(SPL_ERROR_INVALID_ARGUMENT, "raster_longest_first must be 0 or 1");
  return SPL_OK;
}

// Defaults, digit passes, sort block shape and the per-view word counts of
// the workspace for `cap` splats and `pair_cap` (splat, tile) pairs
// (0: tile_pairs_per_splat per splat).
void size_workspace(spl_renderer* r, const spl_configuration* cfg, uint64_t cap, uint64_t pair_cap) {
  r->cfg = *cfg;
  if (r->cfg.tile_pairs_per_splat == 0) r->cfg.tile_pairs_per_splat = 4;
  if (r->cfg.max_views == 0) r->cfg.max_views = 1;
  r->max_views = r->cfg.max_views;
  r->cap = cap;
  r->depth_bits = (int)cfg->radix_bits_per_digit;
  r->depth_passes = (30 + r->depth_bits - 1) / r->depth_bits;
  if (r->cfg.sort_tile_keys == 0) r->cfg.sort_tile_keys = 4096;
  if (r->cfg.raster_batch == 0) r->cfg.raster_batch = 128;
  r->sort_threads = r->cfg.sort_tile_keys == 16384 ? kSortThreadsXL
                    : r->cfg.sort_tile_keys == 8192 ? kSortThreadsWide : kSortThreads;
  const int radix = 1 << r->depth_bits;
  // a multiple of 4: the tile-range pass reads the pair keys as uint4
  if (pair_cap == 0) pair_cap = r->cap * (uint64_t)r->cfg.tile_pairs_per_splat;
  r->pair_cap = std::min<uint64_t>(pair_cap, (1ull << 30) - (1ull << 20)) & ~3ull;
  if (r->pair_cap == 0) r->pair_cap = 4;
  r->cull_tiles = div_up(r->cap, kCullTile);
  r->sort_words = (size_t)div_up(r->cap, kSortThreads * kSortItems) * radix;
  r->bin_words = div_up(r->cap, kBinThreads * kBinItems);
  r->pair_words = (size_t)div_up(r->pair_cap, kSortThreads * kSortItems) * 256;
}

}  // namespace

namespace spl {

int renderer_create(spl_device* dev, const spl_configuration* cfg, bool own_stream, uint32_t presize_tiles,
                    spl_renderer** out, uint64_t* dry_bytes) {
  if (!dev || !cfg || (!out && !dry_bytes)) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_renderer_create: null argument");
  if (out) *out = nullptr;
  const bool dry = dry_bytes != nullptr;
  int rc0 = validate_config(cfg);
  if (rc0) return rc0;
  HIP_TRY(hipSetDevice(dev->ordinal));
  auto* r = new (std::nothrow) spl_renderer();
  if (!r) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_renderer_create: host allocation");
  r->dev = dev;
  r->stream = dev->stream;
  if (own_stream && !dry) {
    r->stream = acquire_stream(dev);
    r->own_stream = true;
  }
  size_workspace(r, cfg, cfg->max_splat_count, 0);
  const size_t V = r->max_views;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  auto alloc = [&](auto** p, size_t count) {
    r->bytes += std::max<size_t>(count, 1) * sizeof(**p);
    return dry ? hipSuccess : dev_alloc(p, count);
  };
  for (int i = 0; i < 2; ++i) {
    chk(alloc(&r->keys[i], V * r->cap));
    chk(alloc(&r->vals[i], V * r->cap));
    chk(alloc(&r->pair_keys[i], V * r->pair_cap));
    chk(alloc(&r->pair_vals[i], V * r->pair_cap));
  }
  chk(alloc(&r->compact_idx, V * r->cap));
  chk(alloc(&r->records, V * r->cap * kRecordFloat4));
  chk(alloc(&r->boxes, V * r->cap));
  chk(alloc(&r->cull_status, V * r->cull_tiles));
  chk(alloc(&r->tile_prefix, V * r->cull_tiles));
  chk(alloc(&r->sort_status, V * r->sort_words * kMaxDepthPasses));
  chk(alloc(&r->bin_status, V * r->bin_words));
  chk(alloc(&r->pair_status, V * r->pair_words * kMaxPairPasses));
  chk(alloc(&r->counters, V * kCounterWords));
  chk(alloc(&r->depth_hist, V * kMaxDepthPasses * 1024));
  chk(alloc(&r->pair_hist, V * kMaxPairPasses * 256));
  chk(alloc(&r->vis, r->cap));
  chk(alloc(&r->d_uniforms, V));
  chk(alloc(&r->d_pixels, V));
  if (dry) {
    *dry_bytes = r->bytes + (uint64_t)presize_tiles * V * (sizeof(uint2) + sizeof(uint32_t));
    delete r;
    return SPL_OK;
  }
  r->side = acquire_stream(dev);
  chk(hipEventCreateWithFlags(&r->fork, hipEventDisableTiming));
  chk(hipEventCreateWithFlags(&r->join, hipEventDisableTiming));
  for (auto& sl : r->slots) {
    chk(hipHostMalloc((void**)&sl.uniforms, V * sizeof(Uniforms), hipHostMallocDefault));
    chk(hipHostMalloc((void**)&sl.pixels, V * sizeof(uint32_t*), hipHostMallocDefault));
    chk(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  }
  if (e == hipSuccess) {
    // Every stage clears the bookkeeping of the stage before it once that
    // stage is done, so this is the only initialisation ever needed.
    chk(hipMemsetAsync(r->cull_status, 0, V * r->cull_tiles * 4, r->stream));
    chk(hipMemsetAsync(r->sort_status, 0, V * r->sort_words * kMaxDepthPasses * 4, r->stream));
    chk(hipMemsetAsync(r->bin_status, 0, V * r->bin_words * 8, r->stream));
    chk(hipMemsetAsync(r->pair_status, 0, V * r->pair_words * kMaxPairPasses * 4, r->stream));
    chk(hipMemsetAsync(r->counters, 0, V * kCounterWords * 4, r->stream));
    chk(hipMemsetAsync(r->depth_hist, 0, V * kMaxDepthPasses * 1024 * 4, r->stream));
    chk(hipMemsetAsync(r->pair_hist, 0, V * kMaxPairPasses * 256 * 4, r->stream));
    chk(hipStreamSynchronize(r->stream));
  }
  if (e != hipSuccess) {
    renderer_free(r);
    delete r;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
      return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_renderer_create: device memory for max_splat_count x max_views");
    return fail(SPL_ERROR_DEVICE, std::string("spl_renderer_create: ") + hipGetErrorString(e));
  }
  if (presize_tiles) {
    const int rc = ensure_ranges(r, presize_tiles);
    if (rc) {
      renderer_free(r);
      delete r;
      return rc;
    }
  }
  device_retain(dev);
  *out = r;
  return SPL_OK;
}

int render_frame_scratch(spl_device* dev, const spl_configuration* cfg, spl_scene* s, spl_frame* f,
                         const float cam[16], const float view[16], uint8_t* scratch, uint64_t bytes) {
  int rc = validate_config(cfg);
  if (rc) return rc;
  spl_renderer r;  // host-side description only: every buffer below points into `scratch`
  r.dev = dev;
  r.stream = dev->stream;
  const uint64_t cap = std::max<uint64_t>(s->count, 1);
  Uniforms u;
  frame_uniforms(*cfg, cam, view, f->width, f->height, u);
  const uint32_t tiles = u.tiles_x * u.tiles_y;
  uint64_t zero_bytes = 0;
  // Carves the workspace at 256-byte alignment; the words that must start
  // at zero (counters, status words, histograms, tile ranges) come first, so
  // one fill clears them.  Returns the bytes used.
  auto carve = [&](uint8_t* base) -> uint64_t {
    uint64_t off = 0;
    auto take = [&](auto** p, uint64_t count) {
      *p = reinterpret_cast<std::remove_reference_t<decltype(*p)>>(base + off);
      off += (std::max<uint64_t>(count, 1) * sizeof(**p) + 255) & ~255ull;
    };
    take(&r.counters, kCounterWords);
    take(&r.cull_status, r.cull_tiles);
    take(&r.sort_status, r.sort_words * kMaxDepthPasses);
    take(&r.bin_status, r.bin_words);
    take(&r.pair_status, r.pair_words * kMaxPairPasses);
    take(&r.depth_hist, (uint64_t)kMaxDepthPasses * 1024);
    take(&r.pair_hist, (uint64_t)kMaxPairPasses * 256);
    take(&r.ranges, tiles);
    zero_bytes = off;
    take(&r.d_uniforms, 1);
    take(&r.d_pixels, 1);
    take(&r.tile_prefix, r.cull_tiles);
    take(&r.vis, r.cap);
    for (int i = 0; i < 2; ++i) {
      take(&r.keys[i], r.cap);
      take(&r.vals[i], r.cap);
    }
    take(&r.compact_idx, r.cap);
    take(&r.records, r.cap * kRecordFloat4);
    take(&r.boxes, r.cap);
    take(&r.tile_order, tiles);
    for (int i = 0; i < 2; ++i) {
      take(&r.pair_keys[i], r.pair_cap);
      take(&r.pair_vals[i], r.pair_cap);
    }
    return off;
  };
  // The (splat, tile) pairs take whatever the rest of the scratch holds.
  spl_configuration one = *cfg;
  one.max_views = 1;
  size_workspace(&r, &one, cap, 4);
  const uint64_t fixed = carve(scratch);
  if (fixed >= bytes) return fail(SPL_ERROR_CAPACITY, "spl_render_frame_static: scene and frame exceed the scratch");
  uint64_t pairs = (bytes - fixed) / 17;
  for (;;) {
    size_workspace(&r, &one, cap, pairs);
    if (carve(scratch) <= bytes) break;
    pairs -= pairs / 64 + 4;
    if (pairs < 4) return fail(SPL_ERROR_CAPACITY, "spl_render_frame_static: scene and frame exceed the scratch");
  }
  r.ranges_cap = tiles;
  launch_static_setup(u, f->pixels, r.d_uniforms, r.d_pixels, reinterpret_cast<uint32_t*>(scratch),
                      zero_bytes / 4, r.stream);
  BatchShape b;
  b.scene = s;
  b.scene_uid = s->uid;
  b.n = s->count;
  b.views = 1;
  b.max_tiles = tiles;
  b.pair_passes = pair_passes_for(tiles);
  launch_sequence(&r, b, r.stream, nullptr);
  HIP_TRY(hipGetLastError());
  return SPL_OK;
}

hipStream_t renderer_stream(const spl_renderer* r) { return r->stream; }
uint64_t renderer_bytes(const spl_renderer* r) { return r->bytes; }

int renderer_counter_words(const spl_renderer* r, uint32_t view, const uint32_t** words) {
  if (view >= r->max_views) return fail(SPL_ERROR_INVALID_ARGUMENT, "view outside the renderer's batch width");
  *words = r->counters + (size_t)view * kCounterWords;
  return SPL_OK;
}

}  // namespace spl

extern "C" {

int spl_renderer_create(spl_device* dev, const spl_configuration* cfg, spl_renderer** out) {
  return renderer_create(dev, cfg, false, 0, out, nullptr);
}

int spl_renderer_create_async(spl_device* dev, const spl_configuration* cfg, spl_renderer** out) {
  return renderer_create(dev, cfg, true, 0, out, nullptr);
}

int spl_renderer_synchronize(spl_renderer* r) {
  if (!r) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_renderer_synchronize: null renderer");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->stream));
  return SPL_OK;
}

int spl_renderer_destroy(spl_renderer* r) {
  if (!r) return SPL_OK;
  (void)hipSetDevice(r->dev->ordinal);
  (void)hipStreamSynchronize(r->stream);
  renderer_free(r);
  spl_device* d = r->dev;
  delete r;
  spl::device_release(d);
  return SPL_OK;
}

int spl_renderer_get_config(const spl_renderer* r, spl_configuration* out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_renderer_get_config: null argument");
  *out = r->cfg;
  return SPL_OK;
}

int spl_render_views(spl_renderer* r, spl_scene* s, uint32_t view_count, spl_frame* const* frames,
                     const float* cams, const float* views) {
  if (!r || !s || !frames || !cams || !views) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_views: null argument");
  if (s->dev != r->dev) return fail(SPL_ERROR_INVALID_ARGUMENT, "scene and renderer must share a device");
  if (s->count > r->cap) return fail(SPL_ERROR_CAPACITY, "scene has more splats than max_splat_count");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  for (uint32_t first = 0; first < view_count; first += r->max_views) {
    const uint32_t nv = std::min<uint32_t>(r->max_views, view_count - first);
    int rc = enqueue_batch(r, s, nv, frames + first, cams + 16 * (size_t)first, views + 16 * (size_t)first);
    if (rc) return rc;
  }
  return SPL_OK;
}

int spl_render_frame(spl_renderer* r, spl_scene* s, spl_frame* f, const float cam[16], const float view[16]) {
  if (!r || !s || !f || !cam || !view) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_frame: null argument");
  spl_frame* frames[1] = {f};
  return spl_render_views(r, s, 1, frames, cam, view);
}

// ---------------------------------------------------------------------------
// Parity and timing taps
// ---------------------------------------------------------------------------

int spl_debug_select_view(spl_renderer* r, uint32_t view) {
  if (!r) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  if (view >= r->max_views || (r->have_frame && view >= r->last_views))
    return fail(SPL_ERROR_INVALID_ARGUMENT, "view outside the last batch");
  r->debug_view = view;
  return SPL_OK;
}

int spl_debug_visible_count(spl_renderer* r, uint64_t* out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  uint32_t v = 0;
  int rc = read_counter(r, r->debug_view, kVisible, &v);
  *out = v;
  return rc;
}

int spl_debug_pair_count(spl_renderer* r, uint64_t* out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  uint32_t v = 0, o = 0;
  int rc = read_counter(r, r->debug_view, kPairs, &v);
  if (rc) return rc;
  rc = read_counter(r, r->debug_view, kPairOverflow, &o);
  *out = v;
  if (rc) return rc;
  if (o) return fail(SPL_ERROR_CAPACITY, "screen-tile pair capacity exceeded in the last frame");
  return SPL_OK;
}

int spl_debug_read_entries(spl_renderer* r, uint32_t* keys, uint32_t* indices, uint64_t capacity, uint64_t* count) {
  if (!r || !count) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  if (!r->have_frame) return fail(SPL_ERROR_INVALID_ARGUMENT, "no frame rendered yet");
  std::vector<uint32_t> order, k;
  int rc = draw_order(r, order, &k);
  if (rc) return rc;
  *count = order.size();
  std::vector<uint32_t> idx(order.size());
  if (!order.empty())
    HIP_TRY(hipMemcpy(idx.data(), r->compact_idx + (size_t)r->debug_view * r->cap, idx.size() * 4,
                      hipMemcpyDeviceToHost));
  const size_t m = (size_t)std::min<uint64_t>(capacity, order.size());
  for (size_t i = 0; i < m; ++i) {
    if (keys) keys[i] = k[i];
    if (indices) indices[i] = idx[order[i]];
  }
  return SPL_OK;
}

int spl_debug_read_records(spl_renderer* r, float* out, uint64_t capacity, uint64_t* count) {
  if (!r || !count) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  if (!r->have_frame) return fail(SPL_ERROR_INVALID_ARGUMENT, "no frame rendered yet");
  std::vector<uint32_t> order;
  int rc = draw_order(r, order, nullptr);
  if (rc) return rc;
  *count = order.size();
  std::vector<float4> rec(order.size() * kRecordFloat4);
  if (!order.empty())
    HIP_TRY(hipMemcpy(rec.data(), r->records + (size_t)r->debug_view * r->cap * kRecordFloat4, rec.size() * 16,
                      hipMemcpyDeviceToHost));
  const size_t m = (size_t)std::min<uint64_t>(capacity, order.size());
  if (out)
    for (size_t i = 0; i < m; ++i) {
      const float4* p = &rec[(size_t)order[i] * kRecordFloat4];
      const float v[10] = {p[0].x, p[0].y, p[0].z, p[0].w, p[1].x, p[1].y, p[1].z, p[1].w, p[2].x, p[2].y};
      std::memcpy(out + 10 * i, v, sizeof(v));
    }
  return SPL_OK;
}

int spl_debug_enable_conics(spl_renderer* r, int enable) {
  if (!r) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  if (enable && !r->conics) HIP_TRY(dev_alloc(&r->conics, r->cap * 6));
  r->conics_on = enable != 0;
  return SPL_OK;
}

int spl_debug_read_conics(spl_renderer* r, double* out, uint64_t capacity, uint64_t* count) {
  if (!r || !count) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  if (!r->conics_on || !r->have_frame || r->debug_view != 0)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "conic capture covers view 0 once enabled");
  std::vector<uint32_t> order;
  int rc = draw_order(r, order, nullptr);
  if (rc) return rc;
  *count = order.size();
  std::vector<double> cn(order.size() * 6);
  if (!order.empty()) HIP_TRY(hipMemcpy(cn.data(), r->conics, cn.size() * 8, hipMemcpyDeviceToHost));
  const size_t m = (size_t)std::min<uint64_t>(capacity, order.size());
  if (out)
    for (size_t i = 0; i < m; ++i) std::memcpy(out + 6 * i, &cn[(size_t)order[i] * 6], 48);
  return SPL_OK;
}

int spl_debug_enable_tile_stats(spl_renderer* r, int enable) {
  if (!r) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  r->tile_stats_on = enable != 0;
  return SPL_OK;
}

int spl_debug_read_tile_stats(spl_renderer* r, uint32_t* out, uint64_t capacity, uint64_t* count) {
  if (!r || !count) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  if (!r->tile_stats_on || !r->tile_work || !r->have_frame)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "tile statistics not enabled");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->stream));
  *count = r->last_tiles[0];
  const size_t m = (size_t)std::min<uint64_t>(capacity, r->last_tiles[0]);
  if (out && m) HIP_TRY(hipMemcpy(out, r->tile_work, m * sizeof(uint2), hipMemcpyDeviceToHost));
  return SPL_OK;
}

int spl_debug_enable_timing(spl_renderer* r, int enable) {
  if (!r) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->stream));
  r->timing = enable != 0;
  r->ev_used = 0;
  for (double& v : r->stage_sum_ms) v = 0.0;
  r->stage_frames = 0;
  return SPL_OK;
}

int spl_debug_stage_times(spl_renderer* r, float* ms, uint32_t capacity) {
  if (!r || !ms) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->stream));
  for (size_t f = 0; f < r->ev_used; ++f) {
    for (int i = 0; i < kStageCount; ++i) {
      float t = 0.0f;
      HIP_TRY(hipEventElapsedTime(&t, r->ev_pool[f][i], r->ev_pool[f][i + 1]));
      r->stage_sum_ms[i] += t;
    }
    r->stage_frames += r->ev_views[f];
  }
  r->ev_used = 0;
  for (uint32_t i = 0; i < capacity && i < (uint32_t)kStageCount; ++i)
    ms[i] = r->stage_frames ? (float)(r->stage_sum_ms[i] / (double)r->stage_frames) : 0.0f;
  return SPL_OK;
}

}  // extern "C"
