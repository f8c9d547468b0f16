// spl_internal.h -- host-side declarations shared by the C-ABI layer and the
// kernel translation units.
//
// Every stage runs for a batch of views at once: per-view buffers live in
// one allocation each, view v at base + v * stride, and kernels take the view
// from blockIdx.y (or loop over views where one read of the scene serves all
// of them: cull and project).  A single frame is the batch of one.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace spl {

// Frame uniforms: renderer.rs:49-61 plus the frame's tile grid.
struct Uniforms {
  float camera_matrix[16];
  float view_matrix[16];
  float view_projection_matrix[16];
  float view_size[2];
  uint32_t image_size[2];
  float frustum_culling_tolerance;
  float ellipse_size_bias;
  float ellipse_margin;
  float splat_scale;
  uint32_t tiles_x, tiles_y;
};

constexpr int kTileSize = 16;        // raster tile edge in pixels (16x16 = 256 lanes)
constexpr int kMaxDepthPasses = 4;   // 30-bit depth keys: 4 x 8 bits or 3 x 10 bits
constexpr int kMaxPairPasses = 3;    // screen-tile ids up to 24 bits
constexpr int kRecordFloat4 = 3;     // 48-byte record: rgba | A00 A10 A01 A11 | tx ty box box
constexpr int kMaxViews = 32;        // views per launch (one bit each in the visibility masks)

// Counter block layout (u32 words), one block per view.
enum CounterSlot : int {
  kCtrCull = 0,     // view 0's block holds the shared multi-view cull counters
  kCtrSort0 = 1,    // .. kCtrSort0 + 3
  kCtrBin = 5,
  kCtrPair0 = 6,    // .. kCtrPair0 + 2
  kCtrCount = 9,
  kDone = 16,       // done counters mirror tile counters at +16
  kVisible = 32,
  kPairs = 33,
  kPairOverflow = 34,
  kScatterError = 35,
  kHeavyTiles = 36, // tiles the raster takes first (longest lists)
  kSplitTiles = 37, // the first of those, that the raster splits over four blocks
  kCounterWords = 64
};

// Scene data in HBM, one stream per attribute (structure of arrays).
struct SceneView {
  const float* centers;     // 3 floats per splat, packed
  const float4* rotations;  // (r, i, j, k) normalised
  const float4* scale_alpha;// exp-scale xyz, sigmoid alpha
  const float* sh;          // 48 floats per splat, coefficient-major RGB
  uint32_t count;
};

// shaders.wgsl:249-266: the spherical-harmonics basis constants.  A scene
// stores coefficient k of every colour channel multiplied by constant k
// (the product each term of the reference's lookup starts with).
constexpr float kShConstants[16] = {
    0.28209479177387814f, -0.4886025119029199f, 0.4886025119029199f, -0.4886025119029199f,
    1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
    0.5462742152960396f, -0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
    0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f,
};

constexpr int kCullThreads = 256, kCullItems = 16;   // 4096 splats per cull / project tile
constexpr int kSortThreads = 256, kSortItems = 16;   // 4096 keys per sort tile (the narrow form)
constexpr int kSortThreadsWide = 512;                // 8192 keys per sort tile
constexpr int kSortThreadsXL = 1024;                 // 16384 keys per sort tile
constexpr int kBinThreads = 256, kBinItems = 4;      // 1024 draw ranks per bin tile
constexpr int kCullTile = kCullThreads * kCullItems;

struct CullLaunch {
  const float* centers;
  uint32_t n;
  uint32_t views;
  const Uniforms* uniforms;  // [views]
  int gpu_predicate;         // 1: shaders.wgsl:54-61 (DepthSorting::None), 0: renderer.rs:381-387
  int write_keys;
  uint32_t* keys; size_t keys_stride;        // view v: keys + v * keys_stride
  uint32_t* idx; size_t idx_stride;
  uint32_t* vis;                             // [n] visibility mask over views
  uint32_t* tile_prefix; size_t tp_stride;   // [views][cull tiles] exclusive prefix
  uint32_t* status; size_t status_stride;    // [views][cull tiles]
  uint32_t* counters;                        // [views][kCounterWords]
  int grid;
};

struct HistLaunch {
  const uint32_t* keys; size_t keys_stride;
  uint32_t* hist; size_t hist_stride;        // [views][passes << bits]
  uint32_t* counters;
  int passes, bits;
  uint32_t* clear; size_t clear_words; size_t clear_stride;  // per view
  int grid_x, views;
};

struct SortPassLaunch {
  const uint32_t* keys_in;
  const uint32_t* vals_in;   // null: values are the input positions
  uint32_t* keys_out;
  uint32_t* vals_out;
  size_t entry_stride;       // per view, all four arrays
  uint32_t* counters;        // [views][kCounterWords]
  int count_slot;            // kVisible or kPairs
  uint32_t count_cap;
  int shift;
  const uint32_t* hist; size_t hist_stride;  // this pass's histogram, per view
  uint32_t* status; size_t status_stride;    // [tiles][radix], per view
  int tile_slot;             // counter slot of this pass's tile counter
  uint32_t* clear_a; size_t clear_a_words; size_t clear_a_stride;
  uint32_t* clear_b; size_t clear_b_words; size_t clear_b_stride;
  int threads;               // kSortThreads, kSortThreadsWide or kSortThreadsXL lanes, kSortItems keys each
  int grid_x, views;
};

struct ProjectLaunch {
  SceneView scene;
  uint32_t views;
  const Uniforms* uniforms;
  const uint32_t* vis;
  const uint32_t* tile_prefix; size_t tp_stride;
  int sh_order;
  int use_covariance_for_scale;
  int use_unaligned_rectangles;
  float4* records; size_t rec_stride;       // float4 units per view
  uint2* boxes; size_t box_stride;          // pixel box per entry, dense, for the binning
  double* conics;                           // optional, view 0 only, 6 per entry
  uint32_t* clear_a; size_t clear_a_words; size_t clear_a_stride;
  uint32_t* clear_b; size_t clear_b_words; size_t clear_b_stride;
  int grid;
};

struct BinLaunch {
  const uint32_t* order; size_t order_stride;  // record index per draw rank; null = identity
  const uint2* boxes; size_t box_stride;       // pixel box per record
  const Uniforms* uniforms;
  uint32_t* pair_keys;
  uint32_t* pair_vals; size_t pair_stride;
  uint32_t pair_cap;
  uint32_t* hist; size_t hist_stride;          // [pair passes][256] per view, accumulated
  int pair_passes;
  uint64_t* status; size_t status_stride;
  uint32_t* counters;
  // finished bookkeeping of the depth sort, zeroed here (per view)
  uint32_t* clear_a; size_t clear_a_words, clear_a_stride;
  uint32_t* clear_b; size_t clear_b_words, clear_b_stride;
  int grid_x, views;
};

struct RangesLaunch {
  const uint32_t* pair_keys; size_t pair_stride;
  uint32_t* counters;
  uint32_t pair_cap;
  const Uniforms* uniforms;
  uint2* ranges; size_t ranges_stride;
  uint32_t* clear_a; size_t clear_a_words; size_t clear_a_stride;
  uint32_t* clear_b; size_t clear_b_words; size_t clear_b_stride;
  int grid_x, views;
};

struct OrderLaunch {
  const uint2* ranges; size_t ranges_stride;
  const Uniforms* uniforms;
  uint32_t* counters;
  uint32_t pair_cap;
  uint32_t* order; size_t order_stride;      // per view: heavy tiles, longest first, then the rest in order
  uint32_t heavy_cap;                        // at most this many heavy tiles per view
  uint32_t heavy_mult, heavy_min;            // heavy: at least heavy_mult x the mean and heavy_min pairs
  uint32_t split_mult;                       // split: at least split_mult x the mean and 2048 pairs (0: none)
  int views;
};

struct RasterLaunch {
  const uint32_t* order; size_t order_stride;
  uint32_t slots_per_view;   // multiple of 8, >= the largest view's tile count
  uint32_t heavy_cap;
  const uint32_t* pair_vals; size_t pair_stride;
  const float4* records; size_t rec_stride;
  uint2* ranges; size_t ranges_stride;
  const Uniforms* uniforms;
  uint32_t* const* pixels;   // [views] frame buffers
  uint32_t* counters;
  int use_unaligned_rectangles;
  int blend_f32;
  int walk;                  // raster mode: 0 entry-major per quadrant, 1 per-pixel walk, 2 per 4x4 sub-block,
                             // 3 per-lane walk over coverage masks
  int batch_entries;         // entries staged in LDS per batch: 128 or 256
  int longest_first;         // take the tiles in the order k_tile_order wrote
  int split_heavy;           // mode 3, longest first: a heavy tile's quadrants on four blocks, two rows a wave
  int bgra;
  uint2* tile_work;          // optional, view 0: per tile (list length, entries consumed)
  int grid_x, views;
};

void launch_cull(const CullLaunch& a, hipStream_t s);
void launch_histograms(const HistLaunch& a, hipStream_t s);
void launch_sort_pass(int bits, const SortPassLaunch& a, hipStream_t s);
void launch_project(const ProjectLaunch& a, hipStream_t s);
void launch_bin(const BinLaunch& a, hipStream_t s);
void launch_ranges(const RangesLaunch& a, hipStream_t s);
void launch_tile_order(const OrderLaunch& a, hipStream_t s);
void launch_raster(const RasterLaunch& a, hipStream_t s);

// spl_render_frame_static: the scratch array's size and the largest scene and
// frame it takes.
constexpr uint64_t kStaticScratchBytes = 48ull << 20;
constexpr uint32_t kStaticMaxSplats = 1u << 17;
constexpr uint32_t kStaticMaxWidth = 3840, kStaticMaxHeight = 2160;
// Writes the frame's uniforms and pixel pointer and zeroes `zero_words`
// words at `zero` (the bookkeeping that starts every frame at zero), with
// one kernel.
void launch_static_setup(const Uniforms& u, uint32_t* pixels, Uniforms* d_uniforms, uint32_t** d_pixels,
                         uint32_t* zero, size_t zero_words, hipStream_t s);

}  // namespace spl
