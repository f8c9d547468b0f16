// spl_internal.h -- host-side declarations shared by the C-ABI layer and the
// kernel translation units.
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
constexpr int kRecordFloat4 = 3;     // 48-byte record: rgba | A00 A10 A01 A11 | tx ty rect rect

// Counter block layout (u32 words).
enum CounterSlot : int {
  kCtrCull = 0,
  kCtrSort0 = 1,    // .. kCtrSort0 + 3
  kCtrBin = 5,
  kCtrPair0 = 6,    // .. kCtrPair0 + 2
  kCtrCount = 9,
  kDone = 16,       // done counters mirror tile counters at +16
  kVisible = 32,
  kPairs = 33,
  kPairOverflow = 34,
  kScatterError = 35,
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

struct CullLaunch {
  const float* centers;
  uint32_t n;
  float vp[16];
  float tol;
  int gpu_predicate;        // 1: shaders.wgsl:54-61 (DepthSorting::None), 0: renderer.rs:381-387
  uint32_t* keys_out;
  uint32_t* idx_out;
  uint32_t* hist;           // [passes][1 << bits], accumulated
  int passes, bits;
  uint32_t* status;
  uint32_t* counters;
  int grid;
};

struct SortPassLaunch {
  const uint32_t* keys_in;
  const uint32_t* vals_in;  // null: values are the input positions
  uint32_t* keys_out;
  uint32_t* vals_out;
  const uint32_t* count;    // device count of entries
  uint32_t count_cap;
  int shift;
  const uint32_t* hist;     // this pass's global histogram
  uint32_t* status;         // [tiles][radix]
  uint32_t* tile_counter;
  uint32_t* done_counter;
  uint32_t* clear_a; size_t clear_a_words;
  uint32_t* clear_b; size_t clear_b_words;
  uint32_t* error_flag;     // set when a scatter position falls outside [0, count)
  int grid;
};

struct ProjectLaunch {
  SceneView scene;
  const uint32_t* compact_idx;
  const uint32_t* count;
  Uniforms u;
  int sh_order;
  int use_covariance_for_scale;
  int use_unaligned_rectangles;
  float4* records;
  double* conics;           // optional, 6 per entry
  uint32_t* clear_a; size_t clear_a_words;
  uint32_t* clear_b; size_t clear_b_words;
  int grid;
};

struct BinLaunch {
  const uint32_t* order;    // record index per draw rank; null = identity
  const float4* records;
  const uint32_t* count;    // visible count
  uint32_t tiles_x;
  uint32_t* pair_keys;
  uint32_t* pair_vals;
  uint32_t pair_cap;
  uint32_t* hist;           // [pair passes][256], accumulated
  int pair_passes;
  uint64_t* status;
  uint32_t* counters;
  int grid;
};

struct RangesLaunch {
  const uint32_t* pair_keys;
  const uint32_t* counters;
  uint32_t pair_cap;
  uint32_t tiles;
  uint2* ranges;
  uint32_t* clear_a; size_t clear_a_words;
  uint32_t* clear_b; size_t clear_b_words;
};

struct RasterLaunch {
  const uint32_t* pair_vals;
  const uint32_t* visible;  // record count: pair values at or past it are ignored
  const float4* records;
  uint2* ranges;
  Uniforms u;
  int use_unaligned_rectangles;
  int blend_f32;
  int bgra;
  uint32_t* pixels;
  uint2* tile_work;         // optional: per tile (list length, entries consumed)
};

constexpr int kCullThreads = 256, kCullItems = 16;
constexpr int kSortThreads = 256, kSortItems = 16;
constexpr int kBinThreads = 256, kBinItems = 4;

void launch_cull(const CullLaunch& a, hipStream_t s);
void launch_sort_pass(int bits, const SortPassLaunch& a, hipStream_t s);
void launch_project(const ProjectLaunch& a, hipStream_t s);
void launch_bin(const BinLaunch& a, hipStream_t s);
void launch_ranges(const RangesLaunch& a, uint32_t max_pairs, hipStream_t s);
void launch_raster(const RasterLaunch& a, hipStream_t s);

}  // namespace spl
