/*
 * splatter.h -- C ABI of the MI355X-native splat pipeline.
 *
 * Drop-in for the per-frame path of Lichtso/splatter: every entry point below
 * names the Rust item it replaces.  A Rust crate keeps the reference's public
 * API (`splatter::renderer::{Renderer, Configuration, DepthSorting}`,
 * `splatter::scene::Scene`) and forwards to these functions through an
 * `extern "C"` block that declares them as written here.
 *
 * Conventions
 *  - Opaque handles; every call returns an int status (SPL_OK == 0, negative
 *    on error).  spl_last_error() returns the calling thread's last message.
 *    No C++ exception crosses the ABI.  (The reference panics instead:
 *    scene.rs:202-224 unwraps, renderer.rs:90 asserts.)
 *  - Matrices are 16 floats, column-major, column c = (x, y, z, w) at
 *    [4c, 4c+4): the layout of `[ppga3d::Point; 4]` produced by
 *    utils.rs:31-42 motor3d_to_mat4.
 *  - Device work is asynchronous on the device's stream, like wgpu's
 *    queue.submit (renderer.rs:465); spl_synchronize() waits for it.
 *  - One renderer renders one frame at a time (the reference shares its
 *    sorting/entry buffers across scenes the same way, renderer.rs:79-83).
 *    Scenes are read-only during rendering and may be shared by renderers
 *    of the same device.
 */
#ifndef SPLATTER_H
#define SPLATTER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spl_device spl_device;
typedef struct spl_renderer spl_renderer;
typedef struct spl_scene spl_scene;
typedef struct spl_frame spl_frame;

enum {
    SPL_OK = 0,
    SPL_ERROR_INVALID_ARGUMENT = -1,
    SPL_ERROR_OUT_OF_MEMORY = -2,
    SPL_ERROR_DEVICE = -3,
    SPL_ERROR_UNSUPPORTED = -4,
    SPL_ERROR_IO = -5,
    SPL_ERROR_CAPACITY = -6
};

/* renderer.rs:14-23 DepthSorting.  NONE draws in splat order and culls with
 * the vertex-stage predicate (shaders.wgsl:521-524).  CPU, GPU and
 * GPU_INDIRECT_DRAW all produce the CPU fallback's visible set and stable
 * order (renderer.rs:372-396), computed on the GPU. */
typedef enum {
    SPL_DEPTH_SORTING_NONE = 0,
    SPL_DEPTH_SORTING_CPU = 1,
    SPL_DEPTH_SORTING_GPU = 2,
    SPL_DEPTH_SORTING_GPU_INDIRECT_DRAW = 3
} spl_depth_sorting;

/* Replaces surface_configuration, of which only .format is read
 * (renderer.rs:259); the showcase uses Bgra8Unorm (application_framework.rs:184). */
typedef enum {
    SPL_FRAME_FORMAT_BGRA8_UNORM = 0,
    SPL_FRAME_FORMAT_RGBA8_UNORM = 1
} spl_frame_format;

/* UNORM8 reproduces an 8-bit unorm render target: the source is clamped to
 * [0,1] and the destination re-quantised after every blend.  F32 keeps the
 * running colour and transmittance in f32 and quantises once at the end. */
typedef enum {
    SPL_BLEND_UNORM8 = 0,
    SPL_BLEND_F32 = 1
} spl_blend_mode;

/* renderer.rs:26-47 Configuration, field by field, plus extensions. */
typedef struct spl_configuration {
    spl_frame_format frame_format;        /* surface_configuration.format */
    spl_depth_sorting depth_sorting;      /* depth_sorting */
    int32_t use_covariance_for_scale;     /* use_covariance_for_scale */
    int32_t use_unaligned_rectangles;     /* use_unaligned_rectangles */
    uint32_t spherical_harmonics_order;   /* spherical_harmonics_order, 0..=3 */
    uint64_t max_splat_count;             /* max_splat_count */
    uint32_t radix_bits_per_digit;        /* radix_bits_per_digit: 8 (default) or 10 */
    float frustum_culling_tolerance;      /* frustum_culling_tolerance */
    float ellipse_margin;                 /* ellipse_margin */
    float splat_scale;                    /* splat_scale */
    /* extensions */
    float field_of_view_y;                /* radians; the reference hard-codes PI/2 (renderer.rs:366) */
    spl_blend_mode blend_mode;
    uint32_t tile_pairs_per_splat;        /* capacity of (splat, screen tile) pairs, per max_splat_count */
    uint32_t max_views;                   /* views per batched launch (1..=32); workspace scales with it */
    int32_t use_hip_graphs;               /* replay each batch shape's launch sequence from a hipGraph */
    uint32_t sort_tile_keys;              /* keys per radix-sort tile: 4096, 8192 or 16384 (256, 512 or
                                             1024 lanes) */
    uint32_t raster_mode;                 /* 0: each wave's lanes test every entry of its 8x8 quadrant;
                                             1: each lane walks the entries covering its pixel;
                                             2: four lane groups step through the lists of their 4x4
                                             sub-blocks side by side;
                                             3 (default): for 32 list slots at a time each lane finds
                                             the entries that can cover its pixel (per row, from the
                                             rectangles' column intervals), then walks them on its own;
                                             one view also splits its heaviest tiles over four blocks
                                             (all the same image) */
    uint32_t raster_batch;                /* entries a raster block stages per batch: 128 or 256
                                             (batches of two or more views; one view uses 256) */
    uint32_t raster_longest_first;        /* 1: tiles with the longest lists start first (same image;
                                             batches of two or more views; one view always does) */
} spl_configuration;

const char* spl_last_error(void);

/* The showcase's Configuration (showcase.rs:39-50) with field_of_view_y = PI/2. */
void spl_configuration_default(spl_configuration* config);

/* wgpu::Device + wgpu::Queue (application_framework.rs:167-177): one HIP
 * device and one stream. */
int spl_device_open(int ordinal, spl_device** out);
int spl_device_close(spl_device* device);
int spl_synchronize(spl_device* device);

/* renderer.rs:88 Renderer::new: allocates the sorting, entry and record
 * workspace for config->max_splat_count splats. */
int spl_renderer_create(spl_device* device, const spl_configuration* config, spl_renderer** out);
int spl_renderer_destroy(spl_renderer* renderer);

/* No counterpart in the reference; for callers that keep several batches in
 * flight.  A renderer whose launches go to a stream of its own instead of
 * the device's, so that the batches of two such renderers run on the device
 * at the same time (the serving form's slots are such renderers).  Scenes,
 * frames and their uploads are complete when their calls return, so it needs
 * no ordering with the device stream.  spl_synchronize does not wait for
 * it: spl_renderer_synchronize does, and must come before reading a frame it
 * rendered. */
int spl_renderer_create_async(spl_device* device, const spl_configuration* config, spl_renderer** out);
int spl_renderer_synchronize(spl_renderer* renderer);
int spl_renderer_get_config(const spl_renderer* renderer, spl_configuration* out);

/* scene.rs:45 Scene::new: splat_count is clamped to max_splat_count (scene.rs:46). */
int spl_scene_create(spl_renderer* renderer, uint64_t splat_count, spl_scene** out);
/* scene.rs:158 Scene::new_dev_test */
int spl_scene_new_dev_test(spl_renderer* renderer, spl_scene** out);
/* scene.rs:40 pub splat_count */
int spl_scene_splat_count(const spl_scene* scene, uint64_t* out);
int spl_scene_destroy(spl_scene* scene);

/* scene.rs:197 Scene::parse_file_header: header length in bytes and the
 * "element vertex" count of a 3DGS PLY file. */
int spl_parse_file_header(const char* path, uint16_t* header_size, uint64_t* splat_count);
/* scene.rs:216 Scene::load_chunk: decodes splats [begin, end) of the file
 * (end clamped to splat_count, scene.rs:217) into the scene. */
int spl_scene_load_chunk(spl_scene* scene, const char* path, uint16_t header_size,
                         uint64_t begin, uint64_t end);
/* The decode of scene.rs:225-255 applied to `count` in-memory 248-byte
 * SerializedSplat records (scene.rs:17-24) written at splat index `first`. */
int spl_scene_load_records(spl_scene* scene, uint64_t first, uint64_t count, const void* records);

/* The render target: replaces the wgpu::TextureView + viewport_size pair
 * (renderer.rs:359-360).  4 bytes per pixel in the renderer's frame_format. */
int spl_frame_create(spl_device* device, uint32_t width, uint32_t height, spl_frame** out);
int spl_frame_destroy(spl_frame* frame);
/* Waits for the device and copies width*height*4 bytes to host memory. */
int spl_frame_read(spl_frame* frame, void* host, uint64_t size);

/* renderer.rs:355 Renderer::render_frame.  camera_matrix and view_matrix
 * are motor3d_to_mat4(motor) and motor3d_to_mat4(motor.inverse())
 * (renderer.rs:364-365); the viewport is the frame's size. */
int spl_render_frame(spl_renderer* renderer, spl_scene* scene, spl_frame* frame,
                     const float camera_matrix[16], const float view_matrix[16]);

/* Batched form (no counterpart in the reference): renders `view_count`
 * views of one scene, view i with matrices at camera_matrices + 16*i and
 * view_matrices + 16*i into frames[i] (frames may differ in size).  Views
 * go through the pipeline max_views at a time, each stage one launch for
 * all of them: the cull reads each splat centre once and the projection
 * each splat's attributes once for every view that sees it.  Same outputs
 * as view_count calls of spl_render_frame. */
int spl_render_views(spl_renderer* renderer, spl_scene* scene, uint32_t view_count,
                     spl_frame* const* frames, const float* camera_matrices,
                     const float* view_matrices);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  spl_render_frame for one view with no renderer: the
 * whole workspace is a fixed 48 MiB scratch array that the library holds on
 * each device it runs on (looked up once, when the device opens), so the
 * call allocates nothing.  Scenes up to 131072 splats and frames up to
 * 3840 x 2160; larger ones fail with SPL_ERROR_CAPACITY.  config is checked
 * as spl_renderer_create checks it; max_views must be 0 or 1, and the
 * (splat, screen tile) pair capacity is what the splats and the frame leave
 * of the scratch (max_splat_count and tile_pairs_per_splat are not used).
 * Asynchronous on the device's stream like spl_render_frame; the image is
 * the same.  Every call on one device shares the scratch array and keeps
 * what the last call left there: calls that overlap on one device (from
 * several threads, or with the device stream's work reordered) race. */
int spl_render_frame_static(spl_device* device, const spl_configuration* config, spl_scene* scene,
                            spl_frame* frame, const float camera_matrix[16], const float view_matrix[16]);

/* ---- serving form (no counterpart in the reference) ---------------------- */
/* A request loop over resident scenes: each request is one view of one scene
 * at its own resolution.  Up to max_in_flight batches run at once, each on a
 * slot of its own (a renderer with its own stream).  When a slot frees up,
 * the oldest pending request and the later pending requests of the same
 * scene (up to renderer.max_views) go out as one spl_render_views batch, so
 * requests that arrive while others are on the device join the next batch.
 * Slots (sized per capacity class of the scene: 2^k or 3*2^(k-1) splats) and
 * frames are pooled within memory_budget_bytes of device memory; idle ones
 * are released least recently used first, and a batch that does not fit
 * waits, in arrival order, for running batches to finish.  Scenes are not
 * counted against the budget.  A request renders exactly what
 * spl_render_frame renders for the same scene, matrices and size.  All calls
 * on one server come from one thread. */
typedef struct spl_server spl_server;

typedef struct spl_server_configuration {
    spl_configuration renderer;   /* pipeline options of every slot; max_splat_count bounds the scenes,
                                     max_views is the widest batch */
    uint32_t max_in_flight;       /* batches on the device at once (1..=8) */
    uint32_t max_width, max_height; /* largest frame a request may ask for */
    uint64_t memory_budget_bytes; /* device memory for slot workspaces and frames */
} spl_server_configuration;

enum { SPL_SERVER_DONE = 0, SPL_SERVER_PENDING = 1, SPL_SERVER_RUNNING = 2 };

typedef struct spl_request_info {
    int status;                   /* SPL_SERVER_* or a negative error code */
    uint32_t width, height;
    uint32_t batch_views;         /* views of the batch the request went out in */
    double submit_ms, start_ms, done_ms;  /* server clock: submitted, launched, seen done */
} spl_request_info;

typedef struct spl_server_stats {
    uint64_t submitted, completed, failed, batches, batched_views, pending;
    uint32_t slots, slots_busy, frames, reserved;
    uint64_t memory_in_use, memory_peak, memory_budget;
} spl_server_stats;

/* max_splat_count 2^23, max_views 8, tile_pairs_per_splat 8, 3 in flight,
 * 3840x2160, 64 GiB. */
void spl_server_configuration_default(spl_server_configuration* config);
int spl_server_create(spl_device* device, const spl_server_configuration* config, spl_server** out);
/* Waits for running batches and frees every slot and frame. */
int spl_server_destroy(spl_server* server);
/* A resident scene for the server's requests (Scene::new without the
 * renderer's clamp; at most renderer.max_splat_count splats).  Loaded and
 * destroyed with the spl_scene_* calls; not destroyed while requests use it. */
int spl_server_scene_create(spl_server* server, uint64_t splat_count, spl_scene** out);
/* Queues a request and runs the scheduler; *ticket names it from then on. */
int spl_server_submit(spl_server* server, spl_scene* scene, uint32_t width, uint32_t height,
                      const float camera_matrix[16], const float view_matrix[16], uint64_t* ticket);
/* Runs the scheduler (reaps finished batches, admits pending requests) and
 * returns the tickets finished since the last poll (at most capacity). */
int spl_server_poll(spl_server* server, uint64_t* finished, uint32_t capacity, uint32_t* count);
/* Runs the scheduler until the request has finished. */
int spl_server_wait(spl_server* server, uint64_t ticket);
int spl_server_request_info(spl_server* server, uint64_t ticket, spl_request_info* info);
/* Waits for the request and copies its width*height*4 bytes to host memory. */
int spl_server_read(spl_server* server, uint64_t ticket, void* host, uint64_t size);
/* Forgets a finished request and returns its frame to the pool. */
int spl_server_release(spl_server* server, uint64_t ticket);
int spl_server_get_stats(spl_server* server, spl_server_stats* out);

/* ---- parity taps (replace the commented readbacks, renderer.rs:456-464) - */
/* The debug reads below describe view `view` (default 0) of the last batch. */
int spl_debug_select_view(spl_renderer* renderer, uint32_t view);
/* Visible count of the last frame (the indirect draw's instance_count,
 * shaders.wgsl:458-461). */
int spl_debug_visible_count(spl_renderer* renderer, uint64_t* out);
/* Sorted (key, splat index) entries of the last frame: entry buffer A of
 * renderer.rs:395 in CPU mode.  Writes min(capacity, visible) entries. */
int spl_debug_read_entries(spl_renderer* renderer, uint32_t* keys, uint32_t* indices,
                           uint64_t capacity, uint64_t* count);
/* The last frame's per-splat vertex outputs in draw order, 10 floats each:
 * rgb, alpha, affine col0.xy, col1.xy, translation.xy. */
int spl_debug_read_records(spl_renderer* renderer, float* records, uint64_t capacity, uint64_t* count);
/* When enabled, the next frames also store each visible splat's contour
 * conic M (shaders.wgsl:165) in FP64: xx, yy, zz, xy, xz, yz, in draw order. */
int spl_debug_enable_conics(spl_renderer* renderer, int enable);
int spl_debug_read_conics(spl_renderer* renderer, double* conics, uint64_t capacity, uint64_t* count);
/* Number of (splat, screen tile) pairs of the last frame. */
int spl_debug_pair_count(spl_renderer* renderer, uint64_t* out);

/* Per screen tile of the last frame: its (splat, tile) list length and how
 * many of those entries the rasteriser consumed before every pixel of the
 * tile was opaque (2 x u32 per tile, row-major tiles of 16x16 pixels). */
int spl_debug_enable_tile_stats(spl_renderer* renderer, int enable);
int spl_debug_read_tile_stats(spl_renderer* renderer, uint32_t* out, uint64_t capacity, uint64_t* count);

/* ---- timing taps for the harness ---------------------------------------- */
/* Per-stage device time per frame in milliseconds, averaged over the batches
 * rendered since timing was enabled (HIP events on the renderer's stream
 * around every stage of every batch, divided by the batch's view count):
 * cull, digit histograms, depth sort, project, bin, pair sort, tile ranges,
 * raster (8 values).  While timing is enabled batches run as direct
 * launches rather than graph replays. */
int spl_debug_enable_timing(spl_renderer* renderer, int enable);
int spl_debug_stage_times(spl_renderer* renderer, float* ms, uint32_t capacity);

#ifdef __cplusplus
}
#endif
#endif
