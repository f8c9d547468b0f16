/*
 * splat_oracle.h -- CPU restatement of splatter's per-frame splat pipeline.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * it may be loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, never by the product library or its Python
 * package.  Everything here is plain C99 compiled with -ffp-contract=off so
 * that every float expression rounds exactly as the reference's Rust/WGSL
 * source spells it (Rust never fuses a*b+c; x86-64 SSE is IEEE-754).
 *
 * Reference files (Lichtso/splatter):
 *   src/scene.rs     record layout + load_chunk decode
 *   src/utils.rs     camera / projection / mat4 helpers
 *   src/renderer.rs  per-frame driver, CPU depth-sort fallback, blend state
 *   src/shaders.wgsl projection, spherical harmonics, rectangle, fragment
 *
 * Parity pins (what checks each function; see tests/test_oracle_*.py):
 *   orc_decode_serialized      closed forms (unit rotor, exp/sigmoid), dev-test scene
 *   orc_camera_uniforms        closed-form perspective, camera*view = I
 *   orc_cpu_depth_sort         closed-form z_ndc, dev-test known answer (ties),
 *                              brute force on tiny N, sortedness/permutation
 *   orc_project_*              sphere closed forms, silhouette tangency p^T M p = 0,
 *                              known-ellipse round trip, FP32 vs FP64 agreement
 *   orc_sh_color               closed-form constants, quadrature orthonormality
 *   orc_exp2                   libm exp2 within 4 ulp over [-125, 0], exact at integers
 *   orc_raster                 isolated-splat analytic footprint, two-layer
 *                              compositing closed form, brute-force bbox check
 * The third-party arithmetic below the reference (geometric_algebra 0.3.0
 * Motor/Rotor algebra, naga/wgpu float semantics, the Vulkan rasteriser's
 * coverage and blender rounding) is restated, not pinned: "parity unpinned"
 * at those boundaries.
 */
#ifndef SPLAT_ORACLE_H
#define SPLAT_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scene.rs:17-24 -- 62 f32 = 248 B, 3DGS PLY vertex order. */
typedef struct {
    float center[3];
    float n[3];
    float color[48];
    float alpha;
    float scale[3];
    float rotation[4];
} orc_serialized_splat;

/* scene.rs:32 / shaders.wgsl:28-35 -- 60 f32 = 240 B. */
typedef struct {
    float rotation[4];
    float center[3];
    float padding;
    float scale[3];
    float alpha;
    float sh[48];
} orc_splat;

/* renderer.rs:49-61 / shaders.wgsl:1-11 -- 224 B.  Matrices column-major,
 * column c = m[4c .. 4c+3] = (x, y, z, w). */
typedef struct {
    float camera_matrix[16];
    float view_matrix[16];
    float view_projection_matrix[16];
    float view_size[2];
    uint32_t image_size[2];
    float frustum_culling_tolerance;
    float ellipse_size_bias;
    float ellipse_margin;
    float splat_scale;
} orc_uniforms;

/* The vertex stage outputs of shaders.wgsl:532,543-547 computed once per
 * splat: flat colour (rgb, alpha) and the 2x3 affine map from the unit
 * rectangle's texture coordinates to view-plane coordinates
 * (col0, col1, translation). 40 B. */
typedef struct {
    float color[4];
    float affine[6];
} orc_record;

/* Blend emulation of the render target. */
enum { ORC_BLEND_UNORM8 = 0, ORC_BLEND_F32 = 1 };

/* ---- scene.rs ---------------------------------------------------------- */
void orc_decode_serialized(const orc_serialized_splat* in, size_t count, orc_splat* out);
void orc_dev_test_scene(orc_splat out[3]);

/* ---- utils.rs / renderer.rs camera ------------------------------------ */
void orc_perspective_projection(float width, float height, float near_, float far_, float out[16]);
void orc_mat4_multiplication(const float a[16], const float b[16], float out[16]);
void orc_mat4_transform(const float a[16], const float b[4], float out[4]);
void orc_camera_uniforms(const float camera_matrix[16], const float view_matrix[16],
                         uint32_t width, uint32_t height, float field_of_view_y,
                         float frustum_culling_tolerance, float ellipse_margin,
                         float splat_scale, orc_uniforms* out);

/* ---- renderer.rs:372-396 CPU depth-sort fallback ----------------------- */
/* positions: xyz per splat.  Writes the visible (key, index) entries sorted
 * stably by key; returns the visible count. */
size_t orc_cpu_depth_sort(const float* positions, size_t count, const float vp[16],
                          float frustum_culling_tolerance, uint32_t* keys, uint32_t* indices);
/* The unsorted variant (index order, keys only) used by DepthSorting::None's
 * vertex-time cull, shaders.wgsl:54-61,521-524 (divides by w + 1e-7). */
size_t orc_gpu_predicate_cull(const float* positions, size_t count, const float vp[16],
                              float frustum_culling_tolerance, uint32_t* indices);

/* Stable bottom-up merge sort of (key, value) pairs by key: the order
 * Rust's slice::sort_by leaves (renderer.rs:394). scratch holds 2*count u32. */
void orc_stable_sort_pairs(uint32_t* keys, uint32_t* values, size_t count, uint32_t* scratch);

/* ---- shaders.wgsl projection ------------------------------------------ */
/* FP64 restatement.  conic_out (optional) receives the 6 unique entries of
 * the contour conic M: xx, yy, zz, xy, xz, yz. */
void orc_project_f64(const orc_splat* splat, const orc_uniforms* u, int sh_order,
                     int use_covariance_for_scale, orc_record* out, double* conic_out);
/* FP32 restatement of the same functions, op for op as the WGSL spells them. */
void orc_project_f32(const orc_splat* splat, const orc_uniforms* u, int sh_order,
                     int use_covariance_for_scale, orc_record* out, float* conic_out);
void orc_contour_f64(const double scale[3], const double rotation[4], const double translation[3],
                     const orc_uniforms* u, double M[9]);
void orc_sh_color(const float ray_direction[3], const float sh[48], int sh_order, float out[3]);
float orc_sh_constant(int k);
/* 2^max(y, -125), y <= 0: the fragment's Gaussian (see orc_raster). */
float orc_exp2(float y);

/* ---- fixed-function raster + fragment + blend -------------------------- */
/* records are in draw order (front to back).  image: w*h*4 bytes, channel
 * order R,G,B,A (A = transmittance). */
void orc_raster(const orc_record* records, size_t count, const orc_uniforms* u,
                int use_unaligned_rectangles, int blend_mode, uint8_t* image,
                int brute_force);

/* Whole frame: cull + sort + project + raster.  Returns visible count. */
size_t orc_render_frame(const orc_splat* splats, size_t count, const orc_uniforms* u,
                        int sh_order, int use_covariance_for_scale,
                        int use_unaligned_rectangles, int blend_mode,
                        uint32_t* keys, uint32_t* indices, orc_record* records,
                        uint8_t* image);

/* ---- CPU baseline ------------------------------------------------------ */
/* Times orc_cpu_depth_sort on one core: runs `repeats` times, returns the
 * median seconds. */
double orc_time_cpu_depth_sort(const float* positions, size_t count, const float vp[16],
                               float tol, int repeats, size_t* visible_out);

/* The same path on `threads` host threads (at most ORC_MAX_THREADS): the
 * transform split over contiguous splat ranges, then a parallel stable merge
 * sort (per-thread merge sort of one chunk each, then merge rounds split by
 * merge path).  Output identical to orc_cpu_depth_sort. */
#define ORC_MAX_THREADS 64
size_t orc_cpu_depth_sort_threads(const float* positions, size_t count, const float vp[16], float tol,
                                  int threads, uint32_t* keys, uint32_t* indices);
/* Median seconds of `repeats` runs after two warm-up runs. */
double orc_time_cpu_depth_sort_threads(const float* positions, size_t count, const float vp[16], float tol,
                                       int threads, int repeats, size_t* visible_out);

#ifdef __cplusplus
}
#endif
#endif
