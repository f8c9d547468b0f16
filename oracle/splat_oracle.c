/*
 * splat_oracle.c -- CPU restatement of splatter's per-frame splat pipeline.
 * TEST INFRASTRUCTURE ONLY (the checker, never the product); see
 * splat_oracle.h for the pins of each function.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off).
 */
#include "splat_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---------------------------------------------------------------------- */
/* scene.rs                                                                */
/* ---------------------------------------------------------------------- */

#define MAX_SIZE_VARIANCE 5.0f /* scene.rs:13 */

/* scene.rs:225-249 (load_chunk decode loop).  Rotor::signum comes from
 * geometric_algebra 0.3.0 (not vendored): restated as r / sqrt(r.r) with the
 * squared magnitude summed left to right -- parity unpinned at that
 * third-party boundary. */
void orc_decode_serialized(const orc_serialized_splat* in, size_t count, orc_splat* out) {
    for (size_t i = 0; i < count; ++i) {
        const orc_serialized_splat* s = &in[i];
        orc_splat* d = &out[i];
        const float* r = s->rotation;
        float sq = ((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]) + r[3] * r[3];
        float inv = 1.0f / sqrtf(sq);
        for (int k = 0; k < 4; ++k) d->rotation[k] = r[k] * inv;
        for (int k = 0; k < 3; ++k) d->center[k] = s->center[k];
        d->padding = 0.0f;
        float average = ((s->scale[0] + s->scale[1]) + s->scale[2]) / 3.0f;
        for (int k = 0; k < 3; ++k)
            d->scale[k] = expf(fminf(fmaxf(s->scale[k], average - MAX_SIZE_VARIANCE),
                                     average + MAX_SIZE_VARIANCE));
        d->alpha = 1.0f / (1.0f + expf(-s->alpha));
        memcpy(d->sh, s->color, sizeof(d->sh));
    }
}

/* scene.rs:157-192 Scene::new_dev_test. */
void orc_dev_test_scene(orc_splat out[3]) {
    memset(out, 0, 3 * sizeof(orc_splat));
    for (int i = 0; i < 3; ++i) {
        out[i].rotation[0] = 1.0f;
        out[i].center[0] = i == 0 ? 1.0f : 0.0f;
        out[i].center[1] = i == 1 ? 1.0f : 0.0f;
        out[i].center[2] = i == 2 ? 1.0f : 0.0f;
        out[i].scale[0] = 0.2f;
        out[i].scale[1] = 0.1f;
        out[i].scale[2] = 0.05f;
        out[i].alpha = 1.0f;
        out[i].sh[0] = i == 0 ? 1.0f : 0.0f;
        out[i].sh[1] = i == 1 ? 1.0f : 0.0f;
        out[i].sh[2] = i == 2 ? 1.0f : 0.0f;
    }
}

/* ---------------------------------------------------------------------- */
/* utils.rs / renderer.rs camera                                           */
/* ---------------------------------------------------------------------- */

/* utils.rs:45-55 */
void orc_perspective_projection(float width, float height, float near_, float far_, float out[16]) {
    float denominator = 1.0f / (near_ - far_);
    memset(out, 0, 16 * sizeof(float));
    out[0] = 1.0f / width;
    out[5] = 1.0f / height;
    out[10] = -far_ * denominator;
    out[11] = 1.0f;
    out[14] = (near_ * far_) * denominator;
}

/* utils.rs:58-65: column j = ((a0*b[j][0] + a1*b[j][1]) + a2*b[j][2]) + a3*b[j][3] */
void orc_mat4_multiplication(const float a[16], const float b[16], float out[16]) {
    float tmp[16];
    for (int j = 0; j < 4; ++j)
        for (int r = 0; r < 4; ++r)
            tmp[4 * j + r] = ((a[r] * b[4 * j + 0] + a[4 + r] * b[4 * j + 1]) +
                              a[8 + r] * b[4 * j + 2]) + a[12 + r] * b[4 * j + 3];
    memcpy(out, tmp, sizeof(tmp));
}

/* utils.rs:68-70 */
void orc_mat4_transform(const float a[16], const float b[4], float out[4]) {
    for (int r = 0; r < 4; ++r)
        out[r] = ((a[r] * b[0] + a[4 + r] * b[1]) + a[8 + r] * b[2]) + a[12 + r] * b[3];
}

/* renderer.rs:364-370,397-409.  field_of_view_y replaces the hard-coded
 * PI*0.5 (renderer.rs:366) and view_width = aspect * view_height replaces
 * aspect / view_height (renderer.rs:368); at the reference's 90 degrees
 * tanf gives exactly 1.0f so both are bit-identical there. */
void orc_camera_uniforms(const float camera_matrix[16], const float view_matrix[16],
                         uint32_t width, uint32_t height, float field_of_view_y,
                         float frustum_culling_tolerance, float ellipse_margin,
                         float splat_scale, orc_uniforms* out) {
    float view_height = tanf(field_of_view_y * 0.5f);
    float view_width = ((float)width / (float)height) * view_height;
    float projection[16];
    orc_perspective_projection(view_width, view_height, 1.0f, 1000.0f, projection);
    memcpy(out->camera_matrix, camera_matrix, 16 * sizeof(float));
    memcpy(out->view_matrix, view_matrix, 16 * sizeof(float));
    orc_mat4_multiplication(projection, view_matrix, out->view_projection_matrix);
    out->view_size[0] = view_width;
    out->view_size[1] = view_height;
    out->image_size[0] = width;
    out->image_size[1] = height;
    out->frustum_culling_tolerance = frustum_culling_tolerance;
    out->ellipse_size_bias = (0.2f * view_width) / (float)width;
    out->ellipse_margin = ellipse_margin;
    out->splat_scale = splat_scale;
}

/* ---------------------------------------------------------------------- */
/* renderer.rs:372-396 CPU depth-sort fallback                             */
/* ---------------------------------------------------------------------- */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Bottom-up stable merge sort on (key, value) by key. */
void orc_stable_sort_pairs(uint32_t* keys, uint32_t* values, size_t n, uint32_t* scratch) {
    uint32_t *sk = keys, *sv = values, *dk = scratch, *dv = scratch + n;
    for (size_t width = 1; width < n; width *= 2) {
        for (size_t lo = 0; lo < n; lo += 2 * width) {
            size_t mid = lo + width < n ? lo + width : n;
            size_t hi = lo + 2 * width < n ? lo + 2 * width : n;
            size_t i = lo, j = mid, o = lo;
            while (i < mid && j < hi) {
                if (sk[j] < sk[i]) { dk[o] = sk[j]; dv[o++] = sv[j++]; }
                else { dk[o] = sk[i]; dv[o++] = sv[i++]; }
            }
            while (i < mid) { dk[o] = sk[i]; dv[o++] = sv[i++]; }
            while (j < hi) { dk[o] = sk[j]; dv[o++] = sv[j++]; }
        }
        uint32_t* t;
        t = sk; sk = dk; dk = t;
        t = sv; sv = dv; dv = t;
    }
    if (sk != keys) {
        memcpy(keys, sk, n * sizeof(uint32_t));
        memcpy(values, sv, n * sizeof(uint32_t));
    }
}

size_t orc_cpu_depth_sort(const float* positions, size_t count, const float vp[16],
                          float tol, uint32_t* keys, uint32_t* indices) {
    size_t v = 0;
    for (size_t i = 0; i < count; ++i) {
        float p[4] = {positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 1.0f};
        float h[4];
        orc_mat4_transform(vp, p, h);
        float r = 1.0f / h[3];
        float c0 = h[0] * r, c1 = h[1] * r, c2 = h[2] * r;
        if (fabsf(c0) < tol && fabsf(c1) < tol && fabsf(c2 - 0.5f) < 0.5f) {
            keys[v] = f2u(c2);
            indices[v] = (uint32_t)i;
            ++v;
        }
    }
    uint32_t* scratch = (uint32_t*)malloc(2 * (v ? v : 1) * sizeof(uint32_t));
    orc_stable_sort_pairs(keys, indices, v, scratch);
    free(scratch);
    return v;
}

/* shaders.wgsl:54-61 worldToClipSpace + isInFrustum, used at vertex time
 * (shaders.wgsl:521-524) when DepthSorting::None. */
size_t orc_gpu_predicate_cull(const float* positions, size_t count, const float vp[16],
                              float tol, uint32_t* indices) {
    size_t v = 0;
    for (size_t i = 0; i < count; ++i) {
        float p[4] = {positions[3 * i], positions[3 * i + 1], positions[3 * i + 2], 1.0f};
        float h[4];
        orc_mat4_transform(vp, p, h);
        float w = h[3] + 0.0000001f;
        float c0 = h[0] / w, c1 = h[1] / w, c2 = h[2] / w;
        if (fabsf(c0) < tol && fabsf(c1) < tol && fabsf(c2 - 0.5f) < 0.5f) indices[v++] = (uint32_t)i;
    }
    return v;
}

/* ---------------------------------------------------------------------- */
/* shaders.wgsl:248-294 spherical harmonics (f32)                          */
/* ---------------------------------------------------------------------- */

static const float kShc[16] = {
    0.28209479177387814f, -0.4886025119029199f, 0.4886025119029199f, -0.4886025119029199f,
    1.0925484305920792f, -1.0925484305920792f, 0.31539156525252005f, -1.0925484305920792f,
    0.5462742152960396f, -0.5900435899266435f, 2.890611442640554f, -0.4570457994644658f,
    0.3731763325901154f, -0.4570457994644658f, 1.445305721320277f, -0.5900435899266435f,
};

float orc_sh_constant(int k) { return kShc[k]; }

void orc_sh_color(const float d[3], const float sh[48], int order, float out[3]) {
    float x = d[0], y = d[1], z = d[2];
    float xx = x * x, yy = y * y, zz = z * z;
    for (int c = 0; c < 3; ++c) {
        float col = 0.5f;
        col = col + kShc[0] * sh[c];
        if (order > 0) {
            col = col + (kShc[1] * sh[3 + c]) * y;
            col = col + (kShc[2] * sh[6 + c]) * z;
            col = col + (kShc[3] * sh[9 + c]) * x;
        }
        if (order > 1) {
            col = col + ((kShc[4] * sh[12 + c]) * x) * y;
            col = col + ((kShc[5] * sh[15 + c]) * y) * z;
            col = col + ((kShc[6] * sh[18 + c]) * ((2.0f * zz - xx) - yy));
            col = col + ((kShc[7] * sh[21 + c]) * x) * z;
            col = col + (kShc[8] * sh[24 + c]) * (xx - yy);
        }
        if (order > 2) {
            col = col + ((kShc[9] * sh[27 + c]) * y) * (3.0f * xx - yy);
            col = col + (((kShc[10] * sh[30 + c]) * x) * y) * z;
            col = col + ((kShc[11] * sh[33 + c]) * y) * ((4.0f * zz - xx) - yy);
            col = col + ((kShc[12] * sh[36 + c]) * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
            col = col + ((kShc[13] * sh[39 + c]) * x) * ((4.0f * zz - xx) - yy);
            col = col + ((kShc[14] * sh[42 + c]) * z) * (xx - yy);
            col = col + ((kShc[15] * sh[45 + c]) * x) * (xx - 3.0f * yy);
        }
        out[c] = col;
    }
}

/* ---------------------------------------------------------------------- */
/* projection (shaders.wgsl:63-246, 530-547) in f64 and f32                */
/* ---------------------------------------------------------------------- */

#define REAL double
#define SUFFIX _f64
#define SQRT sqrt
#define FABS fabs
#include "project.inc"
#undef REAL
#undef SUFFIX
#undef SQRT
#undef FABS

#define REAL float
#define SUFFIX _f32
#define SQRT sqrtf
#define FABS fabsf
#include "project.inc"
#undef REAL
#undef SUFFIX
#undef SQRT
#undef FABS

void orc_project_f64(const orc_splat* splat, const orc_uniforms* u, int sh_order,
                     int use_covariance_for_scale, orc_record* out, double* conic_out) {
    project_f64(splat, u, sh_order, use_covariance_for_scale, out, conic_out);
}

void orc_project_f32(const orc_splat* splat, const orc_uniforms* u, int sh_order,
                     int use_covariance_for_scale, orc_record* out, float* conic_out) {
    project_f32(splat, u, sh_order, use_covariance_for_scale, out, conic_out);
}

void orc_contour_f64(const double scale[3], const double rotation[4], const double translation[3],
                     const orc_uniforms* u, double M[9]) {
    double m[3][3];
    contour_f64(scale, rotation, translation, u, m);
    memcpy(M, m, sizeof(m));
}

/* ---------------------------------------------------------------------- */
/* fragment (shaders.wgsl:580-592) + fixed-function raster and blend       */
/* ---------------------------------------------------------------------- */

/* The fragment's Gaussian e^(-|tc|^2 / 2) (shaders.wgsl:586) is evaluated as
 * 2^-(|S tc|^2) with S = sqrt(log2(e) / 2) folded into each splat's inverse
 * map (prepare below). */
#define ORC_GAUSS_SCALE 0.84932180f

/* 2^max(y, -125) for y <= 0: y = k + f with k = rint(y), |f| <= 1/2, the
 * degree-6 Taylor polynomial of e^(f ln 2) in fused Horner form, then ldexp
 * by k (exact: the result stays normal).  Only correctly rounded IEEE
 * operations (fma, rint, ldexp), so it rounds identically on any IEEE-754
 * machine.  Below -8 every alpha is under the 1/255 threshold. */
float orc_exp2(float y) {
    y = fmaxf(y, -125.0f);
    float k = rintf(y);
    float f = y - k;
    /* relative minimax polynomial of 2^f on [-1/2, 1/2], degree 5, constant
     * term 1 (exact at integers); within 3 ulp of libm exp2 over [-125, 0] */
    float p = 1.32764713e-3f;
    p = fmaf(p, f, 9.67554189e-3f);
    p = fmaf(p, f, 5.55071309e-2f);
    p = fmaf(p, f, 2.40221202e-1f);
    p = fmaf(p, f, 6.93146944e-1f);
    p = fmaf(p, f, 1.0f);
    return ldexpf(p, (int)k);
}

static inline uint8_t q8(float x) {
    x = fminf(fmaxf(x, 0.0f), 1.0f);
    return (uint8_t)floorf(x * 255.0f + 0.5f);
}

typedef struct {
    float i00, i01, i10, i11; /* inverse of the 2x2 affine part */
    float tx, ty;             /* translation */
    float half;               /* coverage half-extent: margin (unaligned) or radius*margin */
    float r, g, b, a;
    /* The same map in integer pixel coordinates (px, py): the pixel centre
     * is q = (sx px + ox, sy py + oy), so with e = o - t
     *   u = au px + bu py + cu,  v = av px + bv py + cv   (S * texture coordinate)
     *   dx = sx px + ex,         dy = sy py + ey          (q - t, view plane) */
    float au, bu, cu, av, bv, cv;
    float sx, sy, ex, ey;
    int valid;
} orc_prepared;

/* Viewport map of pixel centres: q = (sx px + ox, sy py + oy) (inverse of the
 * viewport transform of clip positions (T*(tc,1))/view_size, shaders.wgsl:561). */
static void pixel_map(const orc_uniforms* u, float* sx, float* ox, float* sy, float* oy) {
    float W = (float)u->image_size[0], H = (float)u->image_size[1];
    float vw = u->view_size[0], vh = u->view_size[1];
    *sx = (2.0f * vw) / W;
    *ox = vw / W - vw;
    *sy = -(2.0f * vh) / H;
    *oy = vh - vh / H;
}

/* shaders.wgsl:554-571: the rectangle's inverse map.  The fragment's
 * interpolated texture coordinate equals inverse(A) * (q - t) because the
 * vertex positions are an affine image of the texture coordinates. */
static void prepare(const orc_record* rec, const orc_uniforms* u, int unaligned, orc_prepared* p) {
    float margin = u->ellipse_margin;
    float A00 = rec->affine[0], A10 = rec->affine[1], A01 = rec->affine[2], A11 = rec->affine[3];
    p->valid = 1;
    for (int i = 0; i < 6; ++i)
        if (!isfinite(rec->affine[i])) p->valid = 0;
    float det = A00 * A11 - A10 * A01;
    float rs = (1.0f / det) * ORC_GAUSS_SCALE; /* inverse map scaled by S */
    p->i00 = A11 * rs;
    p->i10 = -A10 * rs;
    p->i01 = -A01 * rs;
    p->i11 = A00 * rs;
    p->tx = rec->affine[4];
    p->ty = rec->affine[5];
    if (unaligned) {
        p->half = margin * ORC_GAUSS_SCALE; /* |tc| < margin in scaled coordinates */
    } else {
        float d0 = A00 * A00 + A10 * A10, d1 = A01 * A01 + A11 * A11;
        p->half = sqrtf(fmaxf(d0, d1)) * margin;
    }
    p->r = rec->color[0]; p->g = rec->color[1]; p->b = rec->color[2]; p->a = rec->color[3];
    float ox, oy;
    pixel_map(u, &p->sx, &ox, &p->sy, &oy);
    p->ex = ox - p->tx;
    p->ey = oy - p->ty;
    p->au = p->i00 * p->sx;
    p->bu = p->i01 * p->sy;
    p->cu = fmaf(p->i00, p->ex, p->i01 * p->ey);
    p->av = p->i10 * p->sx;
    p->bv = p->i11 * p->sy;
    p->cv = fmaf(p->i10, p->ex, p->i11 * p->ey);
}

/* One fragment: coverage, alpha (shaders.wgsl:585-589) and blend
 * (renderer.rs:260-271: colour = src*DstAlpha + dst, alpha = dst*(1-srcA)). */
static inline void shade(const orc_prepared* p, int unaligned, float px, float py,
                         int blend_mode, float* st) {
    float u, v, h = p->half;
    int inside;
    if (unaligned) {
        u = fmaf(p->au, px, fmaf(p->bu, py, p->cu)); /* S * texture coordinate */
        v = fmaf(p->av, px, fmaf(p->bv, py, p->cv));
        inside = fmaxf(fabsf(u), fabsf(v)) < h;
    } else {
        float dx = fmaf(p->sx, px, p->ex), dy = fmaf(p->sy, py, p->ey);
        u = fmaf(p->i00, dx, p->i01 * dy);
        v = fmaf(p->i10, dx, p->i11 * dy);
        inside = fmaxf(fabsf(dx), fabsf(dy)) < h;
    }
    if (!inside) return;
    float alpha = p->a * orc_exp2(-fmaf(u, u, v * v));
    if (!(alpha >= 1.0f / 255.0f)) return;
    float sr = p->r * alpha, sg = p->g * alpha, sb = p->b * alpha;
    if (blend_mode == ORC_BLEND_UNORM8) {
        /* st holds the stored unorm8 codes C, T as floats 0..255.  The ideal
         * unorm blender stores rint(255 * (s * T/255 + C/255)) =
         * rint(s * T + C) with s clamped to [0,1] and the sum to 255, and
         * rint(255 * (T/255) * (1 - a)) = rint(T * (1 - a)) for alpha. */
        float t = st[3];
        float sa = fminf(fmaxf(alpha, 0.0f), 1.0f);
        float s[3] = {fminf(fmaxf(sr, 0.0f), 1.0f), fminf(fmaxf(sg, 0.0f), 1.0f), fminf(fmaxf(sb, 0.0f), 1.0f)};
        for (int c = 0; c < 3; ++c) st[c] = fminf(rintf(fmaf(s[c], t, st[c])), 255.0f);
        st[3] = rintf(fmaf(-sa, t, t));
    } else {
        float da = st[3];
        st[0] = fmaf(sr, da, st[0]);
        st[1] = fmaf(sg, da, st[1]);
        st[2] = fmaf(sb, da, st[2]);
        st[3] = fmaf(-alpha, da, da);
    }
}

void orc_raster(const orc_record* records, size_t count, const orc_uniforms* u,
                int unaligned, int blend_mode, uint8_t* image, int brute_force) {
    int W = (int)u->image_size[0], H = (int)u->image_size[1];
    float vw = u->view_size[0], vh = u->view_size[1];
    float margin = u->ellipse_margin;
    float* state = (float*)malloc((size_t)W * H * 4 * sizeof(float));
    for (size_t i = 0; i < (size_t)W * H; ++i) {
        state[4 * i + 0] = 0.0f;
        state[4 * i + 1] = 0.0f;
        state[4 * i + 2] = 0.0f;
        state[4 * i + 3] = blend_mode == ORC_BLEND_UNORM8 ? 255.0f : 1.0f;
    }
    for (size_t s = 0; s < count; ++s) {
        orc_prepared p;
        prepare(&records[s], u, unaligned, &p);
        if (!p.valid) continue;
        int x0 = 0, x1 = W - 1, y0 = 0, y1 = H - 1;
        if (!brute_force) {
            /* conservative pixel box of the covered region, in double */
            double ex, ey;
            if (unaligned) {
                ex = (fabs((double)records[s].affine[0]) + fabs((double)records[s].affine[2])) * margin;
                ey = (fabs((double)records[s].affine[1]) + fabs((double)records[s].affine[3])) * margin;
            } else {
                ex = ey = p.half;
            }
            double cx = ((double)p.tx / vw + 1.0) * 0.5 * W, cy = (1.0 - (double)p.ty / vh) * 0.5 * H;
            double rx = ex / vw * 0.5 * W + 2.0, ry = ey / vh * 0.5 * H + 2.0;
            double fx0 = floor(cx - rx), fx1 = ceil(cx + rx), fy0 = floor(cy - ry), fy1 = ceil(cy + ry);
            if (fx1 < 0 || fy1 < 0 || fx0 > W - 1 || fy0 > H - 1) continue;
            x0 = fx0 < 0 ? 0 : (int)fx0;
            y0 = fy0 < 0 ? 0 : (int)fy0;
            x1 = fx1 > W - 1 ? W - 1 : (int)fx1;
            y1 = fy1 > H - 1 ? H - 1 : (int)fy1;
        }
        for (int py = y0; py <= y1; ++py) {
            for (int px = x0; px <= x1; ++px) {
                float* st = &state[4 * ((size_t)py * W + px)];
                /* transmittance 0: every later fragment adds src * 0 */
                if (st[3] == 0.0f) continue;
                shade(&p, unaligned, (float)px, (float)py, blend_mode, st);
            }
        }
    }
    for (size_t i = 0; i < (size_t)W * H; ++i)
        for (int c = 0; c < 4; ++c)
            image[4 * i + c] = blend_mode == ORC_BLEND_UNORM8 ? (uint8_t)state[4 * i + c] : q8(state[4 * i + c]);
    free(state);
}

size_t orc_render_frame(const orc_splat* splats, size_t count, const orc_uniforms* u,
                        int sh_order, int use_covariance_for_scale,
                        int use_unaligned_rectangles, int blend_mode,
                        uint32_t* keys, uint32_t* indices, orc_record* records,
                        uint8_t* image) {
    float* pos = (float*)malloc((count ? count : 1) * 3 * sizeof(float));
    for (size_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) pos[3 * i + k] = splats[i].center[k];
    size_t v = orc_cpu_depth_sort(pos, count, u->view_projection_matrix, u->frustum_culling_tolerance,
                                  keys, indices);
    free(pos);
    for (size_t r = 0; r < v; ++r)
        orc_project_f64(&splats[indices[r]], u, sh_order, use_covariance_for_scale, &records[r], NULL);
    if (image) orc_raster(records, v, u, use_unaligned_rectangles, blend_mode, image, 0);
    return v;
}

/* ---------------------------------------------------------------------- */
/* CPU baseline timing                                                     */
/* ---------------------------------------------------------------------- */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int cmp_double(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

double orc_time_cpu_depth_sort(const float* positions, size_t count, const float vp[16],
                               float tol, int repeats, size_t* visible_out) {
    uint32_t* keys = (uint32_t*)malloc((count ? count : 1) * sizeof(uint32_t));
    uint32_t* idx = (uint32_t*)malloc((count ? count : 1) * sizeof(uint32_t));
    double* t = (double*)malloc((size_t)(repeats > 0 ? repeats : 1) * sizeof(double));
    size_t v = 0;
    for (int w = 0; w < 1; ++w) v = orc_cpu_depth_sort(positions, count, vp, tol, keys, idx);
    for (int i = 0; i < repeats; ++i) {
        double t0 = now_s();
        v = orc_cpu_depth_sort(positions, count, vp, tol, keys, idx);
        t[i] = now_s() - t0;
    }
    qsort(t, (size_t)repeats, sizeof(double), cmp_double);
    double med = repeats > 0 ? t[repeats / 2] : 0.0;
    if (visible_out) *visible_out = v;
    free(keys);
    free(idx);
    free(t);
    return med;
}

/* ---------------------------------------------------------------------- */
/* CPU baseline on all host cores: renderer.rs:372-396 with the transform  */
/* split over threads and std's stable sort replaced by a parallel stable  */
/* merge sort.  Same visible set and order as orc_cpu_depth_sort.          */
/* ---------------------------------------------------------------------- */

typedef struct {
    const float* positions;
    size_t lo, hi;
    const float* vp;
    float tol;
    uint32_t* keys;  /* this thread's visible entries, in splat order */
    uint32_t* idx;
    size_t count;
} orc_cull_job;

static void* cull_range(void* arg) {
    orc_cull_job* j = (orc_cull_job*)arg;
    size_t v = 0;
    for (size_t i = j->lo; i < j->hi; ++i) {
        float p[4] = {j->positions[3 * i], j->positions[3 * i + 1], j->positions[3 * i + 2], 1.0f};
        float h[4];
        orc_mat4_transform(j->vp, p, h);
        float r = 1.0f / h[3];
        float c0 = h[0] * r, c1 = h[1] * r, c2 = h[2] * r;
        if (fabsf(c0) < j->tol && fabsf(c1) < j->tol && fabsf(c2 - 0.5f) < 0.5f) {
            j->keys[v] = f2u(c2);
            j->idx[v] = (uint32_t)i;
            ++v;
        }
    }
    j->count = v;
    return NULL;
}

/* Merge-path co-rank: how many of the first k outputs of the stable merge of
 * a[0..m) and b[0..n) come from a (equal keys: a's first). */
static size_t co_rank(size_t k, const uint32_t* a, size_t m, const uint32_t* b, size_t n) {
    size_t i = k < m ? k : m, j = k - i;
    size_t i_lo = k > n ? k - n : 0, j_lo = k > m ? k - m : 0;
    for (;;) {
        if (i > 0 && j < n && a[i - 1] > b[j]) {
            size_t d = (i - i_lo + 1) / 2;
            j_lo = j;
            i -= d;
            j += d;
        } else if (j > 0 && i < m && b[j - 1] >= a[i]) {
            size_t d = (j - j_lo + 1) / 2;
            i_lo = i;
            i += d;
            j -= d;
        } else {
            return i;
        }
    }
}

typedef struct {
    const uint32_t *ak, *av, *bk, *bv;
    size_t m, n;
    uint32_t *ok, *ov;
    size_t out_lo, out_hi;   /* slice of the merged output this job writes */
    uint32_t* sort_k;        /* or, when set: stable-sort this chunk in place */
    uint32_t* sort_v;
    uint32_t* scratch;
    size_t sort_n;
} orc_merge_job;

static void* merge_slice(void* arg) {
    orc_merge_job* j = (orc_merge_job*)arg;
    if (j->sort_k) {
        orc_stable_sort_pairs(j->sort_k, j->sort_v, j->sort_n, j->scratch);
        return NULL;
    }
    size_t i = co_rank(j->out_lo, j->ak, j->m, j->bk, j->n);
    size_t jj = j->out_lo - i;
    for (size_t o = j->out_lo; o < j->out_hi; ++o) {
        if (jj < j->n && (i >= j->m || j->bk[jj] < j->ak[i])) {
            j->ok[o] = j->bk[jj];
            j->ov[o] = j->bv[jj++];
        } else {
            j->ok[o] = j->ak[i];
            j->ov[o] = j->av[i++];
        }
    }
    return NULL;
}

static void run_jobs(void* (*fn)(void*), void* jobs, size_t stride, int count) {
    pthread_t th[ORC_MAX_THREADS];
    int started[ORC_MAX_THREADS];
    for (int t = 0; t < count; ++t)
        started[t] = pthread_create(&th[t], NULL, fn, (char*)jobs + (size_t)t * stride) == 0;
    for (int t = 0; t < count; ++t) {
        if (started[t]) pthread_join(th[t], NULL);
        else fn((char*)jobs + (size_t)t * stride);  /* no thread to be had: run it here */
    }
}

size_t orc_cpu_depth_sort_threads(const float* positions, size_t count, const float vp[16], float tol,
                                  int threads, uint32_t* keys, uint32_t* indices) {
    if (threads < 1) threads = 1;
    if (threads > ORC_MAX_THREADS) threads = ORC_MAX_THREADS;
    orc_cull_job cj[ORC_MAX_THREADS];
    uint32_t* tk = (uint32_t*)malloc(2 * (count ? count : 1) * sizeof(uint32_t));
    uint32_t* ti = tk + (count ? count : 1);
    for (int t = 0; t < threads; ++t) {
        cj[t].positions = positions;
        cj[t].lo = count * (size_t)t / (size_t)threads;
        cj[t].hi = count * (size_t)(t + 1) / (size_t)threads;
        cj[t].vp = vp;
        cj[t].tol = tol;
        cj[t].keys = tk + cj[t].lo;
        cj[t].idx = ti + cj[t].lo;
    }
    run_jobs(cull_range, cj, sizeof(orc_cull_job), threads);
    /* compact the per-thread runs in splat order: filter_map's order */
    size_t v = 0, run_lo[ORC_MAX_THREADS + 1];
    for (int t = 0; t < threads; ++t) {
        run_lo[t] = v;
        memcpy(keys + v, cj[t].keys, cj[t].count * sizeof(uint32_t));
        memcpy(indices + v, cj[t].idx, cj[t].count * sizeof(uint32_t));
        v += cj[t].count;
    }
    run_lo[threads] = v;
    /* each thread sorts one chunk, then chunks merge pairwise; every merge
     * round splits its output evenly over the threads by merge path */
    uint32_t* sk = (uint32_t*)malloc(2 * (v ? v : 1) * sizeof(uint32_t));
    uint32_t* sv = sk + (v ? v : 1);
    orc_merge_job mj[ORC_MAX_THREADS];
    memset(mj, 0, sizeof(mj));
    for (int t = 0; t < threads; ++t) {
        mj[t].sort_k = keys + run_lo[t];
        mj[t].sort_v = indices + run_lo[t];
        mj[t].sort_n = run_lo[t + 1] - run_lo[t];
        mj[t].scratch = tk + 2 * run_lo[t];  /* 2 * chunk words, inside the transform buffers */
    }
    run_jobs(merge_slice, mj, sizeof(orc_merge_job), threads);
    uint32_t *ck = keys, *cv = indices, *nk = sk, *nv = sv;
    size_t bounds[ORC_MAX_THREADS + 1];
    int runs = threads;
    for (int t = 0; t <= threads; ++t) bounds[t] = run_lo[t];
    while (runs > 1) {
        int jobs = 0;
        size_t nb[ORC_MAX_THREADS + 1];
        int nr = 0;
        for (int r = 0; r < runs; r += 2) {
            size_t a0 = bounds[r], a1 = bounds[r + 1];
            size_t b1 = r + 1 < runs ? bounds[r + 2] : a1;
            size_t total = b1 - a0;
            int share = threads / ((runs + 1) / 2);
            if (share < 1) share = 1;
            for (int s = 0; s < share; ++s) {
                orc_merge_job* j = &mj[jobs++];
                memset(j, 0, sizeof(*j));
                j->ak = ck + a0; j->av = cv + a0; j->m = a1 - a0;
                j->bk = ck + a1; j->bv = cv + a1; j->n = b1 - a1;
                j->ok = nk + a0; j->ov = nv + a0;
                j->out_lo = total * (size_t)s / (size_t)share;
                j->out_hi = total * (size_t)(s + 1) / (size_t)share;
            }
            nb[nr++] = a0;
        }
        nb[nr] = v;
        run_jobs(merge_slice, mj, sizeof(orc_merge_job), jobs);
        uint32_t* t;
        t = ck; ck = nk; nk = t;
        t = cv; cv = nv; nv = t;
        runs = nr;
        for (int r = 0; r <= nr; ++r) bounds[r] = nb[r];
    }
    if (ck != keys) {
        memcpy(keys, ck, v * sizeof(uint32_t));
        memcpy(indices, cv, v * sizeof(uint32_t));
    }
    free(sk);
    free(tk);
    return v;
}

double orc_time_cpu_depth_sort_threads(const float* positions, size_t count, const float vp[16], float tol,
                                       int threads, int repeats, size_t* visible_out) {
    uint32_t* keys = (uint32_t*)malloc((count ? count : 1) * sizeof(uint32_t));
    uint32_t* idx = (uint32_t*)malloc((count ? count : 1) * sizeof(uint32_t));
    double* t = (double*)malloc((size_t)(repeats > 0 ? repeats : 1) * sizeof(double));
    size_t v = 0;
    for (int w = 0; w < 2; ++w) v = orc_cpu_depth_sort_threads(positions, count, vp, tol, threads, keys, idx);
    for (int i = 0; i < repeats; ++i) {
        double t0 = now_s();
        v = orc_cpu_depth_sort_threads(positions, count, vp, tol, threads, keys, idx);
        t[i] = now_s() - t0;
    }
    qsort(t, (size_t)repeats, sizeof(double), cmp_double);
    double med = repeats > 0 ? t[repeats / 2] : 0.0;
    if (visible_out) *visible_out = v;
    free(keys);
    free(idx);
    free(t);
    return med;
}
