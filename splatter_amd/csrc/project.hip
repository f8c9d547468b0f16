// project.hip -- per visible splat: spherical-harmonics colour (f32), the
// contour conic of its ellipsoid and the rotated rectangle around it (f64).
//
// This is the per-splat half of the reference's vertex stage
// (shaders.wgsl:507-547) evaluated once per visible splat instead of once per
// quad vertex.  Entry c of a view's compacted list writes that view's
// record c: rgba | A00 A10 A01 A11 | tx ty bbox, 48 B, where A = (col0 col1)
// and t are the rectangle's affine map (shaders.wgsl:543-547) in view-plane
// units and bbox is a conservative pixel box of what the rectangle can cover.
//
// The conic and its decomposition run in FP64: the contour path's semi-axes
// (shaders.wgsl:217-237) lose most of their digits in FP32 to the
// cancellation in e = M.zz - t^T M t.  Products and sums follow the WGSL's
// order; in FP64 a product and the sum it feeds are fused (one rounding
// instead of two, 2^-53 relative), divisions by the same value share one
// reciprocal, and reciprocals and square roots are the hardware estimates
// refined by one Newton step (an ulp or two), all far inside the 1e-5 conic
// tolerance.  The
// f32 colour path is not contracted: it matches the reference's f32
// arithmetic operation for operation.
#include "spl_device.h"
#include "spl_internal.h"

namespace spl {

namespace {

// Matrices are WGSL column-major 3x3: m[c][r].

__device__ __forceinline__ double sgn(double x) { return x > 0.0 ? 1.0 : (x < 0.0 ? -1.0 : 0.0); }

// 1/x: the hardware reciprocal and one Newton step (the error squares, to
// within an ulp or two of the division), IEEE results for 0 and infinity.
// About a third of the instructions of a correctly rounded division.
__device__ __forceinline__ double rcp64(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, r, 1.0);
  return e == e ? fma(r, e, r) : r;
}

// sqrt(x): the hardware reciprocal square root y, h = x y and one Newton
// step h + (x - h^2) y / 2; 0 and infinity as IEEE, NaN below 0.
__device__ __forceinline__ double sqrt64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double h = x * y;
  const double s = fma(fma(-h, h, x), 0.5 * y, h);
  return x == 0.0 || x == __builtin_inf() ? x : s;
}

// shaders.wgsl:63-79 quatToMat
__device__ __forceinline__ void quat_to_mat(const double p[4], double m[3][3]) {
#pragma clang fp contract(fast)
  const double s2 = sqrt(2.0);
  const double q0 = p[0] * s2, q1 = p[1] * s2, q2 = p[2] * s2, q3 = p[3] * s2;
  const double yy = q1 * q1, yz = q1 * q2, yw = q1 * q3, yx = q1 * q0;
  const double zz = q2 * q2, zw = q2 * q3, zx = q2 * q0;
  const double ww = q3 * q3, wx = q3 * q0;
  m[0][0] = (1.0 - zz) - ww; m[0][1] = yz + wx; m[0][2] = yw - zx;
  m[1][0] = yz - wx; m[1][1] = (1.0 - yy) - ww; m[1][2] = zw + yx;
  m[2][0] = yw + zx; m[2][1] = zw - yx; m[2][2] = (1.0 - yy) - zz;
}

// shaders.wgsl:127-130: the view-independent half of
// projectedContourOfEllipsoid, transform = quatToMat(rotation) / scale.
__device__ __forceinline__ void contour_transform(const double scale[3], const double rot[4], double T[3][3]) {
  quat_to_mat(rot, T);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double inv = rcp64(scale[c]);
#pragma unroll
    for (int r = 0; r < 3; ++r) T[c][r] = T[c][r] * inv;
  }
}

// shaders.wgsl:125-169 projectedContourOfEllipsoid, given the transform.
// With B = C^T T and lo = T^T (camera - position) the reference forms
// A = (1 - |lo|^2) I + lo lo^T and M = B A B^T; here M = s B B^T + (B lo)(B lo)^T
// with s = 1 - |lo|^2, the same matrix in fewer FP64 operations (its six
// distinct entries only).
__device__ void contour(const double T[3][3], const double pos[3], const Uniforms& u, double M[3][3]) {
#pragma clang fp contract(fast)
  double C[3][3], B[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) C[c][r] = (double)u.camera_matrix[4 * c + r];
  double ro[3], lo[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) ro[i] = (double)u.camera_matrix[12 + i] - pos[i];
#pragma unroll
  for (int c = 0; c < 3; ++c) lo[c] = (ro[0] * T[c][0] + ro[1] * T[c][1]) + ro[2] * T[c][2];
  const double s = ((1.0 - lo[0] * lo[0]) - lo[1] * lo[1]) - lo[2] * lo[2];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) B[c][r] = (C[r][0] * T[c][0] + C[r][1] * T[c][1]) + C[r][2] * T[c][2];
  double bl[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) bl[r] = (B[0][r] * lo[0] + B[1][r] * lo[1]) + B[2][r] * lo[2];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = c; r < 3; ++r) {
      const double bb = (B[0][r] * B[0][c] + B[1][r] * B[1][c]) + B[2][r] * B[2][c];
      M[c][r] = s * bb + bl[r] * bl[c];
      M[r][c] = M[c][r];
    }
}

// shaders.wgsl:101-104: the view-independent half of
// projectedCovarianceOfEllipsoid, transform = quatToMat(rotation) * scale.
__device__ __forceinline__ void covariance_transform(const double scale[3], const double rot[4], double T[3][3]) {
  quat_to_mat(rot, T);
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) T[c][r] = T[c][r] * scale[c];
}

// shaders.wgsl:99-118 projectedCovarianceOfEllipsoid, given the transform:
// cov = W^T W with W = T^T C J.  Only the top-left 2x2 of cov is used
// (extractScaleOfCovariance, shaders.wgsl:240-246) and J's third column and
// two of its entries are zero, so only W's first two columns and three
// entries of cov are formed.
__device__ void covariance2(const double T[3][3], const double pos[3], const Uniforms& u,
                            double& c00, double& c11, double& c01) {
#pragma clang fp contract(fast)
  double C[3][3], G[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) C[c][r] = (double)u.camera_matrix[4 * c + r];
  const float* V = u.view_matrix;
  double vp[3];
#pragma unroll
  for (int r = 0; r < 3; ++r)
    vp[r] = (((double)V[r] * pos[0] + (double)V[4 + r] * pos[1]) + (double)V[8 + r] * pos[2]) + (double)V[12 + r];
  const double iz = rcp64(vp[2]);
  double qx = vp[0] * iz, qy = vp[1] * iz;
  qx = qx < -1.0 ? -1.0 : (qx > 1.0 ? 1.0 : qx);
  qy = qy < -1.0 ? -1.0 : (qy > 1.0 ? 1.0 : qy);
  const double j02 = -(qx * vp[2]) * (iz * iz), j12 = -(qy * vp[2]) * (iz * iz);
  // G = T^T C
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) G[c][r] = (T[r][0] * C[c][0] + T[r][1] * C[c][1]) + T[r][2] * C[c][2];
  double w0[3], w1[3];
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    w0[r] = G[0][r] * iz + G[2][r] * j02;
    w1[r] = G[1][r] * iz + G[2][r] * j12;
  }
  c00 = (w0[0] * w0[0] + w0[1] * w0[1]) + w0[2] * w0[2];
  c11 = (w1[0] * w1[0] + w1[1] * w1[1]) + w1[2] * w1[2];
  c01 = (w0[0] * w1[0] + w0[1] * w1[1]) + w0[2] * w1[2];
}

// shaders.wgsl:248-294 (f32; same evaluation order).  The scene holds each
// coefficient already multiplied by its basis constant (kShConstants, at
// load): every term of the reference starts with that product, so the f32
// values are the same and a (splat, view) saves 48 multiplications.
template <int L>
__device__ __forceinline__ void sh_color(const float d[3], const float* sh, float out[3]) {
  const float x = d[0], y = d[1], z = d[2];
  const float xx = x * x, yy = y * y, zz = z * z;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float col = 0.5f;
    col = col + sh[c];
    if (L > 0) {
      col = col + sh[3 + c] * y;
      col = col + sh[6 + c] * z;
      col = col + sh[9 + c] * x;
    }
    if (L > 1) {
      col = col + (sh[12 + c] * x) * y;
      col = col + (sh[15 + c] * y) * z;
      col = col + (sh[18 + c] * ((2.0f * zz - xx) - yy));
      col = col + (sh[21 + c] * x) * z;
      col = col + sh[24 + c] * (xx - yy);
    }
    if (L > 2) {
      col = col + (sh[27 + c] * y) * (3.0f * xx - yy);
      col = col + ((sh[30 + c] * x) * y) * z;
      col = col + (sh[33 + c] * y) * ((4.0f * zz - xx) - yy);
      col = col + (sh[36 + c] * z) * ((2.0f * zz - 3.0f * xx) - 3.0f * yy);
      col = col + (sh[39 + c] * x) * ((4.0f * zz - xx) - yy);
      col = col + (sh[42 + c] * z) * (xx - yy);
      col = col + (sh[45 + c] * x) * (xx - 3.0f * yy);
    }
    out[c] = col;
  }
}

__device__ __forceinline__ bool finite(float x) { return (__float_as_uint(x) & 0x7F800000u) != 0x7F800000u; }

// Conservative pixel box of what the rectangle can cover, in f64 so that a
// far-off centre with a huge extent cannot cancel away.  Empty when the
// rectangle is not finite or misses the image.
__device__ void pixel_box(float A00, float A10, float A01, float A11, float tx, float ty, const Uniforms& u,
                          bool unaligned, uint32_t& lo, uint32_t& hi) {
  lo = 0xFFFFFFFFu;  // empty: x0 > x1
  hi = 0u;
  if (!(finite(A00) && finite(A10) && finite(A01) && finite(A11) && finite(tx) && finite(ty))) return;
  const double m = (double)u.ellipse_margin;
  double ex, ey;
  if (unaligned) {
    ex = (fabs((double)A00) + fabs((double)A01)) * m;
    ey = (fabs((double)A10) + fabs((double)A11)) * m;
  } else {
    ex = sqrt64(fmax((double)A00 * A00 + (double)A10 * A10, (double)A01 * A01 + (double)A11 * A11)) * m;
    ey = ex;
  }
  const double W = (double)u.image_size[0], H = (double)u.image_size[1];
  const double ivw = rcp64((double)u.view_size[0]), ivh = rcp64((double)u.view_size[1]);
  const double pcx = ((double)tx * ivw + 1.0) * 0.5 * W, pcy = (1.0 - (double)ty * ivh) * 0.5 * H;
  // +1 px: pixel centres sit half a pixel inside the box edges, the rest
  // absorbs rounding of the f32 record and of the coverage test
  const double rx = ex * ivw * 0.5 * W + 1.0, ry = ey * ivh * 0.5 * H + 1.0;
  const double fx0 = floor(pcx - rx), fx1 = ceil(pcx + rx);
  const double fy0 = floor(pcy - ry), fy1 = ceil(pcy + ry);
  if (fx1 >= 0.0 && fy1 >= 0.0 && fx0 <= W - 1.0 && fy0 <= H - 1.0) {
    const uint32_t x0 = (uint32_t)fmax(fx0, 0.0), y0 = (uint32_t)fmax(fy0, 0.0);
    const uint32_t x1 = (uint32_t)fmin(fx1, W - 1.0), y1 = (uint32_t)fmin(fy1, H - 1.0);
    lo = x0 | (y0 << 16);
    hi = x1 | (y1 << 16);
  }
}

}  // namespace

// One record per (visible splat, view).  Entry c of view v's list (its c-th
// visible splat by index) is found from the visibility masks and the cull's
// per-tile prefixes, so records land where the cull put the keys, with no
// per-view index to read.  The block lists the tile's (splat, view) pairs
// group by group (256 splats each, a scan of the mask popcounts) together
// with each pair's entry c, and projects the list once it holds as many
// pairs as LDS allows: every lane projects one pair, lanes do not idle on
// views that do not see their splat even when a frame has one view, and the
// lanes of one splat read its attributes from the same cache lines.
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_project(ProjectLaunch a) {
  constexpr int THREADS = kCullThreads, ITEMS = kCullItems, WAVES = THREADS / kWave, TILE = kCullTile;
  // Pairs are listed for up to 16 views at a time (a larger batch lists its
  // views in halves), so one group of 256 splats lists at most CAP pairs and
  // a pair packs into 16 bits: slot in the tile << 4 | view - first view.
  constexpr int CAP = THREADS * 16;
  static_assert(TILE <= 4096 && kMaxViews <= 32, "pair packing");
  static_assert(16 * WAVES == kWave, "one wave scans the (view, wave) counts");
  __shared__ uint32_t s_off[kMaxViews][ITEMS * WAVES];
  __shared__ uint32_t s_tp[kMaxViews];
  __shared__ uint16_t s_item[CAP];    // slot in tile << 4 | view - first view
  __shared__ uint32_t s_entry[CAP];   // entry c of that view
  __shared__ uint32_t s_tmp[WAVES];
  // A round's records in view-major order (rank r = pairs of earlier views
  // in the round + earlier pairs of the same view), with their destination:
  // within one view the entries of consecutive listed splats are
  // consecutive, so the records leave in contiguous runs.
  __shared__ float4 s_stage[THREADS][3];
  __shared__ uint2 s_dst[THREADS];      // view, entry
  __shared__ uint32_t s_wcnt[16][WAVES];  // pairs per (view, wave), then their view-major offsets
  // the listed views' uniforms: the FP64 part reads them in LDS instead of
  // waiting on global loads in the middle of its chains
  __shared__ Uniforms s_u[16];
  const uint32_t nv = a.views;
  for (uint32_t v = 0; v < nv; ++v) {
    if (a.clear_a) clear_words(a.clear_a + v * a.clear_a_stride, a.clear_a_words);
    if (a.clear_b) clear_words(a.clear_b + v * a.clear_b_stride, a.clear_b_words);
  }
  const uint32_t tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
  const uint32_t n = a.scene.count;
  const uint32_t num_tiles = (n + TILE - 1) / TILE;
  constexpr int SH_FLOATS = 3 * (L + 1) * (L + 1);
  constexpr int SH_F4 = (SH_FLOATS + 3) / 4;
  const double splat_scale = (double)a.uniforms[0].splat_scale;

  if (tid < 16 * WAVES) s_wcnt[tid / WAVES][tid % WAVES] = 0u;  // zero between rounds from here on

  // Projects listed pairs [0, count) of tile t, views from v0, one round of
  // THREADS pairs at a time.  Lanes take the pairs in list order (a splat's
  // views side by side, so they read its attributes from the same cache
  // lines); the records go through LDS in view-major order and leave as
  // contiguous 16-byte chunks.  Every thread runs every round (barriers).
  auto project_pairs = [&](uint32_t t, uint32_t v0, uint32_t count) {
    for (uint32_t k0 = 0; k0 < count; k0 += THREADS) {
      const uint32_t k = k0 + tid;
      const bool valid = k < count;
      const uint32_t it = valid ? (uint32_t)s_item[k] : 0u;
      const uint32_t h = it & 15u;
      // lanes of this wave holding the same view (four ballots), then the
      // view-major rank through per-(view, wave) counts
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const uint64_t bal = __ballot((h >> b) & 1u);
        peers &= ((h >> b) & 1u) ? bal : ~bal;
      }
      const uint32_t below = lanes_below(peers);
      if (valid && below == 0) s_wcnt[h][w] = (uint32_t)__popcll(peers);
      __syncthreads();
      if (tid < kWave) {  // 16 views x 4 waves, view-major: lane = view * 4 + wave
        const uint32_t x = s_wcnt[tid / WAVES][tid % WAVES];
        uint32_t incl = x;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
          const uint32_t y = __shfl_up(incl, o, kWave);
          if (tid >= (uint32_t)o) incl += y;
        }
        s_wcnt[tid / WAVES][tid % WAVES] = incl - x;
      }
      __syncthreads();
      const uint32_t r = s_wcnt[h][w] + below;
      if (valid) {
        const uint32_t v = v0 + h;
        const uint32_t i = t * TILE + (it >> 4);
        const uint32_t c = s_entry[k];
        const Uniforms& u = s_u[h];

        const float cx = a.scene.centers[3 * (size_t)i + 0];
        const float cy = a.scene.centers[3 * (size_t)i + 1];
        const float cz = a.scene.centers[3 * (size_t)i + 2];
        const float4 rq = a.scene.rotations[i];
        const float4 sa = a.scene.scale_alpha[i];
        float sh[SH_F4 * 4];
        const float4* shp = reinterpret_cast<const float4*>(a.scene.sh + 48 * (size_t)i);
#pragma unroll
        for (int q = 0; q < SH_F4; ++q) {
          const float4 x = shp[q];
          sh[4 * q + 0] = x.x; sh[4 * q + 1] = x.y; sh[4 * q + 2] = x.z; sh[4 * q + 3] = x.w;
        }

        // shaders.wgsl:531-532: colour from the normalised ray direction
        float rd[3] = {cx - u.camera_matrix[12], cy - u.camera_matrix[13], cz - u.camera_matrix[14]};
        const float len = sqrtf((rd[0] * rd[0] + rd[1] * rd[1]) + rd[2] * rd[2]);
        rd[0] = rd[0] / len; rd[1] = rd[1] / len; rd[2] = rd[2] / len;
        float rgb[3];
        sh_color<L>(rd, sh, rgb);
        s_stage[r][0] = make_float4(rgb[0], rgb[1], rgb[2], sa.w);
        s_dst[r] = make_uint2(v, c);
        // the colour's registers are free before the FP64 part starts
        __builtin_amdgcn_sched_barrier(0);

        // shaders.wgsl:533-547
        const double rot[4] = {(double)rq.x, (double)rq.y, (double)rq.z, (double)rq.w};
        const double pos[3] = {(double)cx, (double)cy, (double)cz};
        const double scale[3] = {(double)sa.x * splat_scale, (double)sa.y * splat_scale, (double)sa.z * splat_scale};
        double M[3][3];
        float A00, A10, A01, A11, tx, ty;
        {
#pragma clang fp contract(fast)
          double Tc[3][3];
          contour_transform(scale, rot, Tc);
          contour(Tc, pos, u, M);
          // extractTranslationOfEllipse (shaders.wgsl:179-191)
          const double disc = M[0][0] * M[1][1] - M[0][1] * M[0][1];
          const double inv_disc = rcp64(disc);
          const double t0 = (M[0][1] * M[1][2] - M[1][1] * M[0][2]) * inv_disc;
          const double t1 = (M[0][1] * M[0][2] - M[0][0] * M[1][2]) * inv_disc;
          // extractRotationOfEllipse (shaders.wgsl:193-215)
          const double dxy = M[0][0] - M[1][1];
          const double ra = dxy * dxy;
          const double rb = ra + (4.0 * M[0][1]) * M[0][1];
          // IEEE division and square roots here: ra / rb is exactly 1 for an
          // axis-aligned ellipse, so rc = 1/2 and jj = 0 as in the reference
          const double rc = 0.5 * sqrt(ra / rb);
          double jj = sqrt(0.5 - rc);
          double kk = (-sqrt(0.5 + rc) * sgn(M[0][1])) * sgn(dxy);
          if (M[0][1] < 0.0 || dxy < 0.0) { kk = -kk; jj = -jj; }
          if (dxy < 0.0) { const double tt = jj; jj = -kk; kk = tt; }
          double s0, s1;
          if (a.use_covariance_for_scale) {
            // extractScaleOfCovariance (shaders.wgsl:240-246)
            double Tv[3][3], c00, c11, c01;
            covariance_transform(scale, rot, Tv);
            covariance2(Tv, pos, u, c00, c11, c01);
            const double cd = c00 - c11;
            // IEEE here: sqrt(cd * cd) is exactly |cd|, so a flat splat keeps s1 = 0
            const double bb = sqrt(cd * cd + (4.0 * c01) * c01);
            s0 = sqrt64(((c00 + c11) + bb) * 0.5);
            s1 = sqrt64(((c00 + c11) - bb) * 0.5);
          } else {
            // extractScaleOfEllipse (shaders.wgsl:217-237)
            const double d = ((2.0 * M[0][1]) * jj) * kk;
            const double e = M[2][2] - (((M[0][0] * t0) * t0 + (M[1][1] * t1) * t1) + ((2.0 * M[0][1]) * t0) * t1);
            s0 = sqrt64(fabs(e * rcp64(((M[0][0] * kk) * kk + (M[1][1] * jj) * jj) - d)));
            s1 = sqrt64(fabs(e * rcp64(((M[0][0] * jj) * jj + (M[1][1] * kk) * kk) + d)));
          }
          const double bias = (double)u.ellipse_size_bias;
          const double sa0 = bias + s0, sb0 = bias + s1;
          A00 = (float)(kk * sa0); A10 = (float)(-jj * sa0);
          A01 = (float)(jj * sb0); A11 = (float)(kk * sb0);
          tx = (float)t0; ty = (float)t1;
        }
        uint32_t box_lo, box_hi;
        pixel_box(A00, A10, A01, A11, tx, ty, u, a.use_unaligned_rectangles != 0, box_lo, box_hi);
        s_stage[r][1] = make_float4(A00, A10, A01, A11);
        s_stage[r][2] = make_float4(tx, ty, __uint_as_float(box_lo), __uint_as_float(box_hi));
        if (a.conics && v == 0) {
          double* cn = a.conics + 6 * (size_t)c;
          cn[0] = M[0][0]; cn[1] = M[1][1]; cn[2] = M[2][2];
          cn[3] = M[0][1]; cn[4] = M[0][2]; cn[5] = M[1][2];
        }
      }
      __syncthreads();
      // out in rank order: consecutive lanes write consecutive 16-byte
      // chunks (and boxes) of the same view's records
      const uint32_t n = count - k0 < (uint32_t)THREADS ? count - k0 : (uint32_t)THREADS;
      for (uint32_t q = tid; q < 3 * n; q += THREADS) {
        const uint32_t rr = q / 3, part = q - 3 * rr;
        const uint2 d = s_dst[rr];
        a.records[d.x * a.rec_stride + 3 * (size_t)d.y + part] = s_stage[rr][part];
      }
      if (tid < n) {
        const uint2 d = s_dst[tid];
        const float4 x = s_stage[tid][2];
        a.boxes[d.x * a.box_stride + d.y] = make_uint2(__float_as_uint(x.z), __float_as_uint(x.w));
      }
      if (tid < 16 * WAVES) s_wcnt[tid / WAVES][tid % WAVES] = 0u;
      __syncthreads();
    }
  };

  for (uint32_t t = blockIdx.x; t < num_tiles; t += gridDim.x) {
    // per view: exclusive offsets of the (item, wave) groups inside the tile.
    // The sixteen masks are requested before any is used (past the end a
    // lane reads splat 0's and drops it), so the loads do not wait in turn.
    {
      uint32_t vm[ITEMS];
#pragma unroll
      for (int j = 0; j < ITEMS; ++j) {
        const uint32_t idx = t * TILE + j * THREADS + tid;
        vm[j] = a.vis[idx < n ? idx : 0u];
      }
      // (keeps the scheduler from interleaving each load with its use)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int j = 0; j < ITEMS; ++j)
        if (t * TILE + j * THREADS + tid >= n) vm[j] = 0u;
      for (uint32_t v = 0; v < nv; ++v) {
#pragma unroll
        for (int j = 0; j < ITEMS; ++j) {
          const uint64_t b = __ballot((vm[j] >> v) & 1u);
          if (lane == 0) s_off[v][j * WAVES + w] = (uint32_t)__popcll(b);
        }
      }
    }
    __syncthreads();
    for (uint32_t v = w; v < nv; v += WAVES) {
      const uint32_t c = s_off[v][lane];
      uint32_t incl = c;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, kWave);
        if (lane >= (uint32_t)o) incl += y;
      }
      s_off[v][lane] = incl - c;
      if (lane == 0) s_tp[v] = a.tile_prefix[v * a.tp_stride + t];
    }
    __syncthreads();

    for (uint32_t v0 = 0; v0 < nv; v0 += 16) {
      const uint32_t nh = min(16u, nv - v0);
      {
        constexpr uint32_t UW = sizeof(Uniforms) / 4;
        static_assert(sizeof(Uniforms) % 4 == 0, "uniforms copy as words");
        const uint32_t* src = reinterpret_cast<const uint32_t*>(a.uniforms + v0);
        uint32_t* dst = reinterpret_cast<uint32_t*>(s_u);
        for (uint32_t i = tid; i < nh * UW; i += THREADS) dst[i] = src[i];
      }
      __syncthreads();
      uint32_t listed = 0;  // pairs in the list (uniform across the block)
      // each group's masks are requested one group ahead
      uint32_t ahead = a.vis[t * TILE + tid < n ? t * TILE + tid : 0u];
      for (int j = 0; j < ITEMS; ++j) {
        const uint32_t base = j * THREADS;
        const uint32_t all = t * TILE + base + tid < n ? ahead : 0u;
        if (j + 1 < ITEMS) {
          const uint32_t nx = t * TILE + base + THREADS + tid;
          ahead = a.vis[nx < n ? nx : 0u];
        }
        const uint32_t mine = (all >> v0) & ((1u << nh) - 1u);
        uint32_t total;
        uint32_t slot = block_exclusive_scan<THREADS>((uint32_t)__popc(mine), s_tmp, total);
        if (listed + total > (uint32_t)CAP) {  // project what is listed, then start over
          project_pairs(t, v0, listed);
          __syncthreads();
          listed = 0;
        }
        slot += listed;
        for (uint32_t h = 0; h < nh; ++h) {
          const bool vis = (mine >> h) & 1u;
          const uint64_t b = __ballot(vis);
          if (vis) {
            s_item[slot] = (uint16_t)(((base + tid) << 4) | h);
            s_entry[slot] = s_tp[v0 + h] + s_off[v0 + h][j * WAVES + w] + lanes_below(b);
            ++slot;
          }
        }
        listed += total;
        __syncthreads();
      }
      project_pairs(t, v0, listed);
      __syncthreads();
    }
  }
}

void launch_project(const ProjectLaunch& a, hipStream_t s) {
  switch (a.sh_order) {
    case 0: hipLaunchKernelGGL((k_project<0>), dim3(a.grid), dim3(256), 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_project<1>), dim3(a.grid), dim3(256), 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_project<2>), dim3(a.grid), dim3(256), 0, s, a); break;
    default: hipLaunchKernelGGL((k_project<3>), dim3(a.grid), dim3(256), 0, s, a); break;
  }
}

}  // namespace spl
