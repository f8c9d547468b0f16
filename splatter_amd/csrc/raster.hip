// raster.hip -- compute rasteriser for the projected rectangles.
//
// Replaces the fixed-function half of the reference's render pass
// (renderer.rs:247-291,435-455) and its fragment stage (shaders.wgsl:580-592):
//  bin     one (screen tile, record) pair per 16x16 tile a splat's pixel box
//          touches, written in draw order (single-pass scan with decoupled
//          look-back, pairs emitted by a balanced block-wide expansion), plus
//          the digit histograms of the tile ids for the stable pair sort.
//  ranges  [start, end) of each tile's run in the tile-sorted pair list.
//  raster  one 256-lane block per tile, one lane per pixel, one wave per 8x8
//          quadrant (one view: the heaviest tiles on four blocks, one per
//          quadrant, two rows a wave): records staged in LDS 128 or 256 at a
//          time, each wave keeping the ones that can reach its
//          still-transparent pixels, blended front to back in draw order with
//          the reference's blend state (colour = src*DstAlpha + dst, alpha =
//          dst*(1-srcA), cleared to (0,0,0,1)), the unorm8 target emulated by
//          clamping the source and re-quantising after every blend; a wave
//          stops once its pixels' transmittance is 0, a tile once all four
//          have.
#include "spl_device.h"
#include "spl_internal.h"

namespace spl {

namespace {

// The fragment's Gaussian e^(-|tc|^2 / 2) (shaders.wgsl:586) is evaluated as
// 2^-(|S tc|^2) with S = sqrt(log2(e) / 2) folded into each entry's inverse
// map, so a pixel costs two fused products per coordinate and one exp2.
constexpr float kGaussScale = 0.84932180f;

// 2^y for y <= 0: y = k + f with k = rint(y), |f| <= 1/2, a degree-5
// relative minimax polynomial of 2^f (constant term 1, so exact at
// integers; within 3 ulp of exp2) in fused Horner form, then ldexp by k
// (exact while the result stays normal).  Only correctly rounded IEEE
// operations (fma, rint, ldexp), so it rounds the same on the CPU and the
// GPU.  The oracle clamps y at -125 first; below -8 every alpha is under the
// 1/255 fragment threshold either way, so the pixels agree without it.
__device__ __forceinline__ float spl_exp2(float y) {
  const float k = rintf(y);
  const float f = y - k;
  float p = 1.32764713e-3f;
  p = fmaf(p, f, 9.67554189e-3f);
  p = fmaf(p, f, 5.55071309e-2f);
  p = fmaf(p, f, 2.40221202e-1f);
  p = fmaf(p, f, 6.93146944e-1f);
  p = fmaf(p, f, 1.0f);
  return ldexpf(p, (int)k);
}

__device__ __forceinline__ float q8(float x) {
  x = fminf(fmaxf(x, 0.0f), 1.0f);
  return floorf(x * 255.0f + 0.5f);
}

__device__ __forceinline__ float clamp01(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

// Pixel centres in the view plane are q = (sx px + ox, sy py + oy)
// (shaders.wgsl:561 inverted).  Each staged entry folds this map into its
// own: with e = o - t the coverage coordinate is affine in the integer
// pixel coordinates,
//   unaligned:     u = au px + bu py + cu,  w = av px + bv py + cv
//   axis-aligned:  dx = sx px + ex,  dy = sy py + ey  (then u, w from dx, dy)
// so a pixel costs two fused products per coordinate.
struct PixelMap {
  float sx, ox, sy, oy;
};

__device__ __forceinline__ PixelMap pixel_map(const Uniforms& u) {
  const float Wf = (float)u.image_size[0], Hf = (float)u.image_size[1], vw = u.view_size[0], vh = u.view_size[1];
  PixelMap p;
  p.sx = (2.0f * vw) / Wf;
  p.ox = vw / Wf - vw;
  p.sy = -(2.0f * vh) / Hf;
  p.oy = vh - vh / Hf;
  return p;
}

// The staged form of a record (g0: colour and opacity, g1: the rectangle's
// axes, g2: its centre and pixel box): e0 the inverse map scaled by S (and,
// unaligned, folded with the pixel map), e1 its translation, the coverage
// half-extent and the opacity, e2 the colour.
template <bool UNALIGNED>
__device__ __forceinline__ void prepare_entry(const float4& g0, const float4& g1, const float4& g2, const PixelMap& pm,
                                              float margin, float4& e0, float4& e1, float4& e2) {
  const float A00 = g1.x, A10 = g1.y, A01 = g1.z, A11 = g1.w;
  const float det = A00 * A11 - A10 * A01;
  const float rs = (1.0f / det) * kGaussScale;  // inverse map scaled by S
  const float i00 = A11 * rs, i01 = -A01 * rs, i10 = -A10 * rs, i11 = A00 * rs;
  const float ex = pm.ox - g2.x, ey = pm.oy - g2.y;
  // unaligned: the rectangle |tc| < margin in scaled coordinates;
  // axis-aligned: a square of half-side radius * margin in the view plane
  const float half = UNALIGNED ? margin * kGaussScale
                               : sqrtf(fmaxf(A00 * A00 + A10 * A10, A01 * A01 + A11 * A11)) * margin;
  if constexpr (UNALIGNED) {
    e0 = make_float4(i00 * pm.sx, i01 * pm.sy, i10 * pm.sx, i11 * pm.sy);
    e1 = make_float4(fmaf(i00, ex, i01 * ey), fmaf(i10, ex, i11 * ey), half, g0.w);
  } else {
    e0 = make_float4(i00, i01, i10, i11);
    e1 = make_float4(ex, ey, half, g0.w);
  }
  e2 = make_float4(g0.x, g0.y, g0.z, 0.0f);
}

// Can entry (m, t) cover a pixel centre of the pixel box [x0, x1] x
// [y0, y1]?  The tested coordinates are affine in the pixel, so their exact
// range over the box sits at the corners; the slack (relative to the
// magnitudes the per-pixel evaluation adds up) is far above its rounding.
template <bool UNALIGNED>
__device__ __forceinline__ bool range_meets(const float4& m, const float4& t, float x0, float x1, float y0, float y1,
                                            const PixelMap& pm) {
  const float h = t.z;
  float a0, a1, b0, b1, sa, sb;  // ranges of the tested coordinates, and slack
  if constexpr (UNALIGNED) {
    a0 = (fminf(m.x * x0, m.x * x1) + fminf(m.y * y0, m.y * y1)) + t.x;
    a1 = (fmaxf(m.x * x0, m.x * x1) + fmaxf(m.y * y0, m.y * y1)) + t.x;
    b0 = (fminf(m.z * x0, m.z * x1) + fminf(m.w * y0, m.w * y1)) + t.y;
    b1 = (fmaxf(m.z * x0, m.z * x1) + fmaxf(m.w * y0, m.w * y1)) + t.y;
    sa = 1e-4f * (((fabsf(m.x) * x1 + fabsf(m.y) * y1) + fabsf(t.x)) + h);
    sb = 1e-4f * (((fabsf(m.z) * x1 + fabsf(m.w) * y1) + fabsf(t.y)) + h);
  } else {
    a0 = fmaf(pm.sx, x0, t.x); a1 = fmaf(pm.sx, x1, t.x);  // sx > 0
    b0 = fmaf(pm.sy, y1, t.y); b1 = fmaf(pm.sy, y0, t.y);  // sy < 0
    sa = 1e-4f * ((pm.sx * x1 + fabsf(t.x)) + h);
    sb = 1e-4f * ((fabsf(pm.sy) * y1 + fabsf(t.y)) + h);
  }
  return a1 + sa >= -h && a0 - sa < h && b1 + sb >= -h && b0 - sb < h;
}

// Coverage of the pixel centre (pxf, pyf) by entry (m, t): the scaled
// texture coordinate (u, w2) and whether it lies in the rectangle.
template <bool UNALIGNED>
__device__ __forceinline__ bool covers(const float4& m, const float4& t, float pxf, float pyf, const PixelMap& pm,
                                       float& u, float& w2) {
  const float h = t.z;
  if constexpr (UNALIGNED) {
    u = fmaf(m.x, pxf, fmaf(m.y, pyf, t.x));  // S * texture coordinate
    w2 = fmaf(m.z, pxf, fmaf(m.w, pyf, t.y));
    return fmaxf(fabsf(u), fabsf(w2)) < h;
  } else {
    const float dx = fmaf(pm.sx, pxf, t.x), dy = fmaf(pm.sy, pyf, t.y);
    u = fmaf(m.x, dx, m.y * dy);
    w2 = fmaf(m.z, dx, m.w * dy);
    return fmaxf(fabsf(dx), fabsf(dy)) < h;
  }
}

// One pixel's blend state: colour and transmittance, unorm8 codes or f32.
template <bool F32>
struct PixelState {
  float c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, tr = F32 ? 1.0f : 255.0f;

  // One fragment: alpha (shaders.wgsl:585-589) and the blend.  in: the
  // coverage test's verdict, for walks over candidates that did not pass
  // it yet; an uncovered fragment blends nothing.
  __device__ __forceinline__ void shade(float u, float w2, float opacity, const float4& col, bool in) {
    float alpha = opacity * spl_exp2(-fmaf(u, u, w2 * w2));
    if constexpr (!F32) {
      // unorm8 codes C, T: rint(s * T + C) (s clamped, sum capped at
      // 255) and rint(T * (1 - a)): the ideal 8-bit blender.  Below the
      // threshold alpha becomes 0, which leaves C and T as they are
      // (they hold integers; clamp01 of a NaN product is 0), so no
      // branch is needed.
      alpha = in && alpha >= 1.0f / 255.0f ? alpha : 0.0f;
      const float sr = col.x * alpha, sg = col.y * alpha, sb = col.z * alpha;
      // the source alpha needs no clamp: the opacity is a sigmoid (or 1
      // in the dev-test scene) and spl_exp2 of a non-positive argument is
      // at most 1 (p(f) <= p(0) = 1 when rint(y) = 0, p(f) / 2 < 1 below),
      // so alpha is in [0, 1] after the threshold
      const float sa = alpha;
      c0 = fminf(rintf(fmaf(clamp01(sr), tr, c0)), 255.0f);
      c1 = fminf(rintf(fmaf(clamp01(sg), tr, c1)), 255.0f);
      c2 = fminf(rintf(fmaf(clamp01(sb), tr, c2)), 255.0f);
      tr = rintf(fmaf(-sa, tr, tr));
    } else if (in && alpha >= 1.0f / 255.0f) {
      const float sr = col.x * alpha, sg = col.y * alpha, sb = col.z * alpha;
      const float da = tr;
      c0 = fmaf(sr, da, c0);
      c1 = fmaf(sg, da, c1);
      c2 = fmaf(sb, da, c2);
      tr = fmaf(-alpha, da, da);
    }
  }

  // The pixel's final value, as the frame stores it.
  __device__ __forceinline__ uint32_t packed(bool bgra) const {
    uint32_t r, g, b, al;
    if (!F32) {
      r = (uint32_t)c0; g = (uint32_t)c1; b = (uint32_t)c2; al = (uint32_t)tr;
    } else {
      r = (uint32_t)q8(c0); g = (uint32_t)q8(c1); b = (uint32_t)q8(c2); al = (uint32_t)q8(tr);
    }
    return bgra ? (b | (g << 8) | (r << 16) | (al << 24)) : (r | (g << 8) | (b << 16) | (al << 24));
  }
};

// Per-lane walk over one run of cnt <= 32 entries (cm, ct, cc: their three
// staged rows, in list order, in LDS) of a wave's 8x8 region with top-left
// pixel (fqx0, fqy0); rows: 64 words of LDS for the coverage rows.  The wave
// first finds, for every pixel, the run's entries that can cover it, then
// every lane walks only those, in list order, each lane on its own: the
// wave takes as many steps as its busiest pixel has covering entries, not
// as many as the run has.
//
// The candidates come per row, entry-parallel: 2, 4 or 8 lanes per entry
// each take 4, 2 or 1 of the region's rows and solve the rectangle's two
// slabs |a px + b py + c| < h for the row's column interval, widened by far
// more than the rounding of either that solution or the per-pixel test
// (2^-19 of the magnitudes, against about 12 roundings of 2^-24), so it
// holds every column whose centre the per-pixel test accepts.  A slab that
// hardly varies along x (|a| < h 1e-6), or with a coefficient that is not
// finite, is taken as all columns.  Each pixel lane then gathers its row's
// 32 bits from LDS (row y: byte 3 - (k >> 3) of word k & 7 holds slot k's
// row mask, bit x = column x), and the walk applies the exact per-pixel
// test of the other modes to every candidate (a candidate that fails it
// blends with alpha 0, which leaves the pixel as it is), so the pixels are
// the same.  Runs of up to 12 entries are tested per pixel instead.
template <bool UNALIGNED, bool F32>
__device__ __forceinline__ void walk_run(const float4* cm, const float4* ct, const float4* cc, uint32_t* rows,
                                         uint32_t cnt, uint32_t lane, bool inside, float pxf, float pyf, float fqx0,
                                         float fqy0, float span, const PixelMap& pm, float margin,
                                         PixelState<F32>& pix) {
  uint8_t* const rows8 = reinterpret_cast<uint8_t*>(rows);
  constexpr uint32_t kRowsMin = 12;  // runs up to this long are tested per pixel
  // column interval of row y (local, 0..7) as [ceil(lo), floor(hi1) - 1],
  // clamped to the region, as a row mask
  auto row_mask = [&](float lo, float hi1) -> uint32_t {
    const uint32_t first = (uint32_t)__builtin_amdgcn_fmed3f(ceilf(lo), 0.0f, 8.0f);
    const uint32_t end = (uint32_t)__builtin_amdgcn_fmed3f(floorf(hi1), 0.0f, 8.0f);
    return ((1u << end) - 1u) & ~((1u << first) - 1u);
  };
  // slab |a px + b py + c| < h in local columns: lo(y) = P y + lo0,
  // hi(y) + 1 = P y + hi0
  auto slab = [&](float a_, float b_, float c_, float h_, float& P, float& lo0, float& hi0) {
    const float ia = __builtin_amdgcn_rcpf(a_);
    const float R = h_ * fabsf(ia);
    const float Pv = -b_ * ia;
    const float Qg = -c_ * ia;
    const float Q = fmaf(Pv, fqy0, Qg) - fqx0;
    const float eps = 0x1p-19f * (((fabsf(Pv) * span + fabsf(Qg)) + R) + span) + 0x1p-19f;
    // (R = 0 for an infinite a, whose products with a zero
    // coordinate are NaN in the per-pixel test)
    const bool ok = R > 0.0f && R < 1e6f && fabsf(Pv) < 1e6f && fabsf(Qg) < 1e9f;
    P = ok ? Pv : 0.0f;
    lo0 = ok ? (Q - R) - eps : -1e9f;
    hi0 = ok ? ((Q + R) + eps) + 1.0f : 1e9f;
  };
  uint32_t mask = 0;
  if (cnt <= kRowsMin) {
    // a short run: every lane tests its pixel against each slot
    // (exact), slot k at bit K - 1 - k for K = cnt rounded up to 4
    for (uint32_t k = 0; k < cnt; k += 4) {
      uint32_t nib = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float u, w2;
        float4 t4;
        if constexpr (UNALIGNED) {
          // the half-extent is margin * S for every entry
          const float2 t2 = *reinterpret_cast<const float2*>(&ct[k + q]);
          t4 = make_float4(t2.x, t2.y, margin * kGaussScale, 0.0f);
        } else {
          t4 = ct[k + q];
        }
        nib = (nib << 1) | (covers<UNALIGNED>(cm[k + q], t4, pxf, pyf, pm, u, w2) ? 1u : 0u);
      }
      mask = (mask << 4) | nib;
    }
    mask <<= 32u - ((cnt + 3u) & ~3u);
  } else {
    {
      const uint32_t lpe = cnt <= 8u ? 3u : (cnt <= 16u ? 2u : 1u);  // log2 lanes per entry
      const uint32_t k = lane >> lpe;
      if (k < cnt) {
        const uint32_t rpl = 8u >> lpe;  // rows per lane
        const uint32_t y0 = (lane & ((1u << lpe) - 1u)) * rpl;
        const float4 m = cm[k];
        const float4 t = ct[k];
        uint8_t* const dst = rows8 + (k & 7u) * 4u + (3u - (k >> 3));
        if constexpr (UNALIGNED) {
          const float h = margin * kGaussScale;
          float P1, l1, e1, P2, l2, e2;
          slab(m.x, m.y, t.x, h, P1, l1, e1);
          slab(m.z, m.w, t.y, h, P2, l2, e2);
          for (uint32_t i = 0; i < rpl; ++i) {
            const float y = (float)(y0 + i);
            const float lo = fmaxf(fmaf(P1, y, l1), fmaf(P2, y, l2));
            const float hi1 = fminf(fmaf(P1, y, e1), fmaf(P2, y, e2));
            dst[(y0 + i) * 32u] = (uint8_t)row_mask(lo, hi1);
          }
        } else {
          // dx = sx px + t.x bounds the columns, dy = sy py + t.y the rows
          float P, lx, ex, Py, ly, ey;
          slab(pm.sx, 0.0f, t.x, t.z, P, lx, ex);
          slab(pm.sy, 0.0f, t.y, t.z, Py, ly, ey);
          // rows: the same interval, in rows (py in place of px)
          ly += fqx0 - fqy0;
          ey += fqx0 - fqy0;
          const uint32_t cols = row_mask(lx, ex);
          const uint32_t rmask = row_mask(ly, ey);
          for (uint32_t i = 0; i < rpl; ++i)
            dst[(y0 + i) * 32u] = (uint8_t)(((rmask >> (y0 + i)) & 1u) ? cols : 0u);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    // this pixel's candidates: slot k at bit 31 - k
    const uint32_t xl = lane & 7u;
    const uint4 ra = *reinterpret_cast<const uint4*>(&rows[(lane >> 3) * 8u]);
    const uint4 rb = *reinterpret_cast<const uint4*>(&rows[(lane >> 3) * 8u + 4u]);
    mask = (mask << 1) | ((ra.x >> xl) & 0x01010101u);
    mask = (mask << 1) | ((ra.y >> xl) & 0x01010101u);
    mask = (mask << 1) | ((ra.z >> xl) & 0x01010101u);
    mask = (mask << 1) | ((ra.w >> xl) & 0x01010101u);
    mask = (mask << 1) | ((rb.x >> xl) & 0x01010101u);
    mask = (mask << 1) | ((rb.y >> xl) & 0x01010101u);
    mask = (mask << 1) | ((rb.z >> xl) & 0x01010101u);
    mask = (mask << 1) | ((rb.w >> xl) & 0x01010101u);
  }
  mask &= ~0u << (32u - cnt);  // slots at or past cnt hold stale rows
  if (pix.tr == 0.0f || !inside) mask = 0u;
  while (mask) {
    const uint32_t k = (uint32_t)__builtin_clz(mask);
    mask &= ~(0x80000000u >> k);
    const float4 t = ct[k];
    float u, w2;
    const bool in = covers<UNALIGNED>(cm[k], t, pxf, pyf, pm, u, w2);
    pix.shade(u, w2, t.w, cc[k], in);
    if (pix.tr == 0.0f) mask = 0u;
  }
}

}  // namespace

template <int THREADS, int ITEMS>
__global__ __launch_bounds__(THREADS) void k_bin(BinLaunch a) {
  constexpr int TILE = THREADS * ITEMS;
  constexpr int WAVES = THREADS / kWave;
  static_assert(ITEMS == 4, "a thread's items are one 16-byte LDS row");
  // Per entry of the tile: first output slot, first tile (x | y << 16),
  // width in tiles, record.  Thread t's four entries are one 16-byte row of
  // each array, written with one store (consecutive threads, consecutive
  // banks; four separate words a thread would put 8 lanes on each bank).
  __shared__ __align__(16) uint32_t s_off[TILE + 4];
  __shared__ __align__(16) uint32_t s_org[TILE];
  __shared__ __align__(16) uint32_t s_w[TILE];
  __shared__ __align__(16) uint32_t s_rec[TILE];
  // the entry that owns each output slot of a tile with at most OWN pairs
  constexpr uint32_t OWN = 4096, PER = OWN / THREADS;
  static_assert(PER == 16, "a thread's owner slots are two 16-byte LDS rows");
  __shared__ __align__(16) uint16_t s_own[OWN];
  __shared__ uint32_t s_hist[kMaxPairPasses * 256];
  __shared__ uint32_t s_tmp[WAVES];
  __shared__ uint32_t s_tile;
  __shared__ uint64_t s_prefix;

  const uint32_t tid = threadIdx.x;
  const uint32_t v = blockIdx.y;
  uint32_t* counters = a.counters + (size_t)v * kCounterWords;
  const uint32_t* order = a.order ? a.order + v * a.order_stride : nullptr;
  const uint2* boxes = a.boxes + v * a.box_stride;
  uint32_t* pair_keys = a.pair_keys + v * a.pair_stride;
  uint32_t* pair_vals = a.pair_vals + v * a.pair_stride;
  uint64_t* status = a.status + v * a.status_stride;
  if (a.clear_a) clear_words(a.clear_a + v * a.clear_a_stride, a.clear_a_words);
  if (a.clear_b) clear_words(a.clear_b + v * a.clear_b_stride, a.clear_b_words);
  for (int i = tid; i < a.pair_passes * 256; i += THREADS) s_hist[i] = 0u;
  const uint32_t n = counters[kVisible];
  const uint32_t num_tiles = (n + TILE - 1) / TILE;
  if (num_tiles == 0 && blockIdx.x == 0 && tid == 0) {
    counters[kPairs] = 0u;
    counters[kPairOverflow] = 0u;
  }
  const uint32_t tiles_x = a.uniforms[v].tiles_x;

  while (true) {
    if (tid == 0) s_tile = atomicAdd(&counters[kCtrBin], 1u);
    __syncthreads();
    const uint32_t t = s_tile;
    if (t >= num_tiles) break;
    // The items' record indices, then their boxes, are all requested before
    // any is used (past the end a lane reads item 0 and drops it), so that
    // no load waits for the one before.
    uint32_t cnt[ITEMS], sum = 0, rec[ITEMS];
    uint2 box[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const uint32_t r = t * TILE + tid * ITEMS + k;
      rec[k] = order ? order[r < n ? r : 0u] : r;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const uint32_t r = t * TILE + tid * ITEMS + k;
      if (r < n && rec[k] >= n) counters[kScatterError] = 1u;  // only after an upstream error
      if (r >= n || rec[k] >= n) rec[k] = 0u;  // never read past the boxes
      box[k] = boxes[rec[k]];
    }
    uint32_t org[ITEMS], wid[ITEMS], off[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      const uint32_t r = t * TILE + tid * ITEMS + k;
      cnt[k] = 0;
      org[k] = 0;
      wid[k] = 0;
      if (r < n) {
        const uint2 q = box[k];
        const uint32_t lo = q.x, hi = q.y;
        const uint32_t x0 = (lo & 0xFFFFu) / kTileSize, y0 = (lo >> 16) / kTileSize;
        const uint32_t x1 = (hi & 0xFFFFu) / kTileSize, y1 = (hi >> 16) / kTileSize;
        if (lo != 0xFFFFFFFFu && x1 >= x0 && y1 >= y0) {
          cnt[k] = (x1 - x0 + 1) * (y1 - y0 + 1);
          org[k] = x0 | (y0 << 16);
          wid[k] = x1 - x0 + 1;
        }
      }
      sum += cnt[k];
    }
    // (an entry without pairs owns no slot, so its row is never read)
    reinterpret_cast<uint4*>(s_org)[tid] = make_uint4(org[0], org[1], org[2], org[3]);
    reinterpret_cast<uint4*>(s_w)[tid] = make_uint4(wid[0], wid[1], wid[2], wid[3]);
    reinterpret_cast<uint4*>(s_rec)[tid] = make_uint4(rec[0], rec[1], rec[2], rec[3]);
    uint32_t total;
    uint32_t excl = block_exclusive_scan<THREADS>(sum, s_tmp, total);
#pragma unroll
    for (int k = 0; k < ITEMS; ++k) {
      off[k] = excl;
      excl += cnt[k];
    }
    reinterpret_cast<uint4*>(s_off)[tid] = make_uint4(off[0], off[1], off[2], off[3]);
    if (tid < kWave) {
      const uint64_t prefix = wave_lookback64(status, t, (uint64_t)total);
      if (tid == 0) {
        s_off[TILE] = total;
        s_prefix = prefix;
      }
      if (tid == 0 && t == num_tiles - 1) {
        const uint64_t k_all = prefix + total;
        counters[kPairs] = (uint32_t)(k_all < a.pair_cap ? k_all : a.pair_cap);
        counters[kPairOverflow] = k_all > a.pair_cap ? 1u : 0u;
      }
    }
    __syncthreads();
    const uint64_t prefix = s_prefix;
    // Owner of each output slot: every entry with pairs writes its index at
    // its first slot, then an inclusive max-scan over the slots carries it
    // forward, so the expansion reads a slot's owner instead of
    // binary-searching the offsets (ten dependent LDS reads).  A tile with
    // more than OWN pairs searches.
    // (Slots past the total hold stale owners: the scan carries them only
    // forward, into slots that no pair reads.)
    const bool table = total <= OWN;
    if (table) {
      const uint32_t lane = tid & (kWave - 1), wv = tid / kWave;
      uint4* const own4 = reinterpret_cast<uint4*>(s_own);  // eight slots a row
      for (uint32_t q = tid; q < (total + 7u) / 8u; q += THREADS) own4[q] = make_uint4(0u, 0u, 0u, 0u);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < ITEMS; ++k)
        if (cnt[k] > 0u) s_own[off[k]] = (uint16_t)(tid * ITEMS + k);
      __syncthreads();
      // this thread's sixteen slots, two rows; slot 2i + h is half h of word i
      const uint4 ra = own4[2 * tid], rb = own4[2 * tid + 1];
      uint32_t wd[8] = {ra.x, ra.y, ra.z, ra.w, rb.x, rb.y, rb.z, rb.w};
      uint32_t run = 0;
#pragma unroll
      for (int i = 0; i < 8; ++i) run = max(run, max(wd[i] & 0xFFFFu, wd[i] >> 16));
      uint32_t incl = run;
#pragma unroll
      for (int o = 1; o < kWave; o <<= 1) {
        const uint32_t y = __shfl_up(incl, o, kWave);
        if (lane >= (uint32_t)o) incl = max(incl, y);
      }
      if (lane == kWave - 1) s_tmp[wv] = incl;
      __syncthreads();
      uint32_t carry = __shfl_up(incl, 1, kWave);
      if (lane == 0) carry = 0;
      for (uint32_t i = 0; i < wv; ++i) carry = max(carry, s_tmp[i]);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t lo = max(carry, wd[i] & 0xFFFFu);
        carry = max(lo, wd[i] >> 16);
        wd[i] = lo | (carry << 16);
      }
      own4[2 * tid] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
      own4[2 * tid + 1] = make_uint4(wd[4], wd[5], wd[6], wd[7]);
      __syncthreads();
    }
    for (uint32_t s = tid; s < total; s += THREADS) {
      uint32_t lo = 0, hi = TILE - 1;
      if (table) {
        lo = s_own[s];
      } else {
        while (lo < hi) {
          const uint32_t mid = (lo + hi + 1) >> 1;
          if (s_off[mid] <= s) lo = mid; else hi = mid - 1;
        }
      }
      const uint64_t pos = prefix + s;
      if (pos < a.pair_cap) {
        const uint32_t k = s - s_off[lo];
        const uint32_t w = s_w[lo], org = s_org[lo];
        const uint32_t tile = ((org >> 16) + k / w) * tiles_x + (org & 0xFFFFu) + k % w;
        pair_keys[pos] = tile;
        pair_vals[pos] = s_rec[lo];
        for (int q = 0; q < a.pair_passes; ++q) atomicAdd(&s_hist[q * 256 + ((tile >> (8 * q)) & 255u)], 1u);
      }
    }
    __syncthreads();
  }
  for (int i = tid; i < a.pair_passes * 256; i += THREADS)
    if (s_hist[i]) atomicAdd(&a.hist[v * a.hist_stride + i], s_hist[i]);
  finish_persistent(&counters[kCtrBin], &counters[kDone + kCtrBin]);
}

// Run boundaries of the tile-sorted pair keys, four keys per thread.
__global__ __launch_bounds__(256) void k_tile_ranges(RangesLaunch a) {
  const uint32_t v = blockIdx.y;
  if (a.clear_a) clear_words(a.clear_a + v * a.clear_a_stride, a.clear_a_words);
  if (a.clear_b) clear_words(a.clear_b + v * a.clear_b_stride, a.clear_b_words);
  const uint32_t* keys = a.pair_keys + v * a.pair_stride;  // 16-byte aligned: pair_stride % 4 == 0
  uint2* ranges = a.ranges + v * a.ranges_stride;
  const uint32_t tiles = a.uniforms[v].tiles_x * a.uniforms[v].tiles_y;
  uint32_t n = a.counters[(size_t)v * kCounterWords + kPairs];
  if (n > a.pair_cap) n = a.pair_cap;
  const uint32_t step = gridDim.x * blockDim.x * 4;
  for (uint32_t i0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4; i0 < n; i0 += step) {
    // the three loads go out together: out of range they read key 0 (or
    // the run's first key) and the value is replaced after
    uint32_t k[6];
    const uint32_t prev = keys[i0 > 0 ? i0 - 1 : 0u];
    const uint32_t next = keys[i0 + 4 < n ? i0 + 4 : 0u];
    if (i0 + 4 <= n) {
      const uint4 q = *reinterpret_cast<const uint4*>(keys + i0);
      k[1] = q.x; k[2] = q.y; k[3] = q.z; k[4] = q.w;
    } else {
#pragma unroll
      for (int j = 0; j < 4; ++j) k[1 + j] = i0 + j < n ? keys[i0 + j] : 0xFFFFFFFFu;
    }
    k[0] = i0 > 0 ? prev : 0xFFFFFFFFu;
    k[5] = i0 + 4 < n ? next : 0xFFFFFFFFu;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t i = i0 + j, t = k[1 + j];
      if (i >= n || t >= tiles) continue;
      if (k[j] != t) ranges[t].x = i;
      if (k[2 + j] != t) ranges[t].y = i + 1;
    }
  }
}

// Raster order of a view's tiles: the tiles with the longest pair lists
// ("heavy": at least heavy_mult x the mean and heavy_min pairs, at most
// heavy_cap of them, whole log2 classes only) first, longest class first,
// then every other tile in row-major order.  A heavy tile blends thousands
// of entries per pixel, one after another; started last, it would outlast
// the rest of the frame.  The heavy tiles in classes of at least
// split_mult x the mean (and 2048 pairs) are counted apart: the raster may
// give each of them four blocks.
__global__ __launch_bounds__(1024) void k_tile_order(OrderLaunch a) {
  constexpr int THREADS = 1024;
  __shared__ uint32_t s_cnt[32];
  __shared__ uint32_t s_cur[32];
  __shared__ uint32_t s_tmp[THREADS / kWave];
  __shared__ uint32_t s_cut;
  const uint32_t v = blockIdx.y, tid = threadIdx.x;
  const uint32_t T = a.uniforms[v].tiles_x * a.uniforms[v].tiles_y;
  const uint2* ranges = a.ranges + v * a.ranges_stride;
  uint32_t* order = a.order + v * a.order_stride;
  uint32_t* counters = a.counters + (size_t)v * kCounterWords;
  uint32_t k = counters[kPairs];
  if (k > a.pair_cap) k = a.pair_cap;
  const uint32_t mean = T ? k / T : 0u;
  const uint32_t thr = max(a.heavy_mult * mean, a.heavy_min);
  if (tid < 32) s_cnt[tid] = 0u;
  __syncthreads();
  const uint32_t chunk = (T + THREADS - 1) / THREADS;
  const uint32_t t0 = min(tid * chunk, T), t1 = min(t0 + chunk, T);
  // log2 classes of a heavy candidate (-1 otherwise) for the eight tiles
  // from tb on, past t1 -2: their ranges are requested together (past the
  // end a lane reads tile 0's and drops it) instead of one after another
  constexpr uint32_t G = 8;
  auto classes = [&](uint32_t tb, int (&c)[G]) {
    uint2 r[G];
#pragma unroll
    for (uint32_t g = 0; g < G; ++g) r[g] = ranges[tb + g < t1 ? tb + g : 0u];
#pragma unroll
    for (uint32_t g = 0; g < G; ++g) {
      const uint32_t len = r[g].y - r[g].x;
      c[g] = tb + g >= t1 ? -2 : (len >= thr ? 31 - __clz(len) : -1);
    }
  };
  for (uint32_t tb = t0; tb < t1; tb += G) {
    int c[G];
    classes(tb, c);
#pragma unroll
    for (uint32_t g = 0; g < G; ++g)
      if (c[g] >= 0) atomicAdd(&s_cnt[c[g]], 1u);
  }
  __syncthreads();
  if (tid == 0) {
    // whole classes, longest first, while they fit under the cap
    uint32_t h = 0, cut = 32;
    for (int c = 31; c >= 0; --c) {
      if (h + s_cnt[c] > a.heavy_cap) break;
      s_cur[c] = h;
      h += s_cnt[c];
      cut = (uint32_t)c;
    }
    s_cut = cut;
    counters[kHeavyTiles] = h;
    uint32_t split = 0;
    if (a.split_mult) {
      const uint32_t c_split = 31u - __clz(max(a.split_mult * mean, 2048u));
      for (uint32_t c = max(c_split, cut); c < 32u; ++c) split += s_cnt[c];
    }
    counters[kSplitTiles] = split;
  }
  __syncthreads();
  const uint32_t cut = s_cut;
  uint32_t light = 0;
  for (uint32_t tb = t0; tb < t1; tb += G) {
    int c[G];
    classes(tb, c);
#pragma unroll
    for (uint32_t g = 0; g < G; ++g) {
      if (c[g] == -2) continue;
      if (c[g] >= (int)cut) order[atomicAdd(&s_cur[c[g]], 1u)] = tb + g;
      else ++light;
    }
  }
  uint32_t total;
  uint32_t pos = block_exclusive_scan<THREADS>(light, s_tmp, total);
  const uint32_t heavy = T - total;
  for (uint32_t tb = t0; tb < t1; tb += G) {
    int c[G];
    classes(tb, c);
#pragma unroll
    for (uint32_t g = 0; g < G; ++g)
      if (c[g] != -2 && c[g] < (int)cut) order[heavy + pos++] = tb + g;
  }
}

void launch_tile_order(const OrderLaunch& a, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_order, dim3(1, a.views), dim3(1024), 0, s, a);
}

// One 256-lane block per 16x16 tile; wave w owns the 8x8 quadrant
// (w & 1, w >> 1), one lane per pixel.  Entries arrive 256 at a time in draw
// order: every lane gathers one record and stores the prepared entry (inverse
// affine map, translation, coverage half-extent, colour, pixel box) in LDS,
// while the next batch is already being fetched into registers.  Each wave
// then keeps, in order, only the entries whose pixel box meets its quadrant
// (ballot + mbcnt) and blends those; a wave whose 64 pixels are all opaque
// stops blending, and the block stops when all four have.
template <bool UNALIGNED, bool F32, int MODE, int B>
// (mode 3 occupancy: 128-entry batches at seven waves per SIMD, 72 VGPRs,
// one to three registers spilled outside the loops; the f32 blend needs
// more and keeps six; 256-entry batches fit six at 84)
__global__ __launch_bounds__(256, MODE == 3 ? (B == 128 ? (F32 ? 6 : 7) : 5) : 1) void k_raster(RasterLaunch a) {
  static_assert(B == 128 || B == 256, "entries staged per batch");
  constexpr int U = B / 64;  // list slots per step of the entry-major walks
  constexpr bool WALK = MODE == 1, SUBS = MODE == 2, LANES = MODE == 3;
  constexpr bool LANES_ = LANES;
  // per staged entry: [0] i00 i01 i10 i11, [1] tx ty half alpha, [2] r g b -
  // (one array, so an entry's three rows share one LDS address)
  __shared__ float4 s_ent[2][3][B];
  __shared__ uint2 s_box[2][B];    // pixel box lo / hi (x | y << 16)
  __shared__ __align__(16) uint8_t s_list[4][B];  // per-wave compacted entry slots, read 4 at a time
  // mode 3: per wave, the coverage of 32 run slots as 8 rows x 8 words; byte
  // 3 - (k >> 3) of word k & 7 of row y holds slot k's row mask (bit x =
  // column x of the quadrant)
  __shared__ __align__(16) uint32_t s_rows[LANES_ ? 4 : 1][LANES_ ? 64 : 4];
  __shared__ __align__(16) uint8_t s_sub[SUBS ? 4 : 1][SUBS ? 4 : 1][SUBS ? B : 4];  // mode 2: per 4x4 sub-block

  // Block -> (view, tile).  First the heavy tiles of every view (their
  // count rounded up to 8 blocks; with split_heavy, mode 3 only, four blocks
  // per heavy tile, one per 8x8 quadrant, each wave two of its rows), then
  // each view's other tiles in a slot range of slots_per_view blocks,
  // banded per XCD: blocks are dealt round-robin to the 8 XCDs, and each
  // XCD takes a contiguous run of rows (neighbouring tiles share records in
  // its L2).  A heavy tile's pixels wait on their own long chains of
  // blends; with one view on the device the other blocks are soon done, and
  // sixteen waves on a tile instead of four shorten its chains' wave-wide
  // steps (each step waits for the busiest of 16 pixels, not 64).
  const uint32_t nv = (uint32_t)a.views;
  uint32_t v = 0, pos = 0, quad = 4;  // quad < 4: the quadrant of a split heavy tile
  {
    uint32_t heavy_total = 0, split_total = 0;
    if (a.longest_first)
      for (uint32_t i = 0; i < nv; ++i) heavy_total += a.counters[(size_t)i * kCounterWords + kHeavyTiles];
    // split tiles: the first of view 0's heavy tiles (splitting is for one view)
    if (LANES && a.split_heavy && a.longest_first) split_total = a.counters[kSplitTiles];
    const uint32_t split_blocks = ((split_total + 7u) & ~7u) * 4u;
    const uint32_t heavy_blocks = split_blocks + ((heavy_total - split_total + 7u) & ~7u);
    uint32_t b = blockIdx.x;
    if (b < heavy_blocks) {
      if (b < split_blocks) {
        quad = b & 3u;
        b >>= 2;
        if (b >= split_total) return;
      } else {
        b = split_total + (b - split_blocks);
        if (b >= heavy_total) return;
      }
      for (;; ++v) {
        const uint32_t h = a.counters[(size_t)v * kCounterWords + kHeavyTiles];
        if (b < h) break;
        b -= h;
      }
      pos = b;
    } else {
      b -= heavy_blocks;
      v = b / a.slots_per_view;
      if (v >= nv) return;
      const uint32_t k = b % a.slots_per_view, per = a.slots_per_view / 8;
      const uint32_t p = (k & 7u) * per + (k >> 3);
      const uint32_t h = a.longest_first ? a.counters[(size_t)v * kCounterWords + kHeavyTiles] : 0u;
      const uint32_t tv = a.uniforms[v].tiles_x * a.uniforms[v].tiles_y;
      if (p >= tv - h) return;
      pos = h + p;
    }
  }
  const Uniforms& u = a.uniforms[v];
  const uint2* ranges = a.ranges + v * a.ranges_stride;
  const uint32_t* pair_vals = a.pair_vals + v * a.pair_stride;
  const float4* records = a.records + v * a.rec_stride;
  const uint32_t tile = a.longest_first ? a.order[v * a.order_stride + pos] : pos;
  const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
  // this wave's 8x8 region: quadrant w, or for a split tile rows 2w and
  // 2w + 1 of quadrant quad (lanes 0-15; the others hold no pixel)
  const uint32_t qd = quad < 4u ? quad : w;
  const uint32_t qx0 = (tile % u.tiles_x) * kTileSize + (qd & 1u) * 8u;
  const uint32_t qy0 = (tile / u.tiles_x) * kTileSize + (qd >> 1) * 8u + (quad < 4u ? 2u * w : 0u);
  // mode 2: lane group g = lane / 16 owns the 4x4 sub-block (g & 1, g >> 1)
  const uint32_t g = lane >> 4;
  const uint32_t px = SUBS ? qx0 + (g & 1u) * 4u + (lane & 3u) : qx0 + (lane & 7u);
  const uint32_t py = SUBS ? qy0 + (g >> 1) * 4u + ((lane >> 2) & 3u) : qy0 + (lane >> 3);
  const uint32_t W = u.image_size[0], H = u.image_size[1];
  const bool inside = px < W && py < H && (quad == 4u || lane < 16u);
  // (the ranges are cleared for the next frame by its pair sort)
  const uint2 range = ranges[tile];

  const PixelMap pm = pixel_map(u);
  const float pxf = (float)px, pyf = (float)py;
  const float margin = u.ellipse_margin;
  PixelState<F32> pix;
  const uint32_t visible = a.counters[(size_t)v * kCounterWords + kVisible];

  // register stage of one entry per lane
  float4 g0 = make_float4(0, 0, 0, 0), g1 = g0, g2 = g0;
  auto fetch = [&](uint32_t base) {
    if (tid < (uint32_t)B && base + tid < range.y) {
      uint32_t c = pair_vals[base + tid];
      if (c >= visible) c = 0;  // only after an upstream error; never read past the records
      g0 = records[3 * (size_t)c + 0];
      g1 = records[3 * (size_t)c + 1];
      g2 = records[3 * (size_t)c + 2];
    }
  };
  auto stage = [&](int buf) {
    if (tid >= (uint32_t)B) return;
    prepare_entry<UNALIGNED>(g0, g1, g2, pm, margin, s_ent[buf][0][tid], s_ent[buf][1][tid], s_ent[buf][2][tid]);
    s_box[buf][tid] = make_uint2(__float_as_uint(g2.z), __float_as_uint(g2.w));
  };

  // lanes whose pixel still takes fragments (inside the image, not yet
  // opaque); in modes 0 and 1 lane y * 8 + x holds pixel (x, y) of the quadrant
  uint64_t alive = __ballot(inside);
  bool wave_done = alive == 0ull;
  uint32_t consumed = 0;
  int buf = 0;
  if (range.x < range.y) {
    fetch(range.x);
    stage(0);
  }
  __syncthreads();
  for (uint32_t base = range.x; base < range.y; base += B) {
    const uint32_t cnt = (range.y - base) < (uint32_t)B ? (range.y - base) : (uint32_t)B;
    const bool more = base + B < range.y;
    if (more) fetch(base + B);
    consumed += cnt;
    if (!wave_done) {
      // Entries of this batch that can cover a pixel centre still taking
      // fragments, in order: pixel box against the box of the live pixels,
      // then the exact range of the coverage coordinate over that box's
      // pixel centres (it is affine in the centre, so its extremes sit at the
      // corner centres).  An entry that meets only opaque pixels would add
      // s * 0 to them, so leaving it out changes nothing.  The live box's
      // columns come from the OR of the mask's eight rows, its rows from the
      // lowest and highest set bits.  Mode 2 lays its lanes out per sub-block
      // (bits 16 g .. 16 g + 15 are sub-block g): it keeps the quadrant's box
      // and skips opaque sub-blocks instead.
      uint32_t ax0 = qx0, ax1 = min(qx0 + 7u, W - 1u), ay0 = qy0, ay1 = min(qy0 + 7u, H - 1u);
      if constexpr (!SUBS) {
        uint32_t cols = (uint32_t)alive | (uint32_t)(alive >> 32);
        cols |= cols >> 16;
        cols |= cols >> 8;
        ax0 = qx0 + (uint32_t)__builtin_ctz(cols & 0xFFu);
        ax1 = qx0 + 31u - (uint32_t)__builtin_clz(cols & 0xFFu);
        ay0 = qy0 + ((uint32_t)__builtin_ctzll(alive) >> 3);
        ay1 = qy0 + ((63u - (uint32_t)__builtin_clzll(alive)) >> 3);
      }
      uint32_t n = 0;
      uint32_t ns[4] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int k = 0; k < B / 64; ++k) {
        const uint32_t j = k * 64 + lane;
        bool ov = false;
        if (j < cnt) {
          const uint2 bx = s_box[buf][j];
          const uint32_t x0 = bx.x & 0xFFFFu, y0 = bx.x >> 16, x1 = bx.y & 0xFFFFu, y1 = bx.y >> 16;
          ov = bx.x != 0xFFFFFFFFu && x0 <= ax1 && x1 >= ax0 && y0 <= ay1 && y1 >= ay0;
          if (ov && !SUBS)
            ov = range_meets<UNALIGNED>(s_ent[buf][0][j], s_ent[buf][1][j], (float)ax0, (float)ax1, (float)ay0,
                                        (float)ay1, pm);
        }
        if constexpr (!SUBS) {
          const uint64_t bm = __ballot(ov);
          if (ov) s_list[w][n + lanes_below(bm)] = (uint8_t)j;
          n += (uint32_t)__popcll(bm);
        } else {
          // the same exact test per 4x4 sub-block, one list each
#pragma unroll
          for (int sb = 0; sb < 4; ++sb) {
            // a sub-block whose pixels are all opaque takes no entries
            bool o = ov && ((alive >> (16 * sb)) & 0xFFFFull) != 0ull;
            if (o) {
              const uint32_t bx0 = qx0 + (sb & 1) * 4u, by0 = qy0 + (sb >> 1) * 4u;
              const uint2 bx = s_box[buf][j];
              o = bx0 < W && by0 < H && (bx.x & 0xFFFFu) <= bx0 + 3u && (bx.y & 0xFFFFu) >= bx0 &&
                  (bx.x >> 16) <= by0 + 3u && (bx.y >> 16) >= by0;
              if (o) {
                const uint32_t bx1 = min(bx0 + 3u, W - 1u), by1 = min(by0 + 3u, H - 1u);
                o = range_meets<UNALIGNED>(s_ent[buf][0][j], s_ent[buf][1][j], (float)bx0, (float)bx1, (float)by0,
                                           (float)by1, pm);
              }
            }
            const uint64_t bm = __ballot(o);
            if (o) s_sub[w][sb][ns[sb] + lanes_below(bm)] = (uint8_t)j;
            ns[sb] += (uint32_t)__popcll(bm);
          }
        }
      }
      __builtin_amdgcn_wave_barrier();
      // one fragment at this lane's pixel: coverage, alpha (shaders.wgsl:585-589)
      // and the blend; identical arithmetic on every walk, so they give
      // identical pixels
      auto covers_ = [&](const float4& m, const float4& t, float& u, float& w2) -> bool {
        return covers<UNALIGNED>(m, t, pxf, pyf, pm, u, w2);
      };
      auto shade = [&](float u, float w2, float opacity, const float4& col, bool in) {
        pix.shade(u, w2, opacity, col, in);
      };
      if constexpr (SUBS) {
        // entry-major per 4x4 sub-block: the four lane groups step through
        // their own lists side by side, so one step blends up to four entries
        const uint32_t mine = g == 0 ? ns[0] : g == 1 ? ns[1] : g == 2 ? ns[2] : ns[3];
        const uint32_t most = max(max(ns[0], ns[1]), max(ns[2], ns[3]));
        for (uint32_t i = 0; i < most; i += U) {
          uint32_t js;
          if constexpr (U == 4) js = *reinterpret_cast<const uint32_t*>(&s_sub[w][g][i]);
          else js = *reinterpret_cast<const uint16_t*>(&s_sub[w][g][i]);
          float4 m[U], t[U], col[U];
#pragma unroll
          for (int k = 0; k < U; ++k) {
            // slots past the list may be stale or, on a block's first batch,
            // uninitialised: masked into [0, B) they still index s_ent, and
            // their fragments are never blended (i + k < mine)
            const uint32_t j = (js >> (8 * k)) & (uint32_t)(B - 1);
            m[k] = s_ent[buf][0][j];
            t[k] = s_ent[buf][1][j];
            col[k] = s_ent[buf][2][j];
          }
#pragma unroll
          for (int k = 0; k < U; ++k) {
            float u, w2;
            if (i + k < mine && covers_(m[k], t[k], u, w2) && pix.tr != 0.0f) shade(u, w2, t[k].w, col[k], true);
          }
        }
      } else if constexpr (LANES) {
        // per-lane walk: for 32 list slots at a time the wave first copies
        // the slots' entries into a compact run (its 32 entries of the
        // other staging buffer, free until the next batch is staged), then
        // walks the run (walk_run: each pixel only its covering entries)
        float4* const cm = &s_ent[buf ^ 1][0][32 * w];
        float4* const ct = &s_ent[buf ^ 1][1][32 * w];
        float4* const cc = &s_ent[buf ^ 1][2][32 * w];
        const float fqx0 = (float)qx0, fqy0 = (float)qy0;
        const float span = (float)W + (float)H;  // bounds the pixel coordinates' magnitudes
        for (uint32_t c = 0; c < n; c += 32) {
          const uint32_t cnt = min(32u, n - c);
          if (lane < cnt) {
            const uint32_t j = (uint32_t)s_list[w][c + lane] & (uint32_t)(B - 1);
            cm[lane] = s_ent[buf][0][j];
            ct[lane] = s_ent[buf][1][j];
            cc[lane] = s_ent[buf][2][j];
          }
          __builtin_amdgcn_wave_barrier();
          walk_run<UNALIGNED, F32>(cm, ct, cc, &s_rows[w][0], cnt, lane, inside, pxf, pyf, fqx0, fqy0, span, pm,
                                   margin, pix);
          __builtin_amdgcn_wave_barrier();
        }
      } else if constexpr (!WALK) {
        // entry-major: the wave steps through its list, every lane testing
        // its pixel; four list slots per step so their LDS reads go together
        for (uint32_t i = 0; i < n; i += U) {
          uint32_t js;
          if constexpr (U == 4) js = *reinterpret_cast<const uint32_t*>(&s_list[w][i]);
          else js = *reinterpret_cast<const uint16_t*>(&s_list[w][i]);
          float4 m[U], t[U], col[U];
#pragma unroll
          for (int k = 0; k < U; ++k) {
            // slots past n may be stale or, on a block's first batch,
            // uninitialised: masked into [0, B) they still index s_ent, and
            // the loop below stops at n before using them
            const uint32_t j = (js >> (8 * k)) & (uint32_t)(B - 1);
            m[k] = s_ent[buf][0][j];
            t[k] = s_ent[buf][1][j];
            col[k] = s_ent[buf][2][j];
          }
#pragma unroll
          for (int k = 0; k < U; ++k) {
            if (i + k >= n) break;
            float u, w2;
            if (covers_(m[k], t[k], u, w2) && pix.tr != 0.0f) shade(u, w2, t[k].w, col[k], true);
          }
        }
      } else {
        // pixel-major: for 64 list slots at a time every lane first marks the
        // entries that cover its pixel, then walks only those, in order, so
        // lanes are not held up by entries that cover few of the 64 pixels
        for (uint32_t c = 0; c < n; c += 64) {
          const uint32_t cnt = min(64u, n - c);
          uint32_t lo = 0, hi = 0;
          for (uint32_t k = 0; k < cnt; k += 4) {
            const uint32_t js = *reinterpret_cast<const uint32_t*>(&s_list[w][c + k]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const uint32_t j = (js >> (8 * q)) & 255u;
              float u, w2;
              const uint32_t in = covers_(s_ent[buf][0][j], s_ent[buf][1][j], u, w2) && k + q < cnt ? 1u : 0u;
              if (k < 32) lo |= in << (k + q);
              else hi |= in << (k + q - 32);
            }
          }
          if (pix.tr == 0.0f) lo = hi = 0u;
          while (lo | hi) {
            uint32_t k;
            if (lo) { k = (uint32_t)__builtin_ctz(lo); lo &= lo - 1u; }
            else { k = 32u + (uint32_t)__builtin_ctz(hi); hi &= hi - 1u; }
            const uint32_t j = s_list[w][c + k];
            const float4 t = s_ent[buf][1][j];
            float u, w2;
            (void)covers_(s_ent[buf][0][j], t, u, w2);
            shade(u, w2, t.w, s_ent[buf][2][j], true);
            if (pix.tr == 0.0f) lo = hi = 0u;
          }
        }
      }
      alive = __ballot(inside && pix.tr != 0.0f);
      wave_done = alive == 0ull;
    }
    if (more) {
      // mode 3's compact runs live in the other buffer until every wave has
      // walked
      if constexpr (LANES) __syncthreads();
      stage(buf ^ 1);
    }
    buf ^= 1;
    if (__syncthreads_and(wave_done)) break;
  }
  if (a.tile_work && v == 0 && tid == 0) a.tile_work[tile] = make_uint2(range.y - range.x, consumed);
  if (inside) a.pixels[v][(size_t)py * W + px] = pix.packed(a.bgra != 0);
}

void launch_bin(const BinLaunch& a, hipStream_t s) {
  hipLaunchKernelGGL((k_bin<kBinThreads, kBinItems>), dim3(a.grid_x, a.views), dim3(kBinThreads), 0, s, a);
}

void launch_ranges(const RangesLaunch& a, hipStream_t s) {
  hipLaunchKernelGGL(k_tile_ranges, dim3(a.grid_x, a.views), dim3(256), 0, s, a);
}

template <bool UNALIGNED, bool F32, int MODE>
void launch_raster_as(bool small, dim3 grid, dim3 block, hipStream_t s, const RasterLaunch& a) {
  if (small) hipLaunchKernelGGL((k_raster<UNALIGNED, F32, MODE, 128>), grid, block, 0, s, a);
  else hipLaunchKernelGGL((k_raster<UNALIGNED, F32, MODE, 256>), grid, block, 0, s, a);
}

template <bool UNALIGNED, bool F32>
void launch_raster_mode(bool small, dim3 grid, dim3 block, hipStream_t s, const RasterLaunch& a) {
  if (a.walk == 3) launch_raster_as<UNALIGNED, F32, 3>(small, grid, block, s, a);
  else if (a.walk == 2) launch_raster_as<UNALIGNED, F32, 2>(small, grid, block, s, a);
  else if (a.walk == 1) launch_raster_as<UNALIGNED, F32, 1>(small, grid, block, s, a);
  else launch_raster_as<UNALIGNED, F32, 0>(small, grid, block, s, a);
}

void launch_raster(const RasterLaunch& a, hipStream_t s) {
  const dim3 grid(a.grid_x), block(256);
  const bool small = a.batch_entries == 128;
  if (a.use_unaligned_rectangles) {
    if (a.blend_f32) launch_raster_mode<true, true>(small, grid, block, s, a);
    else launch_raster_mode<true, false>(small, grid, block, s, a);
  } else {
    if (a.blend_f32) launch_raster_mode<false, true>(small, grid, block, s, a);
    else launch_raster_mode<false, false>(small, grid, block, s, a);
  }
}

}  // namespace spl
