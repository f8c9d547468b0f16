// api.cpp -- the C ABI of include/splatter.h: device, renderer, scene and
// frame objects, the per-frame host math of renderer.rs:364-409 and the
// launch sequence of the splat pipeline.  Built with -ffp-contract=off so
// the host-side f32 camera math rounds as the reference's Rust does.
#include "splatter.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "spl_internal.h"

using namespace spl;

namespace {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

#define HIP_TRY(expr)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(SPL_ERROR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <typename T>
hipError_t dev_alloc(T** p, size_t count) {
  *p = nullptr;
  if (count == 0) count = 1;
  return hipMalloc((void**)p, count * sizeof(T));
}

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

struct spl_device {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
};

struct spl_scene {
  spl_device* dev = nullptr;
  uint64_t count = 0;
  float* centers = nullptr;
  float4* rotations = nullptr;
  float4* scale_alpha = nullptr;
  float* sh = nullptr;
};

struct spl_frame {
  spl_device* dev = nullptr;
  uint32_t width = 0, height = 0;
  uint32_t* pixels = nullptr;
};

struct spl_renderer {
  spl_device* dev = nullptr;
  spl_configuration cfg{};
  uint64_t cap = 0;
  int depth_bits = 8, depth_passes = 4;
  uint32_t* keys[2] = {nullptr, nullptr};
  uint32_t* vals[2] = {nullptr, nullptr};
  uint32_t* compact_idx = nullptr;
  float4* records = nullptr;
  uint64_t pair_cap = 0;
  uint32_t* pair_keys[2] = {nullptr, nullptr};
  uint32_t* pair_vals[2] = {nullptr, nullptr};
  uint32_t* cull_status = nullptr;
  size_t cull_status_words = 0;
  uint32_t* sort_status = nullptr;
  size_t sort_status_per_pass = 0;
  uint64_t* bin_status = nullptr;
  size_t bin_status_words = 0;
  uint32_t* pair_status = nullptr;
  size_t pair_status_per_pass = 0;
  uint32_t* counters = nullptr;
  uint32_t* depth_hist = nullptr;
  uint32_t* pair_hist = nullptr;
  uint2* ranges = nullptr;
  uint32_t ranges_cap = 0;
  double* conics = nullptr;
  bool conics_on = false;
  uint2* tile_work = nullptr;
  uint32_t tile_work_cap = 0;
  bool tile_stats_on = false;
  uint32_t last_tiles = 0;
  // last frame
  bool last_sorted = false;
  int last_final = 0;
  bool have_frame = false;
  // stage timing: one set of 8 events per frame, read out in bulk
  bool timing = false;
  std::vector<std::array<hipEvent_t, 8>> ev_pool;
  size_t ev_used = 0;
  double stage_sum_ms[7] = {};
  uint64_t stage_frames = 0;
};

extern "C" {

const char* spl_last_error(void) { return g_error.c_str(); }

void spl_configuration_default(spl_configuration* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->frame_format = SPL_FRAME_FORMAT_BGRA8_UNORM;
  c->depth_sorting = SPL_DEPTH_SORTING_GPU;
  c->use_covariance_for_scale = 1;
  c->use_unaligned_rectangles = 1;
  c->spherical_harmonics_order = 2;
  c->max_splat_count = 1024ull * 1024ull * 4ull;
  c->radix_bits_per_digit = 8;
  c->frustum_culling_tolerance = 1.1f;
  c->ellipse_margin = 2.0f;
  c->splat_scale = 1.0f;
  c->field_of_view_y = 3.14159265358979323846f * 0.5f;
  c->blend_mode = SPL_BLEND_UNORM8;
  c->tile_pairs_per_splat = 4;
}

int spl_device_open(int ordinal, spl_device** out) {
  if (!out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_device_open: null output");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (ordinal < 0 || ordinal >= n) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_device_open: no such device");
  HIP_TRY(hipSetDevice(ordinal));
  auto* d = new (std::nothrow) spl_device();
  if (!d) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_device_open: host allocation");
  d->ordinal = ordinal;
  hipError_t e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete d;
    return fail(SPL_ERROR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  (void)hipDeviceGetAttribute(&d->cu_count, hipDeviceAttributeMultiprocessorCount, ordinal);
  if (d->cu_count <= 0) d->cu_count = 256;
  *out = d;
  return SPL_OK;
}

int spl_device_close(spl_device* d) {
  if (!d) return SPL_OK;
  (void)hipSetDevice(d->ordinal);
  (void)hipStreamSynchronize(d->stream);
  (void)hipStreamDestroy(d->stream);
  delete d;
  return SPL_OK;
}

int spl_synchronize(spl_device* d) {
  if (!d) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_synchronize: null device");
  HIP_TRY(hipSetDevice(d->ordinal));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return SPL_OK;
}

// ---------------------------------------------------------------------------
// Renderer
// ---------------------------------------------------------------------------

static void renderer_free(spl_renderer* r) {
  for (int i = 0; i < 2; ++i) {
    dev_free(r->keys[i]);
    dev_free(r->vals[i]);
    dev_free(r->pair_keys[i]);
    dev_free(r->pair_vals[i]);
  }
  dev_free(r->compact_idx);
  dev_free(r->records);
  dev_free(r->cull_status);
  dev_free(r->sort_status);
  dev_free(r->bin_status);
  dev_free(r->pair_status);
  dev_free(r->counters);
  dev_free(r->depth_hist);
  dev_free(r->pair_hist);
  dev_free(r->ranges);
  dev_free(r->conics);
  dev_free(r->tile_work);
  for (auto& set : r->ev_pool)
    for (auto& e : set)
      if (e) (void)hipEventDestroy(e), e = nullptr;
  r->ev_pool.clear();
  r->ev_used = 0;
}

int spl_renderer_create(spl_device* dev, const spl_configuration* cfg, spl_renderer** out) {
  if (!dev || !cfg || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_renderer_create: null argument");
  *out = nullptr;
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
  HIP_TRY(hipSetDevice(dev->ordinal));
  auto* r = new (std::nothrow) spl_renderer();
  if (!r) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_renderer_create: host allocation");
  r->dev = dev;
  r->cfg = *cfg;
  if (r->cfg.tile_pairs_per_splat == 0) r->cfg.tile_pairs_per_splat = 4;
  r->cap = cfg->max_splat_count;
  r->depth_bits = (int)cfg->radix_bits_per_digit;
  r->depth_passes = (30 + r->depth_bits - 1) / r->depth_bits;
  const int radix = 1 << r->depth_bits;
  const uint64_t pc = r->cap * (uint64_t)r->cfg.tile_pairs_per_splat;
  r->pair_cap = std::min<uint64_t>(pc, (1ull << 30) - (1ull << 20));
  r->cull_status_words = div_up(r->cap, kCullThreads * kCullItems);
  r->sort_status_per_pass = (size_t)div_up(r->cap, kSortThreads * kSortItems) * radix;
  r->bin_status_words = div_up(r->cap, kBinThreads * kBinItems);
  r->pair_status_per_pass = (size_t)div_up(r->pair_cap, kSortThreads * kSortItems) * 256;
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  for (int i = 0; i < 2; ++i) {
    chk(dev_alloc(&r->keys[i], r->cap));
    chk(dev_alloc(&r->vals[i], r->cap));
    chk(dev_alloc(&r->pair_keys[i], r->pair_cap));
    chk(dev_alloc(&r->pair_vals[i], r->pair_cap));
  }
  chk(dev_alloc(&r->compact_idx, r->cap));
  chk(dev_alloc(&r->records, r->cap * kRecordFloat4));
  chk(dev_alloc(&r->cull_status, r->cull_status_words));
  chk(dev_alloc(&r->sort_status, r->sort_status_per_pass * kMaxDepthPasses));
  chk(dev_alloc(&r->bin_status, r->bin_status_words));
  chk(dev_alloc(&r->pair_status, r->pair_status_per_pass * kMaxPairPasses));
  chk(dev_alloc(&r->counters, kCounterWords));
  chk(dev_alloc(&r->depth_hist, (size_t)kMaxDepthPasses * 1024));
  chk(dev_alloc(&r->pair_hist, (size_t)kMaxPairPasses * 256));
  if (e == hipSuccess) {
    // Every stage clears the bookkeeping of the stage before it once that
    // stage is done, so this is the only initialisation ever needed.
    chk(hipMemset(r->cull_status, 0, r->cull_status_words * 4));
    chk(hipMemset(r->sort_status, 0, r->sort_status_per_pass * kMaxDepthPasses * 4));
    chk(hipMemset(r->bin_status, 0, r->bin_status_words * 8));
    chk(hipMemset(r->pair_status, 0, r->pair_status_per_pass * kMaxPairPasses * 4));
    chk(hipMemset(r->counters, 0, kCounterWords * 4));
    chk(hipMemset(r->depth_hist, 0, (size_t)kMaxDepthPasses * 1024 * 4));
    chk(hipMemset(r->pair_hist, 0, (size_t)kMaxPairPasses * 256 * 4));
    chk(hipDeviceSynchronize());
  }
  if (e != hipSuccess) {
    renderer_free(r);
    delete r;
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
      return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_renderer_create: device memory for max_splat_count");
    return fail(SPL_ERROR_DEVICE, std::string("spl_renderer_create: ") + hipGetErrorString(e));
  }
  *out = r;
  return SPL_OK;
}

int spl_renderer_destroy(spl_renderer* r) {
  if (!r) return SPL_OK;
  (void)hipSetDevice(r->dev->ordinal);
  (void)hipStreamSynchronize(r->dev->stream);
  renderer_free(r);
  delete r;
  return SPL_OK;
}

int spl_renderer_get_config(const spl_renderer* r, spl_configuration* out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_renderer_get_config: null argument");
  *out = r->cfg;
  return SPL_OK;
}

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------

int spl_scene_create(spl_renderer* r, uint64_t splat_count, spl_scene** out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_create: null argument");
  *out = nullptr;
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  auto* s = new (std::nothrow) spl_scene();
  if (!s) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_scene_create: host allocation");
  s->dev = r->dev;
  s->count = std::min<uint64_t>(splat_count, r->cap);  // scene.rs:46
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t x) { if (e == hipSuccess) e = x; };
  const uint64_t slots = std::max<uint64_t>(s->count, 1);
  chk(dev_alloc(&s->centers, slots * 3));
  chk(dev_alloc(&s->rotations, slots));
  chk(dev_alloc(&s->scale_alpha, slots));
  chk(dev_alloc(&s->sh, slots * 48));
  if (e == hipSuccess) {
    // wgpu zero-initialises new buffers; so do we
    chk(hipMemset(s->centers, 0, slots * 12));
    chk(hipMemset(s->rotations, 0, slots * 16));
    chk(hipMemset(s->scale_alpha, 0, slots * 16));
    chk(hipMemset(s->sh, 0, slots * 192));
  }
  if (e != hipSuccess) {
    dev_free(s->centers);
    dev_free(s->rotations);
    dev_free(s->scale_alpha);
    dev_free(s->sh);
    delete s;
    return fail(e == hipErrorOutOfMemory ? SPL_ERROR_OUT_OF_MEMORY : SPL_ERROR_DEVICE,
                std::string("spl_scene_create: ") + hipGetErrorString(e));
  }
  *out = s;
  return SPL_OK;
}

int spl_scene_splat_count(const spl_scene* s, uint64_t* out) {
  if (!s || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_splat_count: null argument");
  *out = s->count;
  return SPL_OK;
}

int spl_scene_destroy(spl_scene* s) {
  if (!s) return SPL_OK;
  (void)hipSetDevice(s->dev->ordinal);
  (void)hipStreamSynchronize(s->dev->stream);
  dev_free(s->centers);
  dev_free(s->rotations);
  dev_free(s->scale_alpha);
  dev_free(s->sh);
  delete s;
  return SPL_OK;
}

namespace {

// scene.rs:17-24
struct SerializedSplat {
  float center[3];
  float n[3];
  float color[48];
  float alpha;
  float scale[3];
  float rotation[4];
};
static_assert(sizeof(SerializedSplat) == 248, "SerializedSplat is 62 f32");

constexpr float kMaxSizeVariance = 5.0f;  // scene.rs:13

// Host staging of decoded splats in the device layout.
struct Staging {
  std::vector<float> centers, sh;
  std::vector<float4> rot, sa;
  void resize(size_t n) {
    centers.resize(n * 3);
    sh.resize(n * 48);
    rot.resize(n);
    sa.resize(n);
  }
};

// scene.rs:225-249.  Rotor::signum (geometric_algebra 0.3.0) is r / |r|.
void decode(const SerializedSplat* in, size_t n, Staging& st) {
  st.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const SerializedSplat& s = in[i];
    const float* q = s.rotation;
    const float sq = ((q[0] * q[0] + q[1] * q[1]) + q[2] * q[2]) + q[3] * q[3];
    const float inv = 1.0f / sqrtf(sq);
    st.rot[i] = make_float4(q[0] * inv, q[1] * inv, q[2] * inv, q[3] * inv);
    for (int k = 0; k < 3; ++k) st.centers[3 * i + k] = s.center[k];
    const float average = ((s.scale[0] + s.scale[1]) + s.scale[2]) / 3.0f;
    float e[3];
    for (int k = 0; k < 3; ++k)
      e[k] = expf(fminf(fmaxf(s.scale[k], average - kMaxSizeVariance), average + kMaxSizeVariance));
    st.sa[i] = make_float4(e[0], e[1], e[2], 1.0f / (1.0f + expf(-s.alpha)));
    std::memcpy(&st.sh[48 * i], s.color, 48 * sizeof(float));
  }
}

int upload(spl_scene* s, uint64_t first, const Staging& st, size_t n) {
  if (n == 0) return SPL_OK;
  HIP_TRY(hipSetDevice(s->dev->ordinal));
  hipStream_t q = s->dev->stream;
  HIP_TRY(hipMemcpyAsync(s->centers + 3 * first, st.centers.data(), n * 12, hipMemcpyHostToDevice, q));
  HIP_TRY(hipMemcpyAsync(s->rotations + first, st.rot.data(), n * 16, hipMemcpyHostToDevice, q));
  HIP_TRY(hipMemcpyAsync(s->scale_alpha + first, st.sa.data(), n * 16, hipMemcpyHostToDevice, q));
  HIP_TRY(hipMemcpyAsync(s->sh + 48 * first, st.sh.data(), n * 192, hipMemcpyHostToDevice, q));
  HIP_TRY(hipStreamSynchronize(q));
  return SPL_OK;
}

}  // namespace

int spl_scene_load_records(spl_scene* s, uint64_t first, uint64_t count, const void* records) {
  if (!s || (!records && count)) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_load_records: null argument");
  if (first > s->count || count > s->count - first)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_load_records: range beyond splat_count");
  const auto* in = static_cast<const SerializedSplat*>(records);
  const size_t chunk = 1u << 18;
  Staging st;
  for (uint64_t done = 0; done < count; done += chunk) {
    const size_t n = (size_t)std::min<uint64_t>(chunk, count - done);
    decode(in + done, n, st);
    int rc = upload(s, first + done, st, n);
    if (rc) return rc;
  }
  return SPL_OK;
}

int spl_scene_new_dev_test(spl_renderer* r, spl_scene** out) {
  int rc = spl_scene_create(r, 3, out);
  if (rc) return rc;
  spl_scene* s = *out;
  // scene.rs:159-177: raw values, not decoded
  Staging st;
  st.resize(3);
  for (int i = 0; i < 3; ++i) {
    st.rot[i] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
    st.centers[3 * i + 0] = i == 0 ? 1.0f : 0.0f;
    st.centers[3 * i + 1] = i == 1 ? 1.0f : 0.0f;
    st.centers[3 * i + 2] = i == 2 ? 1.0f : 0.0f;
    st.sa[i] = make_float4(0.2f, 0.1f, 0.05f, 1.0f);
    for (int k = 0; k < 48; ++k) st.sh[48 * i + k] = 0.0f;
    st.sh[48 * i + 0] = i == 0 ? 1.0f : 0.0f;
    st.sh[48 * i + 1] = i == 1 ? 1.0f : 0.0f;
    st.sh[48 * i + 2] = i == 2 ? 1.0f : 0.0f;
  }
  rc = upload(s, 0, st, (size_t)std::min<uint64_t>(3, s->count));
  if (rc) {
    spl_scene_destroy(s);
    *out = nullptr;
  }
  return rc;
}

int spl_parse_file_header(const char* path, uint16_t* header_size, uint64_t* splat_count) {
  if (!path || !header_size || !splat_count) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_parse_file_header: null argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(SPL_ERROR_IO, std::string("cannot open ") + path);
  uint64_t count = 0;
  uint64_t consumed = 0;
  std::string line;
  // scene.rs:201-210: line by line until "end_header\n"
  while (true) {
    line.clear();
    char ch;
    bool any = false;
    while (f.get(ch)) {
      any = true;
      line.push_back(ch);
      if (ch == '\n') break;
    }
    if (!any) return fail(SPL_ERROR_IO, "end_header not found");
    consumed += line.size();
    if (line == "end_header\n") break;
    if (line.rfind("element vertex ", 0) == 0 && line.size() > 16) {
      const std::string num = line.substr(15, line.size() - 16);
      char* end = nullptr;
      unsigned long long v = std::strtoull(num.c_str(), &end, 10);
      if (!end || *end != '\0') return fail(SPL_ERROR_IO, "bad element vertex line");
      count = v;
    }
  }
  if (consumed > 0xFFFF) return fail(SPL_ERROR_IO, "header longer than 65535 bytes");
  *header_size = (uint16_t)consumed;
  *splat_count = count;
  return SPL_OK;
}

int spl_scene_load_chunk(spl_scene* s, const char* path, uint16_t header_size, uint64_t begin, uint64_t end) {
  if (!s || !path) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_load_chunk: null argument");
  end = std::min<uint64_t>(end, s->count);  // scene.rs:217
  if (begin >= end) return SPL_OK;
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(SPL_ERROR_IO, std::string("cannot open ") + path);
  f.seekg((std::streamoff)header_size + (std::streamoff)(begin * sizeof(SerializedSplat)));
  const size_t chunk = 1u << 18;
  std::vector<SerializedSplat> buf;
  Staging st;
  for (uint64_t at = begin; at < end; at += chunk) {
    const size_t n = (size_t)std::min<uint64_t>(chunk, end - at);
    buf.resize(n);
    f.read(reinterpret_cast<char*>(buf.data()), (std::streamsize)(n * sizeof(SerializedSplat)));
    if (!f) return fail(SPL_ERROR_IO, "short read in load_chunk");
    decode(buf.data(), n, st);
    int rc = upload(s, at, st, n);
    if (rc) return rc;
  }
  return SPL_OK;
}

// ---------------------------------------------------------------------------
// Frame
// ---------------------------------------------------------------------------

int spl_frame_create(spl_device* d, uint32_t w, uint32_t h, spl_frame** out) {
  if (!d || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_frame_create: null argument");
  *out = nullptr;
  if (w == 0 || h == 0 || w > 16384 || h > 16384) return fail(SPL_ERROR_INVALID_ARGUMENT, "frame size must be 1..=16384");
  HIP_TRY(hipSetDevice(d->ordinal));
  auto* f = new (std::nothrow) spl_frame();
  if (!f) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_frame_create: host allocation");
  f->dev = d;
  f->width = w;
  f->height = h;
  hipError_t e = dev_alloc(&f->pixels, (size_t)w * h);
  if (e != hipSuccess) {
    delete f;
    return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_frame_create: device memory");
  }
  *out = f;
  return SPL_OK;
}

int spl_frame_destroy(spl_frame* f) {
  if (!f) return SPL_OK;
  (void)hipSetDevice(f->dev->ordinal);
  (void)hipStreamSynchronize(f->dev->stream);
  dev_free(f->pixels);
  delete f;
  return SPL_OK;
}

int spl_frame_read(spl_frame* f, void* host, uint64_t size) {
  if (!f || !host) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_frame_read: null argument");
  const uint64_t need = (uint64_t)f->width * f->height * 4;
  if (size < need) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_frame_read: buffer too small");
  HIP_TRY(hipSetDevice(f->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(f->dev->stream));
  HIP_TRY(hipMemcpy(host, f->pixels, need, hipMemcpyDeviceToHost));
  return SPL_OK;
}

// ---------------------------------------------------------------------------
// Per-frame host math (renderer.rs:364-409, utils.rs:45-70)
// ---------------------------------------------------------------------------

static void mat4_multiplication(const float a[16], const float b[16], float out[16]) {
  float t[16];
  for (int j = 0; j < 4; ++j)
    for (int r = 0; r < 4; ++r)
      t[4 * j + r] = ((a[r] * b[4 * j + 0] + a[4 + r] * b[4 * j + 1]) + a[8 + r] * b[4 * j + 2]) + a[12 + r] * b[4 * j + 3];
  std::memcpy(out, t, sizeof(t));
}

static void frame_uniforms(const spl_configuration& cfg, const float cam[16], const float view[16], uint32_t W,
                           uint32_t H, Uniforms& u) {
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

static int ensure_ranges(spl_renderer* r, uint32_t tiles) {
  if (tiles <= r->ranges_cap) return SPL_OK;
  HIP_TRY(hipStreamSynchronize(r->dev->stream));
  dev_free(r->ranges);
  r->ranges_cap = 0;
  HIP_TRY(dev_alloc(&r->ranges, tiles));
  HIP_TRY(hipMemset(r->ranges, 0, (size_t)tiles * sizeof(uint2)));
  HIP_TRY(hipDeviceSynchronize());
  r->ranges_cap = tiles;
  return SPL_OK;
}

static int pair_passes_for(uint32_t tiles) {
  if (tiles <= 256) return 1;
  if (tiles <= 65536) return 2;
  return 3;
}

static int enqueue_frame(spl_renderer* r, spl_scene* s, spl_frame* f, const float cam[16], const float view[16]) {
  const spl_configuration& cfg = r->cfg;
  hipStream_t q = r->dev->stream;
  Uniforms u{};
  frame_uniforms(cfg, cam, view, f->width, f->height, u);
  const uint32_t tiles = u.tiles_x * u.tiles_y;
  int rc = ensure_ranges(r, tiles);
  if (rc) return rc;
  if (r->tile_stats_on && r->tile_work_cap < tiles) {
    HIP_TRY(hipStreamSynchronize(q));
    dev_free(r->tile_work);
    r->tile_work_cap = 0;
    HIP_TRY(dev_alloc(&r->tile_work, tiles));
    r->tile_work_cap = tiles;
  }
  r->last_tiles = tiles;
  const uint32_t n = (uint32_t)s->count;
  const bool sorted = cfg.depth_sorting != SPL_DEPTH_SORTING_NONE;
  const int P = sorted ? r->depth_passes : 0;
  const int bits = r->depth_bits;
  const int radix = 1 << bits;
  const int Q = pair_passes_for(tiles);
  const int cus = r->dev->cu_count;
  const size_t sort_words = (size_t)div_up(n, kSortThreads * kSortItems) * radix;
  const size_t cull_words = div_up(n, kCullThreads * kCullItems);
  hipEvent_t* ev = nullptr;
  if (r->timing) {
    if (r->ev_used == r->ev_pool.size()) {
      std::array<hipEvent_t, 8> set{};
      for (auto& e : set) HIP_TRY(hipEventCreate(&e));
      r->ev_pool.push_back(set);
    }
    ev = r->ev_pool[r->ev_used++].data();
  }
  auto mark = [&](int i) { if (ev) (void)hipEventRecord(ev[i], q); };

  mark(0);
  {
    CullLaunch a{};
    a.centers = s->centers;
    a.n = n;
    std::memcpy(a.vp, u.view_projection_matrix, 64);
    a.tol = u.frustum_culling_tolerance;
    a.gpu_predicate = sorted ? 0 : 1;
    a.keys_out = sorted ? r->keys[0] : nullptr;
    a.idx_out = r->compact_idx;
    a.hist = r->depth_hist;
    a.passes = P;
    a.bits = bits;
    a.status = r->cull_status;
    a.counters = r->counters;
    a.grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, kCullThreads * kCullItems), (uint32_t)cus * 4));
    launch_cull(a, q);
  }
  mark(1);
  for (int p = 0; p < P; ++p) {
    SortPassLaunch a{};
    a.keys_in = r->keys[p & 1];
    a.vals_in = p == 0 ? nullptr : r->vals[p & 1];
    a.keys_out = r->keys[(p + 1) & 1];
    a.vals_out = r->vals[(p + 1) & 1];
    a.count = r->counters + kVisible;
    a.count_cap = n;
    a.shift = p * bits;
    a.hist = r->depth_hist + (size_t)p * radix;
    a.status = r->sort_status + (size_t)p * r->sort_status_per_pass;
    a.tile_counter = r->counters + kCtrSort0 + p;
    a.done_counter = r->counters + kDone + kCtrSort0 + p;
    if (p == 0) {
      a.clear_a = r->cull_status;
      a.clear_a_words = cull_words;
    } else {
      a.clear_a = r->sort_status + (size_t)(p - 1) * r->sort_status_per_pass;
      a.clear_a_words = sort_words;
    }
    a.error_flag = r->counters + kScatterError;
    a.grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, kSortThreads * kSortItems), (uint32_t)cus * 4));
    launch_sort_pass(bits, a, q);
  }
  const int fin = P & 1;
  mark(2);
  {
    ProjectLaunch a{};
    a.scene = SceneView{s->centers, s->rotations, s->scale_alpha, s->sh, n};
    a.compact_idx = r->compact_idx;
    a.count = r->counters + kVisible;
    a.u = u;
    a.sh_order = (int)cfg.spherical_harmonics_order;
    a.use_covariance_for_scale = cfg.use_covariance_for_scale ? 1 : 0;
    a.use_unaligned_rectangles = cfg.use_unaligned_rectangles ? 1 : 0;
    a.records = r->records;
    a.conics = r->conics_on ? r->conics : nullptr;
    if (P > 0) {
      a.clear_a = r->sort_status + (size_t)(P - 1) * r->sort_status_per_pass;
      a.clear_a_words = sort_words;
      a.clear_b = r->depth_hist;
      a.clear_b_words = (size_t)kMaxDepthPasses * 1024;
    } else {
      a.clear_a = r->cull_status;
      a.clear_a_words = cull_words;
    }
    a.grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, 256), (uint32_t)cus * 8));
    launch_project(a, q);
  }
  mark(3);
  {
    BinLaunch a{};
    a.order = sorted ? r->vals[fin] : nullptr;
    a.records = r->records;
    a.count = r->counters + kVisible;
    a.tiles_x = u.tiles_x;
    a.pair_keys = r->pair_keys[0];
    a.pair_vals = r->pair_vals[0];
    a.pair_cap = (uint32_t)r->pair_cap;
    a.hist = r->pair_hist;
    a.pair_passes = Q;
    a.status = r->bin_status;
    a.counters = r->counters;
    a.grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(div_up(n, kBinThreads * kBinItems), (uint32_t)cus * 4));
    launch_bin(a, q);
  }
  mark(4);
  const size_t pair_words = r->pair_status_per_pass;
  for (int p = 0; p < Q; ++p) {
    SortPassLaunch a{};
    a.keys_in = r->pair_keys[p & 1];
    a.vals_in = r->pair_vals[p & 1];
    a.keys_out = r->pair_keys[(p + 1) & 1];
    a.vals_out = r->pair_vals[(p + 1) & 1];
    a.count = r->counters + kPairs;
    a.count_cap = (uint32_t)r->pair_cap;
    a.shift = 8 * p;
    a.hist = r->pair_hist + (size_t)p * 256;
    a.status = r->pair_status + (size_t)p * pair_words;
    a.tile_counter = r->counters + kCtrPair0 + p;
    a.done_counter = r->counters + kDone + kCtrPair0 + p;
    if (p == 0) {
      a.clear_a = reinterpret_cast<uint32_t*>(r->bin_status);
      a.clear_a_words = (size_t)div_up(n, kBinThreads * kBinItems) * 2;
    } else {
      a.clear_a = r->pair_status + (size_t)(p - 1) * pair_words;
      a.clear_a_words = pair_words;
    }
    a.error_flag = r->counters + kScatterError;
    a.grid = (int)std::max<uint32_t>(1, std::min<uint32_t>(div_up(r->pair_cap, kSortThreads * kSortItems), (uint32_t)cus * 4));
    launch_sort_pass(8, a, q);
  }
  const int pfin = Q & 1;
  mark(5);
  {
    RangesLaunch a{};
    a.pair_keys = r->pair_keys[pfin];
    a.counters = r->counters;
    a.pair_cap = (uint32_t)r->pair_cap;
    a.tiles = tiles;
    a.ranges = r->ranges;
    a.clear_a = r->pair_status + (size_t)(Q - 1) * pair_words;
    a.clear_a_words = pair_words;
    a.clear_b = r->pair_hist;
    a.clear_b_words = (size_t)kMaxPairPasses * 256;
    launch_ranges(a, (uint32_t)std::min<uint64_t>(r->pair_cap, (uint64_t)cus * 8 * 256 * 8), q);
  }
  mark(6);
  {
    RasterLaunch a{};
    a.pair_vals = r->pair_vals[pfin];
    a.visible = r->counters + kVisible;
    a.records = r->records;
    a.ranges = r->ranges;
    a.u = u;
    a.use_unaligned_rectangles = cfg.use_unaligned_rectangles ? 1 : 0;
    a.blend_f32 = cfg.blend_mode == SPL_BLEND_F32 ? 1 : 0;
    a.bgra = cfg.frame_format == SPL_FRAME_FORMAT_BGRA8_UNORM ? 1 : 0;
    a.pixels = f->pixels;
    a.tile_work = r->tile_stats_on ? r->tile_work : nullptr;
    launch_raster(a, q);
  }
  mark(7);
  HIP_TRY(hipGetLastError());
  r->last_sorted = sorted;
  r->last_final = fin;
  r->have_frame = true;
  return SPL_OK;
}

int spl_render_frame(spl_renderer* r, spl_scene* s, spl_frame* f, const float cam[16], const float view[16]) {
  if (!r || !s || !f || !cam || !view) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_frame: null argument");
  if (s->dev != r->dev || f->dev != r->dev) return fail(SPL_ERROR_INVALID_ARGUMENT, "scene, frame and renderer must share a device");
  if (s->count > r->cap) return fail(SPL_ERROR_CAPACITY, "scene has more splats than max_splat_count");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  return enqueue_frame(r, s, f, cam, view);
}

int spl_render_views(spl_renderer* r, spl_scene* s, uint32_t view_count, spl_frame* const* frames,
                     const float* cams, const float* views) {
  if (!r || !s || !frames || !cams || !views) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_views: null argument");
  for (uint32_t i = 0; i < view_count; ++i) {
    if (!frames[i]) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_views: null frame");
    int rc = spl_render_frame(r, s, frames[i], cams + 16 * (size_t)i, views + 16 * (size_t)i);
    if (rc) return rc;
  }
  return SPL_OK;
}

// ---------------------------------------------------------------------------
// Parity and timing taps
// ---------------------------------------------------------------------------

static int read_counter(spl_renderer* r, int slot, uint32_t* v) {
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->dev->stream));
  HIP_TRY(hipMemcpy(v, r->counters + slot, 4, hipMemcpyDeviceToHost));
  return SPL_OK;
}

int spl_debug_visible_count(spl_renderer* r, uint64_t* out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  uint32_t v = 0;
  int rc = read_counter(r, kVisible, &v);
  *out = v;
  return rc;
}

int spl_debug_pair_count(spl_renderer* r, uint64_t* out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  uint32_t v = 0, o = 0;
  int rc = read_counter(r, kPairs, &v);
  if (rc) return rc;
  rc = read_counter(r, kPairOverflow, &o);
  *out = v;
  if (rc) return rc;
  if (o) return fail(SPL_ERROR_CAPACITY, "screen-tile pair capacity exceeded in the last frame");
  return SPL_OK;
}

// order[c] for draw rank c: the compacted position of each drawn splat
static int draw_order(spl_renderer* r, std::vector<uint32_t>& order, std::vector<uint32_t>* keys) {
  uint32_t v = 0, err = 0;
  int rc = read_counter(r, kScatterError, &err);
  if (rc) return rc;
  if (err) return fail(SPL_ERROR_DEVICE, "radix sort scatter left its range (inconsistent histogram)");
  rc = read_counter(r, kVisible, &v);
  if (rc) return rc;
  order.resize(v);
  if (keys) keys->assign(v, 0u);
  if (v == 0) return SPL_OK;
  if (r->last_sorted) {
    HIP_TRY(hipMemcpy(order.data(), r->vals[r->last_final], (size_t)v * 4, hipMemcpyDeviceToHost));
    if (keys) HIP_TRY(hipMemcpy(keys->data(), r->keys[r->last_final], (size_t)v * 4, hipMemcpyDeviceToHost));
  } else {
    for (uint32_t i = 0; i < v; ++i) order[i] = i;
  }
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
  if (!order.empty()) HIP_TRY(hipMemcpy(idx.data(), r->compact_idx, idx.size() * 4, hipMemcpyDeviceToHost));
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
  if (!order.empty()) HIP_TRY(hipMemcpy(rec.data(), r->records, rec.size() * 16, hipMemcpyDeviceToHost));
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
  if (!r->conics_on || !r->have_frame) return fail(SPL_ERROR_INVALID_ARGUMENT, "conic capture not enabled");
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
  HIP_TRY(hipStreamSynchronize(r->dev->stream));
  *count = r->last_tiles;
  const size_t m = (size_t)std::min<uint64_t>(capacity, r->last_tiles);
  if (out && m) HIP_TRY(hipMemcpy(out, r->tile_work, m * sizeof(uint2), hipMemcpyDeviceToHost));
  return SPL_OK;
}

int spl_debug_enable_timing(spl_renderer* r, int enable) {
  if (!r) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->dev->stream));
  r->timing = enable != 0;
  r->ev_used = 0;
  for (double& v : r->stage_sum_ms) v = 0.0;
  r->stage_frames = 0;
  return SPL_OK;
}

int spl_debug_stage_times(spl_renderer* r, float* ms, uint32_t capacity) {
  if (!r || !ms) return fail(SPL_ERROR_INVALID_ARGUMENT, "null argument");
  HIP_TRY(hipSetDevice(r->dev->ordinal));
  HIP_TRY(hipStreamSynchronize(r->dev->stream));
  for (size_t f = 0; f < r->ev_used; ++f) {
    for (int i = 0; i < 7; ++i) {
      float t = 0.0f;
      HIP_TRY(hipEventElapsedTime(&t, r->ev_pool[f][i], r->ev_pool[f][i + 1]));
      r->stage_sum_ms[i] += t;
    }
  }
  r->stage_frames += r->ev_used;
  r->ev_used = 0;
  for (uint32_t i = 0; i < capacity && i < 7; ++i)
    ms[i] = r->stage_frames ? (float)(r->stage_sum_ms[i] / (double)r->stage_frames) : 0.0f;
  return SPL_OK;
}

}  // extern "C"
