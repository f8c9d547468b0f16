// api.cpp -- the C ABI of include/splatter.h for devices, scenes and frames:
// HIP device and stream, the scene's device layout and the record decode of
// scene.rs:225-249 (load_chunk), the PLY header parse of scene.rs:197-213,
// and render targets.  The renderer lives in renderer.cpp.  Scene uploads
// read and decode records on host threads into pinned staging and copy 2^18
// records at a time, one chunk's copy overlapping the next chunk's decode.
#include "splatter.h"

#include <hip/hip_runtime.h>

#include <fcntl.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <functional>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "spl_internal.h"
#include "spl_objects.h"

using namespace spl;

namespace spl {

thread_local std::string g_error;

int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}

}  // namespace spl

extern "C" {

const char* spl_last_error(void) { return spl::g_error.c_str(); }

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
  c->max_views = 1;
  c->use_hip_graphs = 1;
  c->sort_tile_keys = 4096;
  c->raster_mode = 3;
  c->raster_batch = 128;
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
  for (int i = 0; i < kPoolStreams; ++i) {
    e = hipStreamCreateWithFlags(&d->pool[i], hipStreamNonBlocking);
    if (e != hipSuccess) {
      for (int j = 0; j < i; ++j) (void)hipStreamDestroy(d->pool[j]);
      (void)hipStreamDestroy(d->stream);
      delete d;
      return fail(SPL_ERROR_DEVICE, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
  }
  (void)hipDeviceGetAttribute(&d->cu_count, hipDeviceAttributeMultiprocessorCount, ordinal);
  if (d->cu_count <= 0) d->cu_count = 256;
  const int rc = spl::static_scratch_setup(d);  // with the device current
  if (rc) {
    for (int j = 0; j < kPoolStreams; ++j) (void)hipStreamDestroy(d->pool[j]);
    (void)hipStreamDestroy(d->stream);
    delete d;
    return rc;
  }
  *out = d;
  return SPL_OK;
}

int spl_device_close(spl_device* d) {
  if (!d) return SPL_OK;
  (void)hipSetDevice(d->ordinal);
  (void)hipStreamSynchronize(d->stream);
  spl::device_release(d);
  return SPL_OK;
}

int spl_synchronize(spl_device* d) {
  if (!d) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_synchronize: null device");
  HIP_TRY(hipSetDevice(d->ordinal));
  HIP_TRY(hipStreamSynchronize(d->stream));
  return SPL_OK;
}

// ---------------------------------------------------------------------------
// Scene
// ---------------------------------------------------------------------------

int spl_scene_create(spl_renderer* r, uint64_t splat_count, spl_scene** out) {
  if (!r || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_create: null argument");
  return scene_create(renderer_device(r), std::min<uint64_t>(splat_count, renderer_capacity(r)), out);  // scene.rs:46
}

}  // extern "C"

void spl::device_retain(spl_device* d) { d->refs.fetch_add(1); }

void spl::device_release(spl_device* d) {
  if (d->refs.fetch_sub(1) != 1) return;
  (void)hipSetDevice(d->ordinal);
  (void)hipStreamSynchronize(d->stream);
  (void)hipStreamDestroy(d->stream);
  for (hipStream_t s : d->pool) {
    (void)hipStreamSynchronize(s);
    (void)hipStreamDestroy(s);
  }
  delete d;
}

hipStream_t spl::acquire_stream(spl_device* d) {
  std::lock_guard<std::mutex> lock(d->pool_mu);
  int best = 0;
  for (int i = 1; i < kPoolStreams; ++i)
    if (d->pool_users[i] < d->pool_users[best]) best = i;
  ++d->pool_users[best];
  return d->pool[best];
}

void spl::release_stream(spl_device* d, hipStream_t s) {
  std::lock_guard<std::mutex> lock(d->pool_mu);
  for (int i = 0; i < kPoolStreams; ++i)
    if (d->pool[i] == s && d->pool_users[i] > 0) {
      --d->pool_users[i];
      return;
    }
}

int spl::scene_create(spl_device* dev, uint64_t count, spl_scene** out) {
  *out = nullptr;
  if (count > (1ull << 30)) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_create: more than 2^30 splats");
  HIP_TRY(hipSetDevice(dev->ordinal));
  auto* s = new (std::nothrow) spl_scene();
  if (!s) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_scene_create: host allocation");
  // A destroyed scene's host address (and device buffers) can come back for
  // the next scene; the id never does, so a renderer's cached launch
  // sequences, which bake the scene's device pointers in, are keyed on it.
  static std::atomic<uint64_t> next_uid{1};
  s->dev = dev;
  s->uid = next_uid.fetch_add(1, std::memory_order_relaxed);
  s->count = count;
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
    // the uploads go to the device's (non-blocking) stream: finish the
    // clears before any of them can start
    chk(hipDeviceSynchronize());
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
  device_retain(dev);
  *out = s;
  return SPL_OK;
}

extern "C" {

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
  spl_device* d = s->dev;
  delete s;
  spl::device_release(d);
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

// Pinned host staging of decoded splats in the device layout, one slot per
// chunk in flight.  `done` is recorded on the device stream after the slot's
// copies, so the slot is refilled only once its previous chunk has landed.
struct StagingSlot {
  void* block = nullptr;
  float* sh = nullptr;       // 48 n
  float4* rot = nullptr;     // n
  float4* sa = nullptr;      // n
  float* centers = nullptr;  // 3 n
  hipEvent_t done = nullptr;
  bool pending = false;
};

// scene.rs:225-249 for records [lo, hi).  Rotor::signum (geometric_algebra
// 0.3.0) is r / |r|.
void decode_range(const SerializedSplat* in, size_t lo, size_t hi, StagingSlot& st) {
  for (size_t i = lo; i < hi; ++i) {
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
    // each coefficient times its basis constant (the product the lookup starts with)
    for (int k = 0; k < 48; ++k) st.sh[48 * i + k] = kShConstants[k / 3] * s.color[k];
  }
}

// Runs fn(lo, hi) over [0, n) split into contiguous ranges on up to 16 host
// threads (the calling thread takes the first range).
void parallel_ranges(size_t n, const std::function<void(size_t, size_t)>& fn) {
  constexpr size_t kPerThread = 1u << 14;
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const size_t workers = std::min<size_t>({(size_t)std::min(hw, 16u), (n + kPerThread - 1) / kPerThread, 16});
  if (workers <= 1) {
    fn((size_t)0, n);
    return;
  }
  const size_t step = (n + workers - 1) / workers;
  std::vector<std::thread> pool;
  pool.reserve(workers - 1);
  size_t w = 1;
  try {
    for (; w < workers; ++w) {
      const size_t lo = std::min(n, w * step), hi = std::min(n, lo + step);
      pool.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
  } catch (const std::system_error&) {
    // no thread to be had: the calling thread takes the remaining ranges
  }
  for (size_t r = w; r < workers; ++r) fn(std::min(n, r * step), std::min(n, r * step + step));
  fn((size_t)0, std::min(n, step));
  for (auto& t : pool) t.join();
}

// Each record is decoded by one thread with the same arithmetic, so the result
// does not depend on the split.
void decode(const SerializedSplat* in, size_t n, StagingSlot& st) {
  parallel_ranges(n, [&](size_t lo, size_t hi) { decode_range(in, lo, hi, st); });
}

// Reads `bytes` at `offset` of fd; false on an error or end of file first.
bool read_at(int fd, void* dst, size_t bytes, uint64_t offset) {
  char* p = static_cast<char*>(dst);
  while (bytes) {
    const ssize_t got = pread(fd, p, bytes, (off_t)offset);
    if (got < 0 && errno == EINTR) continue;
    if (got <= 0) return false;
    p += got;
    bytes -= (size_t)got;
    offset += (uint64_t)got;
  }
  return true;
}

// Streams decoded chunks to a scene's device arrays.  Two pinned slots
// alternate: the copies of chunk k run on the device stream while chunk k+1 is
// decoded on the host.  finish() waits for the last copy; the destructor also
// waits before the slots are freed.
class Uploader {
 public:
  static constexpr size_t kChunk = 1u << 18;
  static constexpr size_t kBytesPerSplat = 48 * 4 + 16 + 16 + 12;

  Uploader(spl_scene* s, size_t max_chunk) : s_(s), cap_(std::max<size_t>(1, std::min(kChunk, max_chunk))) {}
  Uploader(const Uploader&) = delete;
  Uploader& operator=(const Uploader&) = delete;

  ~Uploader() {
    (void)hipSetDevice(s_->dev->ordinal);
    for (auto& sl : slot_) {
      if (sl.done) {
        if (sl.pending) (void)hipEventSynchronize(sl.done);
        (void)hipEventDestroy(sl.done);
      }
      if (sl.block) (void)hipHostFree(sl.block);
    }
  }

  size_t capacity() const { return cap_; }

  // The next slot, once the copies that last read it have finished.
  int acquire(StagingSlot** out) {
    HIP_TRY(hipSetDevice(s_->dev->ordinal));
    StagingSlot& sl = slot_[next_];
    if (!sl.block) {
      hipError_t e = hipHostMalloc(&sl.block, cap_ * kBytesPerSplat, hipHostMallocDefault);
      if (e != hipSuccess) {
        sl.block = nullptr;
        return fail(SPL_ERROR_OUT_OF_MEMORY, std::string("pinned staging: ") + hipGetErrorString(e));
      }
      char* b = static_cast<char*>(sl.block);
      sl.sh = reinterpret_cast<float*>(b);
      sl.rot = reinterpret_cast<float4*>(b + cap_ * 192);
      sl.sa = reinterpret_cast<float4*>(b + cap_ * 208);
      sl.centers = reinterpret_cast<float*>(b + cap_ * 224);
      HIP_TRY(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    }
    if (sl.pending) {
      HIP_TRY(hipEventSynchronize(sl.done));
      sl.pending = false;
    }
    *out = &sl;
    return SPL_OK;
  }

  // Queues the copies of the acquired slot's first n splats to [first, first+n).
  int send(StagingSlot& sl, uint64_t first, size_t n) {
    next_ ^= 1;
    if (n == 0) return SPL_OK;
    hipStream_t q = s_->dev->stream;
    HIP_TRY(hipMemcpyAsync(s_->centers + 3 * first, sl.centers, n * 12, hipMemcpyHostToDevice, q));
    HIP_TRY(hipMemcpyAsync(s_->rotations + first, sl.rot, n * 16, hipMemcpyHostToDevice, q));
    HIP_TRY(hipMemcpyAsync(s_->scale_alpha + first, sl.sa, n * 16, hipMemcpyHostToDevice, q));
    HIP_TRY(hipMemcpyAsync(s_->sh + 48 * first, sl.sh, n * 192, hipMemcpyHostToDevice, q));
    HIP_TRY(hipEventRecord(sl.done, q));
    sl.pending = true;
    return SPL_OK;
  }

  // Decodes n <= capacity() records and queues their upload to [first, first+n).
  int push(uint64_t first, const SerializedSplat* in, size_t n) {
    StagingSlot* sl = nullptr;
    int rc = acquire(&sl);
    if (rc) return rc;
    decode(in, n, *sl);
    return send(*sl, first, n);
  }

  // As push, for n records read from fd at `offset` into buf: each thread
  // reads and decodes its own range.
  int push_file(uint64_t first, int fd, uint64_t offset, SerializedSplat* buf, size_t n) {
    StagingSlot* sl = nullptr;
    int rc = acquire(&sl);
    if (rc) return rc;
    std::atomic<bool> ok{true};
    parallel_ranges(n, [&](size_t lo, size_t hi) {
      if (!read_at(fd, buf + lo, (hi - lo) * sizeof(SerializedSplat), offset + lo * sizeof(SerializedSplat))) {
        ok = false;
        return;
      }
      decode_range(buf, lo, hi, *sl);
    });
    if (!ok) return fail(SPL_ERROR_IO, "short read in load_chunk");
    return send(*sl, first, n);
  }

  int finish() {
    HIP_TRY(hipSetDevice(s_->dev->ordinal));
    HIP_TRY(hipStreamSynchronize(s_->dev->stream));
    for (auto& sl : slot_) sl.pending = false;
    return SPL_OK;
  }

 private:
  spl_scene* s_;
  size_t cap_;
  StagingSlot slot_[2];
  int next_ = 0;
};

}  // namespace

int spl_scene_load_records(spl_scene* s, uint64_t first, uint64_t count, const void* records) {
  if (!s || (!records && count)) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_load_records: null argument");
  if (first > s->count || count > s->count - first)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_scene_load_records: range beyond splat_count");
  if (count == 0) return SPL_OK;
  const auto* in = static_cast<const SerializedSplat*>(records);
  Uploader up(s, (size_t)std::min<uint64_t>(count, Uploader::kChunk));
  for (uint64_t done = 0; done < count; done += up.capacity()) {
    const size_t n = (size_t)std::min<uint64_t>(up.capacity(), count - done);
    int rc = up.push(first + done, in + done, n);
    if (rc) return rc;
  }
  return up.finish();
}

int spl_scene_new_dev_test(spl_renderer* r, spl_scene** out) {
  int rc = spl_scene_create(r, 3, out);
  if (rc) return rc;
  spl_scene* s = *out;
  // scene.rs:159-177: raw values, not decoded
  Uploader up(s, 3);
  StagingSlot* st = nullptr;
  rc = up.acquire(&st);
  if (!rc) {
    for (int i = 0; i < 3; ++i) {
      st->rot[i] = make_float4(1.0f, 0.0f, 0.0f, 0.0f);
      st->centers[3 * i + 0] = i == 0 ? 1.0f : 0.0f;
      st->centers[3 * i + 1] = i == 1 ? 1.0f : 0.0f;
      st->centers[3 * i + 2] = i == 2 ? 1.0f : 0.0f;
      st->sa[i] = make_float4(0.2f, 0.1f, 0.05f, 1.0f);
      // stored times the basis constants, like decoded coefficients
      for (int k = 0; k < 48; ++k) st->sh[48 * i + k] = kShConstants[k / 3] * 0.0f;
      st->sh[48 * i + 0] = kShConstants[0] * (i == 0 ? 1.0f : 0.0f);
      st->sh[48 * i + 1] = kShConstants[0] * (i == 1 ? 1.0f : 0.0f);
      st->sh[48 * i + 2] = kShConstants[0] * (i == 2 ? 1.0f : 0.0f);
    }
    rc = up.send(*st, 0, (size_t)std::min<uint64_t>(3, s->count));
  }
  if (!rc) rc = up.finish();
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
  const int fd = open(path, O_RDONLY | O_CLOEXEC);
  if (fd < 0) return fail(SPL_ERROR_IO, std::string("cannot open ") + path);
  int rc = SPL_OK;
  {
    Uploader up(s, (size_t)std::min<uint64_t>(end - begin, Uploader::kChunk));
    std::vector<SerializedSplat> buf(up.capacity());
    for (uint64_t at = begin; at < end && !rc; at += up.capacity()) {
      const size_t n = (size_t)std::min<uint64_t>(up.capacity(), end - at);
      rc = up.push_file(at, fd, (uint64_t)header_size + at * sizeof(SerializedSplat), buf.data(), n);
    }
    if (!rc) rc = up.finish();
  }
  close(fd);
  return rc;
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
  spl::device_retain(d);
  *out = f;
  return SPL_OK;
}

int spl_frame_destroy(spl_frame* f) {
  if (!f) return SPL_OK;
  (void)hipSetDevice(f->dev->ordinal);
  (void)hipStreamSynchronize(f->dev->stream);
  dev_free(f->pixels);
  spl_device* d = f->dev;
  delete f;
  spl::device_release(d);
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

}  // extern "C"
