// server.cpp -- the serving form: resident scenes, a request queue admitted by
// a device-memory budget, and continuous batching over several streams.
//
// No counterpart in the reference (its showcase renders one frame at a time,
// showcase.rs:101-143).  A request is one view of one resident scene at its
// own resolution.  The scheduler keeps up to max_in_flight batches on the
// device, each on a slot of its own (a renderer with its own stream), so
// their kernels overlap.  Whenever a slot's batch is done, the oldest pending
// request and every later pending request of the same scene (up to the batch
// width) go out together as one spl_render_views batch: requests that arrive
// while others run join the next batch of their scene.  Each batch needs a
// slot whose workspace fits its scene (slots come in capacity classes) and a
// frame per request; slots and frames are pooled and counted against the
// memory budget, idle ones are released least recently used first when a new
// one does not fit, and a request whose batch still does not fit waits (in
// arrival order) for running batches to finish.  A batch renders exactly what
// spl_render_views renders for the same views, so a request's image does not
// depend on what it was batched or overlapped with.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "spl_internal.h"
#include "spl_objects.h"
#include "splatter.h"

using namespace spl;

namespace {

struct Slot {
  spl_renderer* r = nullptr;
  uint64_t cap = 0;       // capacity class (splats)
  uint64_t bytes = 0;
  bool busy = false;
  uint64_t used = 0;      // LRU clock
  hipEvent_t done = nullptr;
  hipEvent_t ready = nullptr;  // device-stream work (scene uploads) before the batch
  uint32_t* flags = nullptr;   // pinned copy of the batch's counter blocks
  std::vector<uint64_t> tickets;
};

struct FrameBuf {
  spl_frame* f = nullptr;
  uint32_t w = 0, h = 0;
  bool in_use = false;
  uint64_t used = 0;
};

struct Request {
  spl_scene* scene = nullptr;
  uint32_t w = 0, h = 0;
  float cam[16], view[16];
  int status = SPL_SERVER_PENDING;
  int frame = -1;
  int slot = -1;
  uint32_t batch_views = 0;
  double t_submit = 0, t_start = 0, t_done = 0;
};

double now_ms(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Capacity class of a scene: the smallest 2^k or 3 * 2^(k-1) at or above its
// splat count (at most 50% headroom), at least 2^16.
uint64_t capacity_class(uint64_t n, uint64_t max_cap) {
  uint64_t c = 1ull << 16;
  while (c < n) {
    const uint64_t mid = c + c / 2;
    if (mid >= n) {
      c = mid;
      break;
    }
    c <<= 1;
  }
  return std::min(c, max_cap);
}

}  // namespace

struct spl_server {
  spl_device* dev = nullptr;
  spl_server_configuration cfg{};
  uint32_t presize_tiles = 0;
  std::chrono::steady_clock::time_point t0;
  std::vector<Slot> slots;
  std::vector<FrameBuf> frames;
  std::unordered_map<uint64_t, Request> requests;
  std::deque<uint64_t> pending;
  std::vector<uint64_t> finished;  // since the last poll
  uint64_t next_ticket = 1;
  uint64_t clock = 0;
  uint64_t mem_used = 0, mem_peak = 0;
  spl_server_stats stats{};
};

namespace {

// Slots and frames keep their index for life; a released entry is empty
// (null renderer / frame) until a new one takes its place.
void release_slot(spl_server* s, size_t i) {
  Slot& sl = s->slots[i];
  if (!sl.r) return;
  spl_renderer_destroy(sl.r);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.ready) (void)hipEventDestroy(sl.ready);
  if (sl.flags) (void)hipHostFree(sl.flags);
  s->mem_used -= sl.bytes;
  sl = Slot{};
}

void release_frame(spl_server* s, size_t i) {
  FrameBuf& fb = s->frames[i];
  if (!fb.f) return;
  s->mem_used -= (uint64_t)fb.w * fb.h * 4;
  spl_frame_destroy(fb.f);
  fb = FrameBuf{};
}

// Frees idle frames, then idle slots, least recently used first, until
// `need` more bytes fit the budget; slot `keep_slot` (if any) stays.
bool make_room(spl_server* s, uint64_t need, int keep_slot) {
  while (s->mem_used + need > s->cfg.memory_budget_bytes) {
    int best_f = -1, best_s = -1;
    for (size_t i = 0; i < s->frames.size(); ++i)
      if (s->frames[i].f && !s->frames[i].in_use && (best_f < 0 || s->frames[i].used < s->frames[best_f].used))
        best_f = (int)i;
    if (best_f >= 0) {
      release_frame(s, best_f);
      continue;
    }
    for (size_t i = 0; i < s->slots.size(); ++i)
      if (s->slots[i].r && !s->slots[i].busy && (int)i != keep_slot &&
          (best_s < 0 || s->slots[i].used < s->slots[best_s].used))
        best_s = (int)i;
    if (best_s < 0) return false;
    release_slot(s, best_s);
  }
  return true;
}

int busy_slots(const spl_server* s) {
  int n = 0;
  for (const auto& sl : s->slots) n += sl.busy ? 1 : 0;
  return n;
}

// Completion of finished batches: their requests become DONE (or fail with
// the pipeline's capacity / consistency flags).
int reap(spl_server* s) {
  for (auto& sl : s->slots) {
    if (!sl.busy) continue;
    const hipError_t q = hipEventQuery(sl.done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) return fail(SPL_ERROR_DEVICE, std::string("spl_server: batch failed: ") + hipGetErrorString(q));
    const double t = now_ms(s->t0);
    for (size_t v = 0; v < sl.tickets.size(); ++v) {
      Request& rq = s->requests[sl.tickets[v]];
      const uint32_t* fl = sl.flags + v * kCounterWords;
      if (fl[kScatterError]) rq.status = SPL_ERROR_DEVICE;
      else if (fl[kPairOverflow]) rq.status = SPL_ERROR_CAPACITY;
      else rq.status = SPL_SERVER_DONE;
      rq.t_done = t;
      rq.slot = -1;
      if (rq.status == SPL_SERVER_DONE) {
        ++s->stats.completed;
      } else {
        ++s->stats.failed;
        s->frames[rq.frame].in_use = false;  // a failed request keeps no frame
        rq.frame = -1;
      }
      s->finished.push_back(sl.tickets[v]);
    }
    sl.tickets.clear();
    sl.busy = false;
  }
  return SPL_OK;
}

int frame_for(spl_server* s, uint32_t w, uint32_t h, int keep_slot, int* out) {
  for (size_t i = 0; i < s->frames.size(); ++i) {
    FrameBuf& fb = s->frames[i];
    if (fb.f && !fb.in_use && fb.w == w && fb.h == h) {
      fb.in_use = true;
      fb.used = ++s->clock;
      *out = (int)i;
      return SPL_OK;
    }
  }
  const uint64_t need = (uint64_t)w * h * 4;
  if (!make_room(s, need, keep_slot)) return SPL_ERROR_OUT_OF_MEMORY;
  FrameBuf fb;
  fb.w = w;
  fb.h = h;
  const int rc = spl_frame_create(s->dev, w, h, &fb.f);
  if (rc) return rc;
  fb.in_use = true;
  fb.used = ++s->clock;
  size_t i = 0;
  while (i < s->frames.size() && s->frames[i].f) ++i;
  if (i == s->frames.size()) s->frames.push_back(fb);
  else s->frames[i] = fb;
  s->mem_used += need;
  s->mem_peak = std::max(s->mem_peak, s->mem_used);
  *out = (int)i;
  return SPL_OK;
}

// A slot of class `cap` that is idle, or a new one if the budget allows.
int slot_for(spl_server* s, uint64_t cap, int* out) {
  for (size_t i = 0; i < s->slots.size(); ++i)
    if (s->slots[i].r && !s->slots[i].busy && s->slots[i].cap == cap) {
      *out = (int)i;
      return SPL_OK;
    }
  spl_configuration c = s->cfg.renderer;
  c.max_splat_count = cap;
  uint64_t bytes = 0;
  int rc = renderer_create(s->dev, &c, true, s->presize_tiles, nullptr, &bytes);
  if (rc) return rc;
  if (!make_room(s, bytes, -1)) return SPL_ERROR_OUT_OF_MEMORY;
  Slot sl;
  rc = renderer_create(s->dev, &c, true, s->presize_tiles, &sl.r, nullptr);
  if (rc) return rc;
  sl.cap = cap;
  sl.bytes = renderer_bytes(sl.r);
  hipError_t e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.ready, hipEventDisableTiming);
  if (e == hipSuccess)
    e = hipHostMalloc((void**)&sl.flags, (size_t)c.max_views * kCounterWords * 4, hipHostMallocDefault);
  if (e != hipSuccess) {
    spl_renderer_destroy(sl.r);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.ready) (void)hipEventDestroy(sl.ready);
    return fail(SPL_ERROR_DEVICE, std::string("spl_server: slot: ") + hipGetErrorString(e));
  }
  size_t i = 0;
  while (i < s->slots.size() && s->slots[i].r) ++i;
  if (i == s->slots.size()) s->slots.push_back(sl);
  else s->slots[i] = sl;
  s->mem_used += sl.bytes;
  s->mem_peak = std::max(s->mem_peak, s->mem_used);
  *out = (int)i;
  return SPL_OK;
}

// Admission: batches from the head of the queue while slots and memory last.
int admit(spl_server* s) {
  while (!s->pending.empty() && busy_slots(s) < (int)s->cfg.max_in_flight) {
    const uint64_t head = s->pending.front();
    spl_scene* scene = s->requests[head].scene;
    const uint64_t cap = capacity_class(scene->count, s->cfg.renderer.max_splat_count);
    int si = -1;
    int rc = slot_for(s, cap, &si);
    if (rc == SPL_ERROR_OUT_OF_MEMORY) return SPL_OK;  // wait for running batches
    if (rc) return rc;
    // the head and later requests of the same scene, in arrival order
    std::vector<uint64_t> batch;
    for (uint64_t t : s->pending) {
      if (batch.size() == s->cfg.renderer.max_views) break;
      if (s->requests[t].scene == scene) batch.push_back(t);
    }
    std::vector<int> fidx;
    for (uint64_t t : batch) {
      const Request& rq = s->requests[t];
      int fi = -1;
      rc = frame_for(s, rq.w, rq.h, si, &fi);
      if (rc == SPL_ERROR_OUT_OF_MEMORY) break;
      if (rc) return rc;
      fidx.push_back(fi);
    }
    if (fidx.empty()) return SPL_OK;  // not even the head's frame fits yet
    batch.resize(fidx.size());
    Slot& sl = s->slots[si];
    const uint32_t nv = (uint32_t)batch.size();
    std::vector<spl_frame*> fr(nv);
    std::vector<float> cams(16 * (size_t)nv), views(16 * (size_t)nv);
    const double t = now_ms(s->t0);
    for (uint32_t v = 0; v < nv; ++v) {
      Request& rq = s->requests[batch[v]];
      rq.frame = fidx[v];
      rq.slot = si;
      rq.status = SPL_SERVER_RUNNING;
      rq.batch_views = nv;
      rq.t_start = t;
      fr[v] = s->frames[fidx[v]].f;
      std::memcpy(&cams[16 * (size_t)v], rq.cam, 64);
      std::memcpy(&views[16 * (size_t)v], rq.view, 64);
    }
    const hipStream_t q = renderer_stream(sl.r);
    HIP_TRY(hipEventRecord(sl.ready, s->dev->stream));
    HIP_TRY(hipStreamWaitEvent(q, sl.ready, 0));
    rc = spl_render_views(sl.r, scene, nv, fr.data(), cams.data(), views.data());
    if (rc) return rc;
    const uint32_t* ctr = nullptr;
    rc = renderer_counter_words(sl.r, 0, &ctr);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(sl.flags, ctr, (size_t)nv * kCounterWords * 4, hipMemcpyDeviceToHost, q));
    HIP_TRY(hipEventRecord(sl.done, q));
    sl.busy = true;
    sl.used = ++s->clock;
    sl.tickets = batch;
    ++s->stats.batches;
    s->stats.batched_views += nv;
    for (uint64_t tk : batch) s->pending.erase(std::find(s->pending.begin(), s->pending.end(), tk));
  }
  return SPL_OK;
}

int pump(spl_server* s) {
  HIP_TRY(hipSetDevice(s->dev->ordinal));
  int rc = reap(s);
  if (rc) return rc;
  return admit(s);
}

Request* find(spl_server* s, uint64_t ticket) {
  auto it = s->requests.find(ticket);
  return it == s->requests.end() ? nullptr : &it->second;
}

}  // namespace

extern "C" {

void spl_server_configuration_default(spl_server_configuration* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  spl_configuration_default(&c->renderer);
  c->renderer.max_splat_count = 1u << 23;
  c->renderer.max_views = 8;
  c->renderer.tile_pairs_per_splat = 8;
  c->max_in_flight = 3;
  c->max_width = 3840;
  c->max_height = 2160;
  c->memory_budget_bytes = 64ull << 30;
}

int spl_server_create(spl_device* dev, const spl_server_configuration* cfg, spl_server** out) {
  if (!dev || !cfg || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_create: null argument");
  *out = nullptr;
  if (cfg->max_in_flight < 1 || cfg->max_in_flight > 8)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "max_in_flight must be in 1..=8");
  if (cfg->max_width == 0 || cfg->max_height == 0 || cfg->max_width > 16384 || cfg->max_height > 16384)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "max_width / max_height must be in 1..=16384");
  if (cfg->renderer.max_views == 0) return fail(SPL_ERROR_INVALID_ARGUMENT, "renderer.max_views must be at least 1");
  // validates the pipeline options once, without allocating
  uint64_t probe = 0;
  spl_configuration c = cfg->renderer;
  c.max_splat_count = std::max<uint64_t>(1, std::min<uint64_t>(c.max_splat_count, 1ull << 16));
  int rc = renderer_create(dev, &c, true, 0, nullptr, &probe);
  if (rc) return rc;
  auto* s = new (std::nothrow) spl_server();
  if (!s) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_server_create: host allocation");
  s->dev = dev;
  device_retain(dev);
  s->cfg = *cfg;
  s->presize_tiles = div_up(cfg->max_width, kTileSize) * div_up(cfg->max_height, kTileSize);
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return SPL_OK;
}

int spl_server_destroy(spl_server* s) {
  if (!s) return SPL_OK;
  (void)hipSetDevice(s->dev->ordinal);
  for (auto& sl : s->slots)
    if (sl.busy) (void)hipEventSynchronize(sl.done);
  for (size_t i = 0; i < s->slots.size(); ++i) release_slot(s, i);
  for (size_t i = 0; i < s->frames.size(); ++i) release_frame(s, i);
  spl_device* d = s->dev;
  delete s;
  spl::device_release(d);
  return SPL_OK;
}

int spl_server_scene_create(spl_server* s, uint64_t splat_count, spl_scene** out) {
  if (!s || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_scene_create: null argument");
  if (splat_count > s->cfg.renderer.max_splat_count)
    return fail(SPL_ERROR_CAPACITY, "spl_server_scene_create: more splats than renderer.max_splat_count");
  return scene_create(s->dev, splat_count, out);
}

int spl_server_submit(spl_server* s, spl_scene* scene, uint32_t width, uint32_t height,
                      const float camera_matrix[16], const float view_matrix[16], uint64_t* ticket) {
  if (!s || !scene || !camera_matrix || !view_matrix || !ticket)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_submit: null argument");
  if (scene->dev != s->dev) return fail(SPL_ERROR_INVALID_ARGUMENT, "scene and server must share a device");
  if (scene->count > s->cfg.renderer.max_splat_count)
    return fail(SPL_ERROR_CAPACITY, "scene has more splats than renderer.max_splat_count");
  if (width == 0 || height == 0 || width > s->cfg.max_width || height > s->cfg.max_height)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "request size outside 1..=max_width x 1..=max_height");
  Request rq;
  rq.scene = scene;
  rq.w = width;
  rq.h = height;
  std::memcpy(rq.cam, camera_matrix, 64);
  std::memcpy(rq.view, view_matrix, 64);
  rq.t_submit = now_ms(s->t0);
  const uint64_t t = s->next_ticket++;
  s->requests[t] = rq;
  s->pending.push_back(t);
  ++s->stats.submitted;
  *ticket = t;
  return pump(s);
}

int spl_server_poll(spl_server* s, uint64_t* finished, uint32_t capacity, uint32_t* count) {
  if (!s || !count || (capacity && !finished)) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_poll: null argument");
  const int rc = pump(s);
  const uint32_t n = (uint32_t)std::min<size_t>(capacity, s->finished.size());
  for (uint32_t i = 0; i < n; ++i) finished[i] = s->finished[i];
  s->finished.erase(s->finished.begin(), s->finished.begin() + n);
  *count = n;
  return rc;
}

int spl_server_wait(spl_server* s, uint64_t ticket) {
  if (!s) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_wait: null argument");
  while (true) {
    int rc = pump(s);
    if (rc) return rc;
    Request* rq = find(s, ticket);
    if (!rq) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_wait: unknown ticket");
    if (rq->status != SPL_SERVER_PENDING && rq->status != SPL_SERVER_RUNNING) return SPL_OK;
    // block on the request's batch, or on the oldest running batch
    const Slot* target = rq->slot >= 0 ? &s->slots[rq->slot] : nullptr;
    if (!target)
      for (const auto& sl : s->slots)
        if (sl.busy && (!target || sl.used < target->used)) target = &sl;
    if (!target) return fail(SPL_ERROR_OUT_OF_MEMORY, "spl_server_wait: request does not fit the memory budget");
    HIP_TRY(hipEventSynchronize(target->done));
  }
}

int spl_server_request_info(spl_server* s, uint64_t ticket, spl_request_info* info) {
  if (!s || !info) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_request_info: null argument");
  const Request* rq = find(s, ticket);
  if (!rq) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_request_info: unknown ticket");
  info->status = rq->status;
  info->width = rq->w;
  info->height = rq->h;
  info->batch_views = rq->batch_views;
  info->submit_ms = rq->t_submit;
  info->start_ms = rq->t_start;
  info->done_ms = rq->t_done;
  return SPL_OK;
}

int spl_server_read(spl_server* s, uint64_t ticket, void* host, uint64_t size) {
  if (!s || !host) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_read: null argument");
  int rc = spl_server_wait(s, ticket);
  if (rc) return rc;
  const Request* rq = find(s, ticket);
  if (rq->status != SPL_SERVER_DONE) return fail(rq->status, "spl_server_read: the request failed");
  const uint64_t bytes = (uint64_t)rq->w * rq->h * 4;
  if (size < bytes) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_read: buffer smaller than width*height*4");
  HIP_TRY(hipMemcpy(host, s->frames[rq->frame].f->pixels, bytes, hipMemcpyDeviceToHost));
  return SPL_OK;
}

int spl_server_release(spl_server* s, uint64_t ticket) {
  if (!s) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_release: null argument");
  Request* rq = find(s, ticket);
  if (!rq) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_release: unknown ticket");
  if (rq->status == SPL_SERVER_PENDING || rq->status == SPL_SERVER_RUNNING)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_release: request not finished");
  if (rq->frame >= 0) s->frames[rq->frame].in_use = false;
  s->requests.erase(ticket);
  return pump(s);
}

int spl_server_get_stats(spl_server* s, spl_server_stats* out) {
  if (!s || !out) return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_server_get_stats: null argument");
  *out = s->stats;
  out->pending = s->pending.size();
  out->slots = 0;
  for (const auto& sl : s->slots) out->slots += sl.r ? 1u : 0u;
  out->slots_busy = (uint32_t)busy_slots(s);
  out->frames = 0;
  for (const auto& fb : s->frames) out->frames += fb.f ? 1u : 0u;
  out->memory_in_use = s->mem_used;
  out->memory_peak = s->mem_peak;
  out->memory_budget = s->cfg.memory_budget_bytes;
  return SPL_OK;
}

}  // extern "C"
