// spl_objects.h -- the objects behind the C ABI's opaque handles and the
// error plumbing shared by api.cpp and renderer.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <string>

#include "splatter.h"

namespace spl {

// Sets the calling thread's message (spl_last_error) and returns `code`.
int fail(int code, const std::string& msg);

// Renderer accessors for the scene functions (Scene::new clamps to the
// renderer's max_splat_count, scene.rs:46).
spl_device* renderer_device(const spl_renderer* r);
uint64_t renderer_capacity(const spl_renderer* r);

// spl_renderer_create with a stream of the renderer's own (instead of the
// device's) and the tile-range buffer sized up front for `presize_tiles`
// screen tiles, so no later frame up to that size allocates.  With
// dry_bytes set nothing is allocated: the device bytes such a renderer would
// hold are written there instead.
int renderer_create(spl_device* dev, const spl_configuration* cfg, bool own_stream, uint32_t presize_tiles,
                    spl_renderer** out, uint64_t* dry_bytes);
hipStream_t renderer_stream(const spl_renderer* r);
uint64_t renderer_bytes(const spl_renderer* r);  // device memory the renderer holds
// Device address of view `view`'s counter block (kCounterWords u32).
int renderer_counter_words(const spl_renderer* r, uint32_t view, const uint32_t** words);

// Scene::new without a renderer: `count` splats, zeroed, on `dev`.
int scene_create(spl_device* dev, uint64_t count, spl_scene** out);

// One frame of spl_render_frame's pipeline with every workspace buffer carved
// from `scratch` (`bytes` of device memory): no allocation.  The (splat,
// tile) pairs get what the splats and the frame leave of it.
int render_frame_scratch(spl_device* dev, const spl_configuration* cfg, spl_scene* s, spl_frame* f,
                         const float cam[16], const float view[16], uint8_t* scratch, uint64_t bytes);
// Looks up this library's scratch array on device `dev` (current) once and
// keeps its address in dev->static_scratch.
int static_scratch_setup(spl_device* dev);

// Every renderer, scene, frame and server holds a reference to its device,
// as does the handle spl_device_open returns: the device (and its stream)
// goes with the last of them, whatever order they are destroyed in.
void device_retain(spl_device* d);
void device_release(spl_device* d);

inline uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

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

}  // namespace spl

#define HIP_TRY(expr)                                                                                   \
  do {                                                                                                  \
    hipError_t e_ = (expr);                                                                             \
    if (e_ != hipSuccess) return spl::fail(SPL_ERROR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Streams a device hands to its renderers (own streams and the projection's
// side streams), all made when the device opens, right after its own: two
// batches in flight overlapped fully only on streams made first in the
// process (1540 against about 1390 frames/s on MI355X once other streams
// had been made and freed before them).  A renderer takes the least-used
// pool streams and gives them back when it is destroyed.
constexpr int kPoolStreams = 8;

struct spl_device {
  int ordinal = 0;
  hipStream_t stream = nullptr;
  int cu_count = 0;
  std::atomic<int> refs{1};  // the opener's handle + one per object created on the device
  hipStream_t pool[kPoolStreams] = {};
  void* static_scratch = nullptr;  // this device's copy of the library's scratch array (spl_render_frame_static)
  int pool_users[kPoolStreams] = {};
  std::mutex pool_mu;
};

namespace spl {
hipStream_t acquire_stream(spl_device* d);
void release_stream(spl_device* d, hipStream_t s);
}  // namespace spl

struct spl_scene {
  spl_device* dev = nullptr;
  uint64_t uid = 0;  // unique for the life of the process: names the scene in cached launch sequences
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
