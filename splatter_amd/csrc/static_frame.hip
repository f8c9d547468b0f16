// static_frame.hip -- spl_render_frame_static: one frame of the pipeline
// with its whole workspace in a device array of this library, for callers
// that cannot allocate device memory (code under stream capture, tight
// serving loops).  No counterpart in the reference.
//
// `scratch` is a file-scope device array of kStaticScratchBytes.  Each GPU
// the library runs on holds its own copy at its own address, looked up once
// per device when the device opens (hipGetSymbolAddress with that device
// current) and kept in spl_device::static_scratch; host code uses only that
// address, never the array's name, which on the host is a stand-in and not
// a device address.  Every call on one GPU shares the array: a call zeroes
// the bookkeeping it needs first, and calls that overlap on one GPU race.
#include <mutex>

#include "spl_device.h"
#include "spl_internal.h"
#include "spl_objects.h"

namespace {

static __device__ __align__(256) uint8_t scratch[spl::kStaticScratchBytes];

constexpr int kMaxOrdinals = 64;
std::mutex g_scratch_mu;
void* g_scratch_at[kMaxOrdinals] = {};

__global__ __launch_bounds__(256) void k_static_setup(spl::Uniforms u, uint32_t* pixels, spl::Uniforms* d_uniforms,
                                                      uint32_t** d_pixels, uint32_t* zero, size_t zero_words) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *d_uniforms = u;
    *d_pixels = pixels;
  }
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < zero_words; i += stride) zero[i] = 0u;
}

}  // namespace

namespace spl {

void launch_static_setup(const Uniforms& u, uint32_t* pixels, Uniforms* d_uniforms, uint32_t** d_pixels,
                         uint32_t* zero, size_t zero_words, hipStream_t s) {
  const size_t blocks = std::min<size_t>(std::max<size_t>((zero_words + 255) / 256, 1), 1024);
  hipLaunchKernelGGL(k_static_setup, dim3((uint32_t)blocks), dim3(256), 0, s, u, pixels, d_uniforms, d_pixels,
                     zero, zero_words);
}

int static_scratch_setup(spl_device* dev) {
  if (dev->ordinal < 0 || dev->ordinal >= kMaxOrdinals)
    return fail(SPL_ERROR_UNSUPPORTED, "static scratch: device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_at[dev->ordinal]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(SPL_ERROR_DEVICE, std::string("static scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
    g_scratch_at[dev->ordinal] = p;
  }
  dev->static_scratch = g_scratch_at[dev->ordinal];
  return SPL_OK;
}

}  // namespace spl

extern "C" int spl_render_frame_static(spl_device* dev, const spl_configuration* cfg, spl_scene* scene,
                                       spl_frame* frame, const float camera_matrix[16],
                                       const float view_matrix[16]) {
  using namespace spl;
  if (!dev || !cfg || !scene || !frame || !camera_matrix || !view_matrix)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_frame_static: null argument");
  if (scene->dev != dev || frame->dev != dev)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_frame_static: scene, frame and device must match");
  if (cfg->max_views > 1)
    return fail(SPL_ERROR_INVALID_ARGUMENT, "spl_render_frame_static: renders one view (max_views 0 or 1)");
  if (scene->count > kStaticMaxSplats)
    return fail(SPL_ERROR_CAPACITY, "spl_render_frame_static: at most 131072 splats");
  if (frame->width > kStaticMaxWidth || frame->height > kStaticMaxHeight)
    return fail(SPL_ERROR_CAPACITY, "spl_render_frame_static: frames up to 3840 x 2160");
  if (!dev->static_scratch) return fail(SPL_ERROR_DEVICE, "spl_render_frame_static: no scratch on this device");
  HIP_TRY(hipSetDevice(dev->ordinal));
  return render_frame_scratch(dev, cfg, scene, frame, camera_matrix, view_matrix,
                              static_cast<uint8_t*>(dev->static_scratch), kStaticScratchBytes);
}
