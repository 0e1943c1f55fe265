// Allocation-free closest-hit ray queries: rtpg_trace_closest on caller-owned
// device buffers.  The traversal's spill stack and its stack-bound
// flag live in `scratch`, a fixed device array of this file; nothing is
// allocated per call (no hipMalloc/hipFree, no staging copies of the rays).
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/rtpg.h"
#include "context.h"

namespace rtpg {
namespace {

// scratch layout (32-bit words): [0, kFlagWords) the stack-bound flag (word 0,
// the rest pads the stacks to a 256-byte boundary); then the spill stacks of
// one launch: 2 x (stack bound - kShortStack) words per ray.
constexpr size_t kScratchBytes = 32u << 20;
constexpr size_t kScratchWords = kScratchBytes / 4u;
constexpr size_t kFlagWords = 64;
constexpr uint32_t kMinChunk = 256;  // a launch of fewer rays than this is not worth it
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint32_t scratch[kScratchWords];

// Device address of `scratch` per device (each device holds its own copy),
// from hipGetSymbolAddress with that device current; null until set up.
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};

template <class F>
int guarded(F&& f) {
  try {
    f();
    return RTPG_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return RTPG_ERR_INVALID;
  } catch (...) {
    set_last_error("unknown error");
    return RTPG_ERR_INVALID;
  }
}

void require(bool c, const char* msg) {
  if (!c) throw Error(RTPG_ERR_INVALID, msg);
}

__global__ void k_clear_flag(uint32_t* flag) {
  if (threadIdx.x == 0) flag[0] = 0u;
}

uint32_t* scratch_of(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  return (device >= 0 && device < kMaxDevices) ? g_scratch[device] : nullptr;
}

int trace_closest_noalloc(rtpg_context ctx, const float* rays, uint32_t n, float* hits) {
  return guarded([&] {
    require(ctx && rays && hits, "NULL argument");
    require(ctx->n_tris > 0, "no scene uploaded");
    if (n == 0) return;
    RTPG_HIP(hipSetDevice(ctx->device));
    uint32_t* base = scratch_of(ctx->device);
    require(base != nullptr, "scratch not set up on this device");
    // each ray needs 2 words per stack entry beyond the LDS short stack
    const size_t per_ray = ctx->max_stack > kShortStack ? 2u * (size_t)(ctx->max_stack - kShortStack) : 0u;
    const size_t room = kScratchWords - kFlagWords;
    const size_t fit = per_ray ? room / per_ray : (size_t)n;
    if (fit < kMinChunk && fit < n) {
      throw Error(RTPG_ERR_LIMIT, "scene's traversal stack bound (" + std::to_string(ctx->max_stack) +
                                      " entries) needs more than the " + std::to_string(kScratchBytes >> 20) +
                                      " MiB scratch for a launch of " + std::to_string(kMinChunk) + " rays");
    }
    const uint32_t chunk = (uint32_t)std::min<size_t>(fit, (size_t)n);
    hipStream_t s = ctx->stream;
    uint32_t* flag = base;
    uint32_t* deep = per_ray ? base + kFlagWords : nullptr;
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag);
    RTPG_HIP(hipGetLastError());
    SceneView sc = ctx->scene_view();
    // launches on one stream run in order, so consecutive chunks reuse the
    // same spill stacks
    for (uint32_t off = 0; off < n; off += chunk) {
      uint32_t m = std::min(chunk, n - off);
      const float4* r = reinterpret_cast<const float4*>(rays) + 2u * (size_t)off;
      float4* h = reinterpret_cast<float4*>(hits) + off;
      RTPG_HIP(launch_trace_rays(sc, r, m, h, nullptr, false, deep, flag, s));
    }
    uint32_t err = 0;
    RTPG_HIP(hipMemcpyAsync(&err, flag, sizeof(err), hipMemcpyDeviceToHost, s));
    RTPG_HIP(hipStreamSynchronize(s));
    if (err) throw Error(RTPG_ERR_LIMIT, "traversal stack bound exceeded");
  });
}

}  // namespace

// Called by rtpg_context_create with the device current: records the device
// address of this device's `scratch` once.
void scratch_setup(int device) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (device < 0 || device >= kMaxDevices) throw Error(RTPG_ERR_INVALID, "device index beyond the scratch table");
  if (g_scratch[device]) return;
  void* p = nullptr;
  hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
  g_scratch[device] = static_cast<uint32_t*>(p);
}

}  // namespace rtpg

extern "C" {

int rtpg_trace_closest_noalloc(rtpg_context ctx, const float* rays, uint32_t n, float* hits) {
  return rtpg::trace_closest_noalloc(ctx, rays, n, hits);
}

}  // extern "C"
