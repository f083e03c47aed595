// Failure injection for tests (bm3d_debug_inject_failure): one process-wide
// one-shot countdown per kind of runtime event. While a countdown is armed,
// each event of its kind decrements it, and the event that finds it at zero
// fails as if the runtime had refused it, without calling the runtime (a
// failed launch launches nothing). Disarmed, an event costs one atomic load.
// Internal.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace bm3d {

enum InjectSite { kInjectDeviceAlloc = 1, kInjectHostAlloc = 2, kInjectLaunch = 3, kInjectSites = 4 };

// True when this event of kind `site` must fail.
bool inject_now(int site);

// hipMalloc / hipHostMalloc with the injection check.
inline hipError_t malloc_device(void** p, size_t bytes) {
    if (inject_now(kInjectDeviceAlloc)) return hipErrorOutOfMemory;
    return hipMalloc(p, bytes);
}
template <typename T>
inline hipError_t malloc_device(T** p, size_t bytes) {
    return malloc_device(reinterpret_cast<void**>(p), bytes);
}
template <typename T>
inline hipError_t malloc_pinned(T** p, size_t bytes) {
    if (inject_now(kInjectHostAlloc)) return hipErrorOutOfMemory;
    return hipHostMalloc(reinterpret_cast<void**>(p), bytes, 0);
}

}  // namespace bm3d

// First statement of every kernel launcher (functions returning hipError_t).
#define BM3D_INJECT_LAUNCH()                                                  \
    do {                                                                      \
        if (::bm3d::inject_now(::bm3d::kInjectLaunch)) return hipErrorLaunchFailure; \
    } while (0)
