// Device runtime behind the C-ABI: device selection, the library stream, memory, events.
// One process drives one GPU; all kernels run on one non-blocking stream created at the
// default priority.  The role matches the reference's cuda.c memory manager
// (src/cuda.c:296-466, :1168-1200), but grids are device-canonical: the host library
// tracks which copy is current and there is no swapping and no CPU fallback.

#include <cstdio>
#include <cstring>
#include <mutex>

#include "grid_hip.h"
#include "rt.hpp"

namespace lg {
static hipStream_t g_stream = nullptr;
static int g_device = -1;
static char g_err[512] = "";
static char g_arch[64] = "";
static std::mutex g_mu;

hipStream_t stream() { return g_stream; }

int check(hipError_t e, const char *what) {
  if (e == hipSuccess) return 0;
  std::snprintf(g_err, sizeof(g_err), "%s: %s", what, hipGetErrorString(e));
  return -1;
}

int check_launch(const char *what) { return check(hipGetLastError(), what); }

int fail(const char *what) {
  std::snprintf(g_err, sizeof(g_err), "%s", what);
  return -1;
}
}  // namespace lg

using lg::check;

extern "C" int hgrid_init(int device) {
  std::lock_guard<std::mutex> lk(lg::g_mu);
  if (lg::g_stream) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    std::snprintf(lg::g_err, sizeof(lg::g_err), "no HIP device present");
    return -1;
  }
  if (device < 0 || device >= n) device = 0;
  if (check(hipSetDevice(device), "hipSetDevice")) return -1;
  hipDeviceProp_t prop;
  if (check(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties")) return -1;
  std::snprintf(lg::g_arch, sizeof(lg::g_arch), "%s", prop.gcnArchName);
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    std::snprintf(lg::g_err, sizeof(lg::g_err), "device %d is %s; libgrid is built for gfx950 only", device,
                  prop.gcnArchName);
    return -2;
  }
  if (check(hipStreamCreateWithFlags(&lg::g_stream, hipStreamNonBlocking), "hipStreamCreate")) return -1;
  lg::g_device = device;
  // the device's fixed scratch of the allocation-free entry point
  if (lg::noalloc_setup(device)) return -1;
  return 0;
}

extern "C" int hgrid_initialized(void) { return lg::g_stream != nullptr; }
extern "C" const char *hgrid_last_error(void) { return lg::g_err; }
extern "C" const char *hgrid_device_name(void) { return lg::g_arch; }

extern "C" int hgrid_malloc(void **ptr, size_t bytes) {
  *ptr = nullptr;
  if (bytes == 0) bytes = 256;
  return check(hipMalloc(ptr, bytes), "hipMalloc");
}

extern "C" int hgrid_free(void *ptr) {
  if (!ptr) return 0;
  // the stream may still use the memory
  if (check(hipStreamSynchronize(lg::g_stream), "hipStreamSynchronize")) return -1;
  return check(hipFree(ptr), "hipFree");
}

extern "C" int hgrid_memcpy_h2d(void *dst, const void *src, size_t bytes) {
  if (check(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, lg::g_stream), "hipMemcpy H2D")) return -1;
  return check(hipStreamSynchronize(lg::g_stream), "hipStreamSynchronize");
}

extern "C" int hgrid_memcpy_d2h(void *dst, const void *src, size_t bytes) {
  if (check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, lg::g_stream), "hipMemcpy D2H")) return -1;
  return check(hipStreamSynchronize(lg::g_stream), "hipStreamSynchronize");
}

extern "C" int hgrid_memcpy_d2d(void *dst, const void *src, size_t bytes) {
  return check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, lg::g_stream), "hipMemcpy D2D");
}

extern "C" int hgrid_sync(void) { return check(hipStreamSynchronize(lg::g_stream), "hipStreamSynchronize"); }

// HIP graphs of the library stream: capture a sequence of library calls once, replay it.
extern "C" int hgrid_graph_begin(void) {
  return check(hipStreamBeginCapture(lg::g_stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
}

extern "C" int hgrid_graph_end(void **exec) {
  hipGraph_t graph = nullptr;
  if (check(hipStreamEndCapture(lg::g_stream, &graph), "hipStreamEndCapture")) return -1;
  hipGraphExec_t ge = nullptr;
  const int rc = check(hipGraphInstantiate(&ge, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
  (void)hipGraphDestroy(graph);
  if (rc) return -1;
  *exec = (void *)ge;
  return 0;
}

extern "C" int hgrid_graph_launch(void *exec) {
  return check(hipGraphLaunch((hipGraphExec_t)exec, lg::g_stream), "hipGraphLaunch");
}

extern "C" int hgrid_graph_free(void *exec) {
  return exec ? check(hipGraphExecDestroy((hipGraphExec_t)exec), "hipGraphExecDestroy") : 0;
}

extern "C" int hgrid_event_create(void **ev) {
  hipEvent_t e;
  if (check(hipEventCreate(&e), "hipEventCreate")) return -1;
  *ev = (void *)e;
  return 0;
}
extern "C" int hgrid_event_destroy(void *ev) { return check(hipEventDestroy((hipEvent_t)ev), "hipEventDestroy"); }
extern "C" int hgrid_event_record(void *ev) {
  return check(hipEventRecord((hipEvent_t)ev, lg::g_stream), "hipEventRecord");
}
extern "C" int hgrid_event_elapsed_ms(void *start, void *stop, float *ms) {
  if (check(hipEventSynchronize((hipEvent_t)stop), "hipEventSynchronize")) return -1;
  return check(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}
extern "C" int hgrid_mem_info(size_t *free_bytes, size_t *total_bytes) {
  return check(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}
