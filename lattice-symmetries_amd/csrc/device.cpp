// device.cpp -- device set-up, stream, timers, errors and the device-memory
// helpers of the additive ABI.
#include <atomic>
#include <cstdlib>
#include <cstring>

#include "lsamd_internal.hpp"

namespace lsamd {
namespace {

std::mutex g_mutex;
bool g_device_ready = false;
bool g_device_failed = false;
hipStream_t g_stream = nullptr;
thread_local std::string g_last_error;

hipEvent_t g_ev_begin[kTimerCount] = {};
hipEvent_t g_ev_end[kTimerCount] = {};
bool g_ev_used[kTimerCount] = {};

} // namespace

void set_last_error(std::string const &msg) { g_last_error = msg; }

bool fail(std::string const &msg) {
  g_last_error = msg;
  return false;
}

bool ensure_device() {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_device_ready) return true;
  if (g_device_failed) {
    g_last_error = "no usable gfx950 device";
    return false;
  }
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
    g_device_failed = true;
    g_last_error = "no HIP device visible";
    return false;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, 0) != hipSuccess) {
    g_device_failed = true;
    g_last_error = "hipGetDeviceProperties failed";
    return false;
  }
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_device_failed = true;
    g_last_error = std::string("device is ") + prop.gcnArchName + ", engine is built for gfx950";
    return false;
  }
  if (hipSetDevice(0) != hipSuccess || hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking) != hipSuccess) {
    g_device_failed = true;
    g_last_error = "stream creation failed";
    return false;
  }
  for (int i = 0; i < kTimerCount; ++i) {
    (void)hipEventCreate(&g_ev_begin[i]);
    (void)hipEventCreate(&g_ev_end[i]);
  }
  g_device_ready = true;
  return true;
}

hipStream_t stream() { return g_stream; }

void timer_begin(int family) {
  if (family < 0 || family >= kTimerCount || !g_device_ready) return;
  (void)hipEventRecord(g_ev_begin[family], g_stream);
}

void timer_end(int family) {
  if (family < 0 || family >= kTimerCount || !g_device_ready) return;
  (void)hipEventRecord(g_ev_end[family], g_stream);
  g_ev_used[family] = true;
}

} // namespace lsamd

using namespace lsamd;

extern "C" int ls_amd_device_available(void) { return ensure_device() ? 0 : 1; }

extern "C" char const *ls_amd_last_error(void) { return g_last_error.c_str(); }

extern "C" void ls_amd_clear_error(void) { g_last_error.clear(); }

extern "C" void *ls_amd_device_malloc(size_t bytes) {
  if (!ensure_device()) return nullptr;
  void *p = nullptr;
  if (hipMalloc(&p, bytes == 0 ? 1 : bytes) != hipSuccess) {
    set_last_error("hipMalloc failed");
    return nullptr;
  }
  return p;
}

extern "C" void ls_amd_device_free(void *ptr) {
  if (ptr != nullptr) (void)hipFree(ptr);
}

extern "C" int ls_amd_memcpy_h2d(void *dst, void const *src, size_t bytes) {
  if (!ensure_device()) return -1;
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()) != hipSuccess) return -1;
  return hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_memcpy_d2h(void *dst, void const *src, size_t bytes) {
  if (!ensure_device()) return -1;
  if (hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()) != hipSuccess) return -1;
  return hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" int ls_amd_synchronize(void) {
  if (!ensure_device()) return -1;
  return hipStreamSynchronize(stream()) == hipSuccess ? 0 : -1;
}

extern "C" double ls_amd_last_kernel_ms(int family) {
  if (family < 0 || family >= kTimerCount || !g_ev_used[family]) return -1.0;
  if (hipEventSynchronize(g_ev_end[family]) != hipSuccess) return -1.0;
  float ms = 0;
  if (hipEventElapsedTime(&ms, g_ev_begin[family], g_ev_end[family]) != hipSuccess) return -1.0;
  return static_cast<double>(ms);
}
