// noalloc.hip -- ls_amd_state_info_noalloc: the state_info of
// ls_hs_state_info_halide_kernel (host arrays in and out) with its device
// staging in one fixed file-scope device array instead of per-call device
// allocations, for callers that cannot allocate (HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The staging holds at most kNoAllocMaxBatch states: their input words, the
// representatives, the characters and the norms (40 bytes per state).  The
// kernel writes every output entry of a call, so nothing is carried over
// from the previous call.  Calls in one process share the array and are
// serialised by a lock; each process has its own copy of it on each device.
#include <mutex>

#include "lsamd_internal.hpp"

namespace {
constexpr int64_t kNoAllocMaxBatch = int64_t{1} << 16;
constexpr size_t kStagingBytesPerState = 8 + 8 + 16 + 8;
constexpr size_t kScratchBytes = static_cast<size_t>(kNoAllocMaxBatch) * kStagingBytesPerState; // 2.5 MiB
constexpr int kMaxDevices = 64;
} // namespace

// The staging array.  In host code this name is a stand-in, not a device
// address: host code uses the address hipGetSymbolAddress returns for it.
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

namespace {
std::mutex g_noalloc_mutex;
uint8_t *g_scratch_ptr[kMaxDevices] = {}; // device address of `scratch` on each device

// Device address of the staging array on the current device, resolved once
// per device.
bool scratch_address(uint8_t *&out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return lsamd::fail("ls_amd_state_info_noalloc: no current device");
  if (g_scratch_ptr[dev] == nullptr) {
    void *ptr = nullptr;
    hipError_t const e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || ptr == nullptr)
      return lsamd::fail(std::string("ls_amd_state_info_noalloc: hipGetSymbolAddress failed: ") +
                         hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<uint8_t *>(ptr);
  }
  out = g_scratch_ptr[dev];
  return true;
}

// Host <-> device copy of `count` 8-byte words with a stride (in words) on
// the host side.
hipError_t copy_words(void *dst, size_t dst_stride, void const *src, size_t src_stride, int64_t count,
                      hipMemcpyKind kind) {
  if (dst_stride == 1 && src_stride == 1)
    return hipMemcpyAsync(dst, src, 8 * static_cast<size_t>(count), kind, lsamd::stream());
  return hipMemcpy2DAsync(dst, 8 * dst_stride, src, 8 * src_stride, 8, static_cast<size_t>(count), kind,
                          lsamd::stream());
}
} // namespace

extern "C" int ls_amd_state_info_noalloc(ptrdiff_t batch_size, uint64_t const *alphas, ptrdiff_t alphas_stride,
                                         uint64_t *betas, ptrdiff_t betas_stride, ls_hs_scalar *characters,
                                         double *norms, void const *private_data) {
  using namespace lsamd;
  auto const *g = static_cast<GroupData const *>(private_data);
  if (g == nullptr || g->magic != kGroupMagic)
    return fail("ls_amd_state_info_noalloc: kernel data was not created by this engine") ? 0 : -1;
  if (batch_size > kNoAllocMaxBatch)
    return fail("ls_amd_state_info_noalloc: a batch of " + std::to_string(batch_size) +
                " states exceeds the maximum of " + std::to_string(kNoAllocMaxBatch))
               ? 0
               : -1;
  if (batch_size <= 0) return 0;
  if (alphas == nullptr || betas == nullptr || characters == nullptr || norms == nullptr || alphas_stride < 1 ||
      betas_stride < 1)
    return fail("ls_amd_state_info_noalloc: null array or stride < 1") ? 0 : -1;
  if (!ensure_device()) return -1;
  std::lock_guard<std::mutex> lock(g_noalloc_mutex);
  uint8_t *base = nullptr;
  if (!scratch_address(base)) return -1;
  auto *d_alphas = reinterpret_cast<uint64_t *>(base);
  auto *d_reps = reinterpret_cast<uint64_t *>(base + 8 * kNoAllocMaxBatch);
  auto *d_chars = reinterpret_cast<double *>(base + 16 * kNoAllocMaxBatch);
  auto *d_norms = reinterpret_cast<double *>(base + 32 * kNoAllocMaxBatch);
  int64_t const n = batch_size;
  if (copy_words(d_alphas, 1, alphas, static_cast<size_t>(alphas_stride), n, hipMemcpyHostToDevice) != hipSuccess)
    return fail("ls_amd_state_info_noalloc: copy of the input failed") ? 0 : -1;
  if (!launch_state_info(*g, n, d_alphas, 1, d_reps, 1, d_chars, d_norms)) return -1;
  if (copy_words(betas, static_cast<size_t>(betas_stride), d_reps, 1, n, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpyAsync(characters, d_chars, 16 * static_cast<size_t>(n), hipMemcpyDeviceToHost, stream()) !=
          hipSuccess ||
      hipMemcpyAsync(norms, d_norms, 8 * static_cast<size_t>(n), hipMemcpyDeviceToHost, stream()) != hipSuccess)
    return fail("ls_amd_state_info_noalloc: copy of the outputs failed") ? 0 : -1;
  if (hipStreamSynchronize(stream()) != hipSuccess)
    return fail("ls_amd_state_info_noalloc: stream synchronisation failed") ? 0 : -1;
  return 0;
}
