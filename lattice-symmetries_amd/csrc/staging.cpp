// staging.cpp -- copies of pageable host arrays to and from device memory
// through two pinned staging buffers.
//
// The reference's entry points hand over plain host arrays (Chapel or numpy
// memory, pageable).  A pageable hipMemcpy goes through the runtime's own
// small staging at a fraction of the link rate.  Here a pool of host threads
// copies one chunk into a pinned buffer while the DMA engine moves the
// previous chunk over PCIe on a copy stream of its own (default priority),
// and the reverse for device-to-host copies.  Small arrays, and a call that
// finds the buffers in use by another thread, take the plain path.
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "lsamd_internal.hpp"

namespace lsamd {
namespace {

constexpr size_t kDefaultStageMiB = 64; // per buffer: per-chunk synchronisation costs more than the pipeline fill
constexpr size_t kSmallCopy = size_t(4) << 20; // below this, one plain copy
constexpr int kMaxCopyThreads = 16;

// LSAMD_STAGING_MIB sets the buffer size (1 to 256 MiB); LSAMD_STAGING=0
// sends every copy the plain way.
size_t stage_bytes() {
  static size_t const b = [] {
    size_t mib = kDefaultStageMiB;
    if (char const *e = std::getenv("LSAMD_STAGING_MIB")) mib = std::strtoull(e, nullptr, 10);
    return std::min<size_t>(256, std::max<size_t>(1, mib)) << 20;
  }();
  return b;
}
bool staging_enabled() {
  static bool const on = [] {
    char const *e = std::getenv("LSAMD_STAGING");
    return e == nullptr || std::strcmp(e, "0") != 0;
  }();
  return on;
}

// Host threads that split one memcpy between them.  Created once and kept
// for the life of the process.
class CopyPool {
public:
  explicit CopyPool(int n) : n_(n) {
    for (int t = 0; t < n_; ++t) threads_.emplace_back([this, t] { work(t); });
  }
  void copy(void *dst, void const *src, size_t bytes) {
    if (n_ <= 1 || bytes < (size_t(1) << 20)) {
      std::memcpy(dst, src, bytes);
      return;
    }
    std::unique_lock<std::mutex> lk(m_);
    dst_ = static_cast<char *>(dst);
    src_ = static_cast<char const *>(src);
    bytes_ = bytes;
    pending_ = n_;
    ++gen_;
    cv_.notify_all();
    done_.wait(lk, [&] { return pending_ == 0; });
  }

private:
  void work(int t) {
    uint64_t seen = 0;
    std::unique_lock<std::mutex> lk(m_);
    for (;;) {
      cv_.wait(lk, [&] { return gen_ != seen; });
      seen = gen_;
      char *const d = dst_;
      char const *const s = src_;
      size_t const b = bytes_;
      lk.unlock();
      size_t const per = ((b + n_ - 1) / n_ + 63) & ~size_t(63);
      size_t const a = std::min(b, per * static_cast<size_t>(t)), e = std::min(b, a + per);
      if (e > a) std::memcpy(d + a, s + a, e - a);
      lk.lock();
      if (--pending_ == 0) done_.notify_all();
    }
  }
  int n_;
  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable cv_, done_;
  uint64_t gen_ = 0;
  int pending_ = 0;
  char *dst_ = nullptr;
  char const *src_ = nullptr;
  size_t bytes_ = 0;
};

struct Staging {
  std::mutex busy;
  bool ready = false, failed = false;
  void *buf[2] = {nullptr, nullptr};
  hipStream_t copy = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  hipEvent_t start = nullptr, done = nullptr;
  std::vector<hipEvent_t> produced; // one per chunk of produce_to_host (grow-only)
  CopyPool *pool = nullptr;

  // Called with `busy` held.
  bool chunk_events(size_t n) {
    while (produced.size() < n) {
      hipEvent_t e;
      if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return false;
      produced.push_back(e);
    }
    return true;
  }

  // Called with `busy` held.
  bool init() {
    if (ready) return true;
    if (failed) return false;
    failed = true;
    for (void *&b : buf)
      if (hipHostMalloc(&b, stage_bytes(), hipHostMallocDefault) != hipSuccess) return false;
    if (hipStreamCreateWithFlags(&copy, hipStreamNonBlocking) != hipSuccess) return false;
    for (hipEvent_t *e : {&ev[0], &ev[1], &start, &done})
      if (hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return false;
    unsigned const hw = std::thread::hardware_concurrency();
    pool = new CopyPool(static_cast<int>(std::max(1u, std::min<unsigned>(hw, kMaxCopyThreads))));
    failed = false;
    ready = true;
    return true;
  }
};

Staging g_staging;

bool plain_to_device(void *d, void const *h, size_t bytes) {
  LSAMD_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  return true;
}

bool plain_to_host(void *h, void const *d, size_t bytes) {
  LSAMD_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, stream()));
  LSAMD_HIP(hipStreamSynchronize(stream()));
  return true;
}

} // namespace

bool copy_to_device(void *d, void const *h, size_t bytes) {
  if (bytes == 0) return true;
  std::unique_lock<std::mutex> lk(g_staging.busy, std::try_to_lock);
  if (bytes < kSmallCopy || !staging_enabled() || !lk.owns_lock() || !g_staging.init())
    return plain_to_device(d, h, bytes);
  Staging &S = g_staging;
  auto body = [&]() -> bool {
    // earlier work on the engine stream may still read the destination
    LSAMD_HIP(hipEventRecord(S.start, stream()));
    LSAMD_HIP(hipStreamWaitEvent(S.copy, S.start, 0));
    char *const dp = static_cast<char *>(d);
    char const *const hp = static_cast<char const *>(h);
    size_t const chunk = stage_bytes();
    size_t k = 0;
    for (size_t off = 0; off < bytes; off += chunk, ++k) {
      int const b = static_cast<int>(k & 1);
      size_t const len = std::min(chunk, bytes - off);
      if (k >= 2) LSAMD_HIP(hipEventSynchronize(S.ev[b])); // its previous chunk has left
      S.pool->copy(S.buf[b], hp + off, len);
      LSAMD_HIP(hipMemcpyAsync(dp + off, S.buf[b], len, hipMemcpyHostToDevice, S.copy));
      LSAMD_HIP(hipEventRecord(S.ev[b], S.copy));
    }
    LSAMD_HIP(hipEventRecord(S.done, S.copy));
    LSAMD_HIP(hipStreamWaitEvent(stream(), S.done, 0));
    return true;
  };
  bool const ok = body();
  // the buffers are free (and the copies complete) before another call may use them
  bool const drained = hipStreamSynchronize(S.copy) == hipSuccess;
  return ok && (drained || fail("staging copy stream failed"));
}

bool copy_to_host(void *h, void const *d, size_t bytes) {
  if (bytes == 0) return true;
  std::unique_lock<std::mutex> lk(g_staging.busy, std::try_to_lock);
  if (bytes < kSmallCopy || !staging_enabled() || !lk.owns_lock() || !g_staging.init())
    return plain_to_host(h, d, bytes);
  Staging &S = g_staging;
  char const *const dp = static_cast<char const *>(d);
  char *const hp = static_cast<char *>(h);
  size_t const chunk = stage_bytes();
  size_t const chunks = (bytes + chunk - 1) / chunk;
  auto issue = [&](size_t k) -> bool {
    int const b = static_cast<int>(k & 1);
    size_t const off = k * chunk, len = std::min(chunk, bytes - off);
    LSAMD_HIP(hipMemcpyAsync(S.buf[b], dp + off, len, hipMemcpyDeviceToHost, S.copy));
    LSAMD_HIP(hipEventRecord(S.ev[b], S.copy));
    return true;
  };
  auto body = [&]() -> bool {
    // the source is written by work already queued on the engine stream
    LSAMD_HIP(hipEventRecord(S.start, stream()));
    LSAMD_HIP(hipStreamWaitEvent(S.copy, S.start, 0));
    if (!issue(0) || (chunks > 1 && !issue(1))) return false;
    for (size_t k = 0; k < chunks; ++k) {
      int const b = static_cast<int>(k & 1);
      size_t const off = k * chunk, len = std::min(chunk, bytes - off);
      LSAMD_HIP(hipEventSynchronize(S.ev[b]));
      S.pool->copy(hp + off, S.buf[b], len);
      if (k + 2 < chunks && !issue(k + 2)) return false;
    }
    return true;
  };
  bool const ok = body();
  bool const drained = hipStreamSynchronize(S.copy) == hipSuccess;
  return ok && (drained || fail("staging copy stream failed"));
}

bool produce_to_host(void *h, void const *d, size_t bytes, size_t unit,
                     std::function<bool(size_t, size_t)> const &produce) {
  if (bytes == 0) return true;
  unit = std::max<size_t>(1, unit);
  std::unique_lock<std::mutex> lk(g_staging.busy, std::try_to_lock);
  size_t const chunk = stage_bytes() / unit * unit;
  size_t const chunks = chunk == 0 ? 0 : (bytes + chunk - 1) / chunk;
  bool const staged = bytes >= kSmallCopy && staging_enabled() && lk.owns_lock() && chunk > 0 &&
                      g_staging.init() && g_staging.chunk_events(chunks);
  if (!staged) {
    if (!produce(0, bytes)) return false;
    return plain_to_host(h, d, bytes);
  }
  Staging &S = g_staging;
  char const *const dp = static_cast<char const *>(d);
  char *const hp = static_cast<char *>(h);
  auto issue = [&](size_t k) -> bool {
    int const b = static_cast<int>(k & 1);
    size_t const off = k * chunk, len = std::min(chunk, bytes - off);
    LSAMD_HIP(hipStreamWaitEvent(S.copy, S.produced[k], 0));
    LSAMD_HIP(hipMemcpyAsync(S.buf[b], dp + off, len, hipMemcpyDeviceToHost, S.copy));
    LSAMD_HIP(hipEventRecord(S.ev[b], S.copy));
    return true;
  };
  auto body = [&]() -> bool {
    // all the producing work is queued first, each chunk's end marked
    for (size_t k = 0; k < chunks; ++k) {
      size_t const off = k * chunk;
      if (!produce(off, std::min(chunk, bytes - off))) return false;
      LSAMD_HIP(hipEventRecord(S.produced[k], stream()));
    }
    if (!issue(0) || (chunks > 1 && !issue(1))) return false;
    for (size_t k = 0; k < chunks; ++k) {
      int const b = static_cast<int>(k & 1);
      size_t const off = k * chunk, len = std::min(chunk, bytes - off);
      LSAMD_HIP(hipEventSynchronize(S.ev[b]));
      S.pool->copy(hp + off, S.buf[b], len);
      if (k + 2 < chunks && !issue(k + 2)) return false;
    }
    return true;
  };
  bool const ok = body();
  bool const drained = hipStreamSynchronize(S.copy) == hipSuccess && hipStreamSynchronize(stream()) == hipSuccess;
  return ok && (drained || fail("staging copy stream failed"));
}

} // namespace lsamd
