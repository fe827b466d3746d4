// Device and pinned host memory for the engines: the process-wide caches of
// device allocations (DevicePool, with the streams a release waits for) and
// of pinned host blocks (PinnedPool), and the growable device buffer DevBuf.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

namespace rsa_amd {

#define HIP_OK(expr)                                                                             \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      throw Error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + #expr);           \
  } while (0)

// Process-wide cache of device allocations.  A released buffer goes back to
// the cache once the device is idle (hipFree would wait for that too) and is
// handed to a later request of at least half its size; hipMalloc runs only
// when nothing fits (the whole cache is freed and the call retried when the
// device is out of memory).  Engines created one after another thus reuse the
// same HBM instead of mapping and unmapping it for every run.
// The streams on which a buffer released by this thread may still be in use:
// an engine names its own for the duration of its calls (StreamScope), so a
// release waits for that engine's work only and engines on other threads run
// on.  Without a scope a release waits for the whole device.
struct RelStreams {
  bool set = false;
  hipStream_t s[2] = {nullptr, nullptr};  // both null with set: nothing in flight
};
inline thread_local RelStreams tl_rel;

class StreamScope {
 public:
  StreamScope(hipStream_t a, hipStream_t b) : prev_(tl_rel) {
    tl_rel.set = true;
    tl_rel.s[0] = a;
    tl_rel.s[1] = b;
  }
  ~StreamScope() { tl_rel = prev_; }
  StreamScope(const StreamScope&) = delete;
  StreamScope& operator=(const StreamScope&) = delete;

 private:
  RelStreams prev_;
};

class DevicePool {
 public:
  static DevicePool& get() {
    static DevicePool* pool = new DevicePool();  // never destroyed: outlives every engine
    return *pool;
  }
  void* alloc(size_t bytes, size_t& got) {
    void* p = nullptr;
    {
      std::lock_guard<std::mutex> lk(m_);
      auto it = free_.lower_bound(bytes);
      if (it != free_.end() && it->first / 2 <= bytes) {
        got = it->first;
        p = it->second;
        free_.erase(it);
      }
    }
    if (p) {
      poison(p, got);
      return p;
    }
    hipError_t e = hipMalloc(&p, bytes);
    if (e == hipErrorOutOfMemory) {
      (void)hipGetLastError();
      trim();
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) throw Error(std::string("HIP error ") + hipGetErrorString(e) + " at hipMalloc");
    got = bytes;
    poison(p, got);
    return p;
  }
  // RSA_GPU_POISON=<byte>: every buffer handed out is first filled with that
  // byte, so that a kernel reading memory nobody wrote in this engine reads the
  // same garbage whatever ran there before (a test setting).
  static int poison_byte() {
    static const int v = [] {
      const char* e = std::getenv("RSA_GPU_POISON");
      return e ? (std::atoi(e) & 0xff) : -1;
    }();
    return v;
  }
  static void poison(void* p, size_t bytes) {
    const int b = poison_byte();
    if (b < 0) return;
    HIP_OK(hipMemsetAsync(p, b, bytes, nullptr));
    HIP_OK(hipStreamSynchronize(nullptr));
  }
  void release(void* p, size_t bytes) {
    if (!tl_rel.set) {
      (void)hipDeviceSynchronize();  // kernels queued earlier may still use it
    } else {
      for (hipStream_t st : tl_rel.s)
        if (st) (void)hipStreamSynchronize(st);
    }
    std::lock_guard<std::mutex> lk(m_);
    free_.emplace(bytes, p);
  }
  void trim() {
    (void)hipDeviceSynchronize();
    std::lock_guard<std::mutex> lk(m_);
    for (auto& kv : free_) (void)hipFree(kv.second);
    free_.clear();
  }

 private:
  std::mutex m_;
  std::multimap<size_t, void*> free_;  // by size in bytes
};

// Process-wide cache of pinned host allocations, by flags and size.  A new
// block is written once when allocated: the first CPU pass over fresh pinned
// pages is several times slower than later ones, and the MT ring is filled by
// a thread that must keep ahead of the draws.
class PinnedPool {
 public:
  static PinnedPool& get() {
    static PinnedPool* pool = new PinnedPool();  // never destroyed: outlives every engine
    return *pool;
  }
  void* alloc(size_t bytes, unsigned flags) {
    {
      std::lock_guard<std::mutex> lk(m_);
      for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->flags == flags && it->bytes >= bytes && it->bytes / 2 <= bytes) {
          void* p = it->p;
          used_[p] = *it;
          free_.erase(it);
          return p;
        }
      }
    }
    void* p = nullptr;
    HIP_OK(hipHostMalloc(&p, bytes, flags));
    std::memset(p, 0, bytes);
    std::lock_guard<std::mutex> lk(m_);
    used_[p] = Block{p, bytes, flags};
    return p;
  }
  void release(void* p) {
    if (!p) return;
    std::lock_guard<std::mutex> lk(m_);
    auto it = used_.find(p);
    if (it == used_.end()) return;
    free_.push_back(it->second);
    used_.erase(it);
  }

 private:
  struct Block {
    void* p;
    size_t bytes;
    unsigned flags;
  };
  std::mutex m_;
  std::vector<Block> free_;
  std::map<void*, Block> used_;
};

template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void reserve(size_t n) {  // contents are not preserved; grows by at least 1.5x
    if (n <= cap && p) return;
    const size_t want = std::max<size_t>({n, cap + cap / 2, 1});
    release();
    size_t got = 0;
    p = static_cast<T*>(DevicePool::get().alloc(want * sizeof(T), got));
    cap = got / sizeof(T);
  }
  void release() {
    if (p) DevicePool::get().release(p, cap * sizeof(T));
    p = nullptr;
    cap = 0;
  }
};

}  // namespace rsa_amd
