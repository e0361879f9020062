// Host <-> device copies of caller arrays (host entry points).
//
// The arrays a caller such as the Fortran SOLVE passes are pageable. For
// those the runtime stages each copy through pinned memory itself, on one
// host thread, which caps a transfer at that thread's memcpy rate. Large
// transfers here go through a pinned ring of their own instead: a pool of
// host threads copies one chunk into pinned memory while the DMA engine moves
// the previous chunk. Small transfers keep the plain asynchronous copy.
//
// Semantics are those of hipMemcpyAsync on pageable memory: to_device
// returns once the caller's array has been read (the device copy is queued
// on the stream); to_host returns with the caller's array written.

#include <condition_variable>
#include <thread>

namespace bhip {

class HostCopier {
 public:
  static constexpr size_t kChunk = (size_t)8 << 20;  // bytes per pinned slot
  static constexpr int kSlots = 3;
  static constexpr size_t kMinStaged = (size_t)4 << 20;

  // process-wide instance, never destroyed (its threads wait on a condition
  // variable when idle; no HIP call is left for exit-time destructors)
  static HostCopier& get() {
    static HostCopier* inst = new HostCopier();
    return *inst;
  }

  // dev: the device of stream s, current on the calling thread
  hipError_t to_device(void* dst, const void* src, size_t bytes, hipStream_t s, int dev) {
    if (bytes < kMinStaged) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    std::lock_guard<std::mutex> use(use_mu_);
    hipError_t e = init(dev);
    if (e != hipSuccess) return e;
    char** pin = ring_[dev].pin;
    hipEvent_t* ev = ring_[dev].ev;
    const char* in = static_cast<const char*>(src);
    char* out = static_cast<char*>(dst);
    const size_t nch = (bytes + kChunk - 1) / kChunk;
    for (size_t k = 0; k < nch; ++k) {
      const int slot = (int)(k % kSlots);
      const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
      // the slot's previous DMA (of this call or an earlier one) has read it
      if ((e = hipEventSynchronize(ev[slot])) != hipSuccess) return e;
      pcopy(pin[slot], in + off, len);
      if ((e = hipMemcpyAsync(out + off, pin[slot], len, hipMemcpyHostToDevice, s)) !=
          hipSuccess)
        return e;
      if ((e = hipEventRecord(ev[slot], s)) != hipSuccess) return e;
    }
    return hipSuccess;
  }

  hipError_t to_host(void* dst, const void* src, size_t bytes, hipStream_t s, int dev) {
    if (bytes < kMinStaged) {
      hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
      return e != hipSuccess ? e : hipStreamSynchronize(s);
    }
    std::lock_guard<std::mutex> use(use_mu_);
    hipError_t e = init(dev);
    if (e != hipSuccess) return e;
    char** pin = ring_[dev].pin;
    hipEvent_t* ev = ring_[dev].ev;
    const char* in = static_cast<const char*>(src);
    char* out = static_cast<char*>(dst);
    const size_t nch = (bytes + kChunk - 1) / kChunk;
    auto issue = [&](size_t k) {
      const int slot = (int)(k % kSlots);
      const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
      hipError_t r = hipEventSynchronize(ev[slot]);
      if (r == hipSuccess)
        r = hipMemcpyAsync(pin[slot], in + off, len, hipMemcpyDeviceToHost, s);
      if (r == hipSuccess) r = hipEventRecord(ev[slot], s);
      return r;
    };
    for (size_t k = 0; k < std::min(nch, (size_t)kSlots); ++k)
      if ((e = issue(k)) != hipSuccess) return e;
    for (size_t k = 0; k < nch; ++k) {
      const int slot = (int)(k % kSlots);
      const size_t off = k * kChunk, len = std::min(kChunk, bytes - off);
      if ((e = hipEventSynchronize(ev[slot])) != hipSuccess) return e;
      pcopy(out + off, pin[slot], len);
      if (k + kSlots < nch && (e = issue(k + kSlots)) != hipSuccess) return e;
    }
    return hipSuccess;
  }

 private:
  HostCopier() = default;

  // the ring of device dev (pinned slots and the events of their last DMA,
  // created with dev current), and the thread pool on first use
  hipError_t init(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    Ring& r = ring_[dev];
    if (!r.pin[0]) {
      void* p = nullptr;
      hipError_t e = hipHostMalloc(&p, kChunk * kSlots, hipHostMallocDefault);
      if (e != hipSuccess) return e;
      for (int k = 0; k < kSlots; ++k) {
        if ((e = hipEventCreateWithFlags(&r.ev[k], hipEventDisableTiming)) != hipSuccess)
          return e;
      }
      for (int k = 0; k < kSlots; ++k) r.pin[k] = static_cast<char*>(p) + k * kChunk;
    }
    if (!pool_started_) {
      const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
      parts_ = (int)std::min(8u, hw);
      for (int k = 1; k < parts_; ++k) {
        std::thread t([this, k] { worker(k); });
        t.detach();
      }
      pool_started_ = true;
    }
    return hipSuccess;
  }

  // part k of parts_ of a copy, split at 4 KiB boundaries
  void piece(char* d, const char* s, size_t n, int k) const {
    const size_t per = ((n + parts_ - 1) / parts_ + 4095) & ~(size_t)4095;
    const size_t a = std::min(n, per * k), b = std::min(n, per * (k + 1));
    if (b > a) std::memcpy(d + a, s + a, b - a);
  }

  void pcopy(char* d, const char* s, size_t n) {
    {
      std::lock_guard<std::mutex> lk(mu_);
      dst_ = d;
      src_ = s;
      len_ = n;
      pending_ = parts_ - 1;
      ++gen_;
    }
    cv_.notify_all();
    piece(d, s, n, 0);
    std::unique_lock<std::mutex> lk(mu_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

  void worker(int k) {
    long seen = 0;
    for (;;) {
      char* d;
      const char* s;
      size_t n;
      {
        std::unique_lock<std::mutex> lk(mu_);
        cv_.wait(lk, [&] { return gen_ != seen; });
        seen = gen_;
        d = dst_;
        s = src_;
        n = len_;
      }
      piece(d, s, n, k);
      std::lock_guard<std::mutex> lk(mu_);
      if (--pending_ == 0) done_.notify_one();
    }
  }

  static constexpr int kMaxDevices = 64;
  struct Ring {
    char* pin[kSlots] = {nullptr, nullptr, nullptr};
    hipEvent_t ev[kSlots] = {nullptr, nullptr, nullptr};
  };
  std::mutex use_mu_;  // one transfer at a time (the thread pool is shared)
  Ring ring_[kMaxDevices];
  bool pool_started_ = false;
  int parts_ = 1;
  // the pool's current copy
  std::mutex mu_;
  std::condition_variable cv_, done_;
  char* dst_ = nullptr;
  const char* src_ = nullptr;
  size_t len_ = 0;
  long gen_ = 0;
  int pending_ = 0;
};

// copies of caller (host) arrays, to and from the context's stream
static inline hipError_t to_device(bhip_ctx* c, void* dst, const void* src, size_t bytes) {
  return HostCopier::get().to_device(dst, src, bytes, c->stream, c->device);
}
static inline hipError_t to_host(bhip_ctx* c, void* dst, const void* src, size_t bytes) {
  return HostCopier::get().to_host(dst, src, bytes, c->stream, c->device);
}

}  // namespace bhip
