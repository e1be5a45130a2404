// tspool.h -- the process-wide cache of device blocks, streams, pinned mailboxes and pinned
// staging blocks shared by every Twist/Shout oracle (one device per process).
//
// Many host threads build, drive and drop oracles at once, so every structure here is touched
// under one mutex and each object is handed to exactly one owner at a time.  The allocator
// calls themselves (which serialize inside the runtime and may wait for the device) run outside
// the lock.  The pool is a template over its allocator backend so that the same code is built
// twice: over HIP in the library (twist.hip) and over malloc in a host-only stress test run
// under ThreadSanitizer and AddressSanitizer (tests/tspool_stress.cpp).
//
// Backend:  using Stream;  static bool dev_alloc(void**, size_t);  static void alloc_failed(size_t)
//           (throws);  static void dev_free(void*);
//           static void host_alloc(void**, size_t) (throws);  static void host_free(void*);
//           static Stream stream_create() (throws);  static void clear_error();
// K: the mailbox element type; a mailbox holds mbox_elems of them.
#pragma once
#include <cstddef>
#include <map>
#include <mutex>
#include <vector>

namespace ns {

template <class Backend, class K>
class BlockPool {
 public:
  using Stream = typename Backend::Stream;
  BlockPool(size_t mbox_elems, size_t max_cached, size_t max_pinned)
      : mbox_elems_(mbox_elems), max_cached_(max_cached), max_pinned_(max_pinned) {}
  BlockPool(const BlockPool&) = delete;
  BlockPool& operator=(const BlockPool&) = delete;

  static size_t size_class(size_t n) {
    size_t c = 256;
    while (c < n) c <<= 1;
    return c;
  }
  // device block of exactly `cap` bytes (a size class); contents undefined
  void* take(size_t cap) {
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = blocks_.find(cap);
      if (it != blocks_.end()) {
        void* p = it->second;
        blocks_.erase(it);
        cached_ -= cap;
        return p;
      }
    }
    void* p = nullptr;
    if (!Backend::dev_alloc(&p, cap)) {
      Backend::clear_error();
      flush();
      if (!Backend::dev_alloc(&p, cap)) Backend::alloc_failed(cap);
    }
    return p;
  }
  // the caller guarantees that no work in flight still uses the block
  void give(void* p, size_t cap) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (cached_ + cap <= max_cached_) {
        blocks_.emplace(cap, p);
        cached_ += cap;
        return;
      }
    }
    Backend::dev_free(p);
  }
  void flush() {
    std::multimap<size_t, void*> all;
    {
      std::lock_guard<std::mutex> g(mu_);
      all.swap(blocks_);
      cached_ = 0;
    }
    for (auto& b : all) Backend::dev_free(b.second);
  }
  Stream stream() {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (!streams_.empty()) {
        Stream s = streams_.back();
        streams_.pop_back();
        return s;
      }
    }
    return Backend::stream_create();
  }
  // the stream must have drained
  void put_stream(Stream s) {
    std::lock_guard<std::mutex> g(mu_);
    streams_.push_back(s);
  }
  K* mailbox() {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (!mboxes_.empty()) {
        K* m = mboxes_.back();
        mboxes_.pop_back();
        return m;
      }
    }
    void* p = nullptr;
    Backend::host_alloc(&p, mbox_elems_ * sizeof(K));
    return (K*)p;
  }
  void put_mailbox(K* m) {
    std::lock_guard<std::mutex> g(mu_);
    mboxes_.push_back(m);
  }
  // pinned host blocks for the constructors' uploads (size classes as for device blocks)
  void* pinned_take(size_t cap) {
    {
      std::lock_guard<std::mutex> g(mu_);
      auto it = pinned_.find(cap);
      if (it != pinned_.end()) {
        void* p = it->second;
        pinned_.erase(it);
        pinned_cached_ -= cap;
        return p;
      }
    }
    void* p = nullptr;
    Backend::host_alloc(&p, cap);
    return p;
  }
  // the copies that read the block must have completed
  void pinned_give(void* p, size_t cap) {
    {
      std::lock_guard<std::mutex> g(mu_);
      if (pinned_cached_ + cap <= max_pinned_) {
        pinned_.emplace(cap, p);
        pinned_cached_ += cap;
        return;
      }
    }
    Backend::host_free(p);
  }
  // cached totals, for tests
  size_t cached_device_bytes() {
    std::lock_guard<std::mutex> g(mu_);
    return cached_;
  }
  size_t cached_pinned_bytes() {
    std::lock_guard<std::mutex> g(mu_);
    return pinned_cached_;
  }

 private:
  const size_t mbox_elems_, max_cached_, max_pinned_;
  std::mutex mu_;
  std::multimap<size_t, void*> blocks_, pinned_;
  size_t cached_ = 0, pinned_cached_ = 0;
  std::vector<Stream> streams_;
  std::vector<K*> mboxes_;
};

}  // namespace ns
