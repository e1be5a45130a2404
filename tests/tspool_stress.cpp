// Host-only stress test of the oracle block pool (nightstream_amd/csrc/tspool.h) under
// ThreadSanitizer / AddressSanitizer: the pool's own code over a malloc backend, driven by 16
// threads that build, round, fold and drop emulated oracles the way TsOracleGpu does.
//
// "Streams" are executor threads that run queued work in order, so work is in flight while the
// owner's thread goes on, as on a GPU.  Every block is stamped with its owner and generation by
// the owner, and the queued work checks the stamp before and after touching the block: a block
// handed to a second owner while the first one's work is still queued shows up as a stamp
// mismatch (and, under TSan, as a data race between the two writers).
//
// The owner protocol restated here is TsOracleGpu's: uploads go through pinned staging blocks
// that return to the pool only after the stream drained; a block that is replaced by a larger
// one is released after a device-wide drain; the destructor drains the stream before its stream,
// staging blocks, mailbox and device blocks go back.
//
// Build: g++ -O1 -g -std=c++17 -pthread -fsanitize=thread (or address) -I<csrc> tspool_stress.cpp
// Run:   ./a.out [threads] [oracles per thread]; exit 0 and "ok" on success.
// -DPLANT_RELEASE_BEFORE_DRAIN builds the destructor without its drain, a mistake the stamps
// (and TSan) must catch.
#include <atomic>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "tspool.h"

static std::atomic<long> g_errors{0};
static void fail(const char* what) {
  if (g_errors.fetch_add(1) < 10) fprintf(stderr, "stamp mismatch: %s\n", what);
}

// ---- emulated streams ----
struct EmuStream {
  std::mutex mu;
  std::condition_variable cv, idle;
  std::deque<std::function<void()>> q;
  bool busy = false, stop = false;
  std::thread th;
  EmuStream() : th([this] { run(); }) {}
  void run() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv.wait(lk, [&] { return stop || !q.empty(); });
      if (q.empty()) return;
      auto f = std::move(q.front());
      q.pop_front();
      busy = true;
      lk.unlock();
      f();
      lk.lock();
      busy = false;
      if (q.empty()) idle.notify_all();
    }
  }
  void enqueue(std::function<void()> f) {
    std::lock_guard<std::mutex> g(mu);
    q.push_back(std::move(f));
    cv.notify_one();
  }
  void sync() {
    std::unique_lock<std::mutex> lk(mu);
    idle.wait(lk, [&] { return q.empty() && !busy; });
  }
  void shutdown() {
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
      cv.notify_one();
    }
    th.join();
  }
};

static std::mutex g_streams_mu;
static std::vector<EmuStream*> g_streams;
static void device_sync() {  // hipDeviceSynchronize: every stream drains
  std::vector<EmuStream*> all;
  {
    std::lock_guard<std::mutex> g(g_streams_mu);
    all = g_streams;
  }
  for (auto* s : all) s->sync();
}

static std::atomic<long> g_dev_allocs{0};
struct MallocBackend {
  using Stream = EmuStream*;
  static bool dev_alloc(void** p, size_t n) {
    // every 97th allocation fails once, to run the pool's flush-and-retry path
    if (g_dev_allocs.fetch_add(1) % 97 == 96) return false;
    *p = malloc(n);
    return *p != nullptr;
  }
  static void alloc_failed(size_t) { throw std::bad_alloc(); }
  static void dev_free(void* p) { free(p); }
  static void host_alloc(void** p, size_t n) {
    *p = malloc(n);
    if (!*p) throw std::bad_alloc();
  }
  static void host_free(void* p) { free(p); }
  static EmuStream* stream_create() {
    auto* s = new EmuStream;
    std::lock_guard<std::mutex> g(g_streams_mu);
    g_streams.push_back(s);
    return s;
  }
  static void clear_error() {}
};

struct K2 {
  uint64_t c0, c1;
};
constexpr size_t MBOX = 41;
using Pool = ns::BlockPool<MallocBackend, K2>;
static Pool* g_pool;

// a block's first 16 bytes: owner id and generation
static void stamp(void* p, uint64_t owner, uint64_t gen) {
  uint64_t* w = (uint64_t*)p;
  w[0] = owner;
  w[1] = gen;
}
static bool stamped(const void* p, uint64_t owner, uint64_t gen) {
  const uint64_t* w = (const uint64_t*)p;
  return w[0] == owner && w[1] == gen;
}

struct Buf {  // TsBuf
  void* p = nullptr;
  size_t bytes = 0;
  ~Buf() {
    if (p) g_pool->give(p, bytes);
  }
  // replacing a block: the old one may still be used by queued work, so the device drains first
  void ensure(size_t n) {
    if (n <= bytes && p) return;
    if (p) {
      device_sync();
      g_pool->give(p, bytes);
      p = nullptr;
      bytes = 0;
    }
    bytes = Pool::size_class(n);
    p = g_pool->take(bytes);
  }
};

struct Oracle {  // TsOracleGpu's ownership protocol
  uint64_t id, gen = 0;
  EmuStream* st;
  K2* mbox;
  Buf bufs[4];
  std::vector<std::pair<void*, size_t>> stage;
  size_t stage_off = 0;

  explicit Oracle(uint64_t id_) : id(id_) {
    st = g_pool->stream();
    mbox = g_pool->mailbox();
  }
  void* stage_take(size_t bytes) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (stage.empty() || stage_off + bytes > stage.back().second) {
      size_t cap = Pool::size_class(bytes < (1 << 12) ? (1 << 12) : bytes);
      stage.emplace_back(g_pool->pinned_take(cap), cap);
      stage_off = 0;
    }
    void* p = (char*)stage.back().first + stage_off;
    stage_off += bytes;
    return p;
  }
  void stage_release() {
    for (auto& b : stage) g_pool->pinned_give(b.first, b.second);
    stage.clear();
    stage_off = 0;
  }
  // constructor upload: host -> pinned staging -> (queued) device block; the staging blocks
  // return after the stream drained
  void build(std::mt19937_64& rng) {
    gen++;
    for (auto& b : bufs) {
      size_t n = 64 + rng() % (1 << 16);
      b.ensure(n);
      void* h = stage_take(n);
      stamp(h, id, gen);
      memset((char*)h + 16, (int)(id & 0xff), n - 16);
      void* d = b.p;
      uint64_t me = id, g = gen;
      st->enqueue([=] {
        if (!stamped(h, me, g)) fail("staging block changed owner while its copy was queued");
        memcpy(d, h, n);
      });
    }
    st->sync();
    stage_release();
  }
  // a round: queued work reads the blocks and writes the mailbox; the host waits and reads it
  void round() {
    uint64_t me = id, g = gen;
    K2* mb = mbox;
    void* ps[4];
    for (int i = 0; i < 4; i++) ps[i] = bufs[i].p;
    st->enqueue([=] {
      for (auto* p : ps)
        if (!stamped(p, me, g)) fail("device block changed owner while a round was queued");
      for (size_t i = 0; i < MBOX; i++) mb[i] = K2{me, g + i};
    });
    st->sync();
    for (size_t i = 0; i < MBOX; i++)
      if (mbox[i].c0 != id || mbox[i].c1 != gen + i) fail("mailbox overwritten by another owner");
  }
  // a deferred fold: queued work rewrites the blocks (new generation) and the support size in
  // the mailbox; the host reads it only at the next round
  void fold() {
    uint64_t me = id, g = gen, g2 = gen + 1;
    K2* mb = mbox;
    void* ps[4];
    for (int i = 0; i < 4; i++) ps[i] = bufs[i].p;
    st->enqueue([=] {
      for (auto* p : ps) {
        if (!stamped(p, me, g)) fail("device block changed owner while a fold was queued");
        stamp(p, me, g2);
      }
      mb[MBOX - 1] = K2{me, g2};
    });
    gen = g2;
  }
  // growing one block while work is in flight
  void grow(size_t i) {
    st->sync();  // the new block is filled from the old one's stamp by queued work below
    Buf& b = bufs[i];
    b.ensure(b.bytes * 2);
    void* d = b.p;
    uint64_t me = id, g = gen;
    st->enqueue([=] { stamp(d, me, g); });
  }
  ~Oracle() {
#ifndef PLANT_RELEASE_BEFORE_DRAIN  // the planted mistake the test must catch
    st->sync();
#endif
    g_pool->put_stream(st);
    stage_release();
    g_pool->put_mailbox(mbox);
    // bufs return to the pool as members are destroyed, after the drain above
  }
};

int main(int argc, char** argv) {
  const int threads = argc > 1 ? atoi(argv[1]) : 16;
  const int per = argc > 2 ? atoi(argv[2]) : 40;
  // small caps so that blocks also go through the free paths
  g_pool = new Pool(MBOX, (size_t)8 << 20, (size_t)2 << 20);
  std::atomic<uint64_t> next_id{1};
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; t++)
    ts.emplace_back([&, t] {
      std::mt19937_64 rng(1000 + t);
      for (int k = 0; k < per; k++) {
        std::vector<Oracle*> live;
        const int n = 1 + (int)(rng() % 4);
        for (int i = 0; i < n; i++) {
          live.push_back(new Oracle(next_id.fetch_add(1)));
          live.back()->build(rng);
        }
        for (int r = 0; r < 6; r++)
          for (auto* o : live) {
            o->round();
            o->fold();
            if (rng() % 11 == 0) o->grow(rng() % 4);
          }
        for (auto* o : live) delete o;
        if (rng() % 7 == 0) g_pool->flush();
      }
    });
  for (auto& t : ts) t.join();
  device_sync();
  for (auto* s : g_streams) {
    s->shutdown();
    delete s;
  }
  delete g_pool;  // cached blocks are leaked on purpose: the product pool is never destroyed
  if (g_errors.load()) {
    fprintf(stderr, "%ld stamp mismatches\n", g_errors.load());
    return 1;
  }
  printf("ok: %d threads x %d rounds of oracles\n", threads, per);
  return 0;
}
