// Recorded device allocations with a block cache (see lsw_device.h).
//
// hipFree waits for the whole device to go idle, and a server whose workers
// load and evict problems while other workers keep the device busy would
// spend its time in those waits.  Freed blocks are therefore kept:
//   * a freed block first goes to the pending list: kernels queued before
//     the free (on any stream) may still use it;
//   * a request above 64 MiB (or one that would pass the cap below) that no
//     ready block fits, but a pending one would, waits once for the device
//     (hipDeviceSynchronize) and moves every pending block to the ready
//     list; a smaller one takes a new block instead;
//   * a ready block serves a request of at most its size and at least half
//     of it (best fit), so a cached block wastes at most half its bytes;
//   * live and cached bytes together stay below what the device had free at
//     the first allocation (less a margin): cached blocks are returned to the
//     runtime, the largest first, before an allocation would pass that;
//   * when hipMalloc fails, the cache is returned to the runtime (after a
//     device synchronisation) and the allocation is tried once more;
//   * otherwise the cache is kept for the life of the process (or until
//     lsw_release_device_cache): returning hundreds of GB to the driver and
//     allocating them again later stalled single hipMalloc calls for
//     1-2 seconds on a busy device.
// The per-thread balance (what a thread allocated minus what it freed) counts
// the requested sizes in whole blocks (256 B, or 2 MiB pages above 2 MiB), so
// the server charges a problem the memory it asked for, independent of which
// larger cached block served a request.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <unordered_map>

#include "lsw_device.h"

namespace lsw {

namespace {
std::mutex gMu;
struct Live {
  size_t block, charged;  // the block's size; what the owner's balance counts
};
std::unordered_map<void*, Live> gSizes;       // live blocks (handed out)
std::multimap<size_t, void*> gReady;          // cached blocks safe to reuse, by size
std::multimap<size_t, void*> gPending;        // cached blocks freed since the last device sync
std::atomic<long long> gTotal{0};             // bytes handed out
std::atomic<long long> gCached{0};            // bytes held in the cache
thread_local long long tBalance = 0;
thread_local int tIdleFrees = 0;  // depth of IdleFrees scopes on this thread
// seconds this thread spent in uploads and in allocations (set-up log)
thread_local double tUploadS = 0, tAllocS = 0;
thread_local long long tUploads = 0, tAllocs = 0;

struct Lap {
  double& acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit Lap(double& a) : acc(a) {}
  ~Lap() { acc += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

constexpr size_t kGrain = 256;

size_t blockSize(size_t bytes) {
  // 256-byte grain below 2 MiB, 2 MiB pages above
  if (bytes < ((size_t)2 << 20)) return std::max(kGrain, (bytes + kGrain - 1) / kGrain * kGrain);
  const size_t page = (size_t)2 << 20;
  return (bytes + page - 1) / page * page;
}

// largest cached block a request may take: twice its size for small
// blocks, a quarter more above 64 MiB (a large block handed out for a much
// smaller request would hold device memory nobody accounted for)
size_t maxFit(size_t size) { return size <= ((size_t)64 << 20) ? 2 * size : size + size / 4; }

// requests above this size wait for pending blocks rather than take new ones
constexpr size_t kWaitMin = (size_t)64 << 20;

// best fit among cached blocks of [size, maxFit(size)]; the lock is held
void* takeFrom(std::multimap<size_t, void*>& m, size_t size, size_t* got) {
  auto it = m.lower_bound(size);
  if (it == m.end() || it->first > maxFit(size)) return nullptr;
  void* p = it->second;
  *got = it->first;
  m.erase(it);
  return p;
}

bool fitsIn(const std::multimap<size_t, void*>& m, size_t size) {
  auto it = m.lower_bound(size);
  return it != m.end() && it->first <= maxFit(size);
}

// Device bytes the library may hold, live and cached together: what was free
// when it first allocated, less a margin.  Going past the device's memory
// does not fail at once: the driver moves buffers out to host memory, and
// allocations and kernels then take seconds.
long long deviceCap() {
  static const long long cap = [] {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess || fr == 0) {
      (void)hipGetLastError();
      return (long long)1 << 62;
    }
    const long long margin = std::max<long long>((long long)4 << 30, (long long)fr / 50);
    return (long long)fr - margin;
  }();
  return cap;
}

void handOut(void* p, size_t size, size_t charged) {
  gSizes[p] = Live{size, charged};
  gTotal += (long long)size;
  tBalance += (long long)charged;
}

// LSW_MEM_LOG=1 prints the slow allocator events (waits, runtime calls)
bool memLog() {
  static const bool v = std::getenv("LSW_MEM_LOG") != nullptr;
  return v;
}
double secondsSince(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// returns every cached block to the runtime; the lock is held
void releaseCache() {
  const auto t0 = std::chrono::steady_clock::now();
  size_t n = 0;
  const long long bytes = gCached.load();
  (void)hipDeviceSynchronize();
  for (auto* m : {&gReady, &gPending}) {
    for (auto& kv : *m) {
      (void)hipFree(kv.second);
      gCached -= (long long)kv.first;
      ++n;
    }
    m->clear();
  }
  if (memLog())
    std::fprintf(stderr, "mem: released %zu cached blocks (%.2f GB) in %.4f s\n", n, bytes / 1e9, secondsSince(t0));
}
// frees cached blocks, the largest first, until the live and cached bytes
// plus `size` fit deviceCap(); the lock is held
void trimFor(size_t size) {
  const long long cap = deviceCap();
  auto over = [&] { return gTotal.load() + gCached.load() + (long long)size > cap; };
  if (!over()) return;
  const auto t0 = std::chrono::steady_clock::now();
  size_t n = 0;
  while (over() && !gReady.empty()) {
    auto it = std::prev(gReady.end());
    (void)hipFree(it->second);
    gCached -= (long long)it->first;
    gReady.erase(it);
    ++n;
  }
  if (over() && !gPending.empty()) {
    (void)hipDeviceSynchronize();
    for (auto& kv : gPending) {
      (void)hipFree(kv.second);
      gCached -= (long long)kv.first;
      ++n;
    }
    gPending.clear();
  }
  if (memLog())
    std::fprintf(stderr, "mem: trimmed %zu cached blocks for %.3f MB in %.4f s (live %.2f GB, cached %.2f GB)\n", n,
                 size / 1e6, secondsSince(t0), gTotal.load() / 1e9, gCached.load() / 1e9);
}
}  // namespace

hipError_t devMallocBytes(void** p, size_t bytes) {
  Lap lap(tAllocS);
  ++tAllocs;
  const size_t size = blockSize(bytes);
  std::unique_lock<std::mutex> lk(gMu);
  size_t got = 0;
  void* q = takeFrom(gReady, size, &got);
  // A small request that only a pending block fits takes a new block while
  // the cap allows: the wait for the device would wait for every stream's
  // queued work (tens of milliseconds on a device other threads keep busy).
  const bool wait = size > kWaitMin || gTotal.load() + gCached.load() + (long long)size > deviceCap();
  if (!q && wait && fitsIn(gPending, size)) {
    // kernels queued before those frees may still read the blocks: wait for
    // the device without the lock (other threads allocate and free in the
    // meantime; what they free during the wait stays pending), then make the
    // blocks freed before the wait reusable
    std::multimap<size_t, void*> synced;
    synced.swap(gPending);
    lk.unlock();
    const auto t0 = std::chrono::steady_clock::now();
    const hipError_t e = hipDeviceSynchronize();
    if (memLog()) std::fprintf(stderr, "mem: device wait for %zu pending blocks %.4f s\n", synced.size(), secondsSince(t0));
    lk.lock();
    gReady.insert(synced.begin(), synced.end());
    if (e != hipSuccess) return e;
    q = takeFrom(gReady, size, &got);
  }
  if (q) {
    gCached -= (long long)got;
    handOut(q, got, size);
    *p = q;
    return hipSuccess;
  }
  trimFor(size);
  const auto t0 = std::chrono::steady_clock::now();
  hipError_t e = hipMalloc(&q, size);
  if (memLog()) {
    const double dt = secondsSince(t0);
    if (dt > 1e-3 || e != hipSuccess)
      std::fprintf(stderr, "mem: hipMalloc %.3f MB %s in %.4f s (cached %.2f GB)\n", size / 1e6,
                   e == hipSuccess ? "ok" : "failed", dt, gCached.load() / 1e9);
  }
  if (e != hipSuccess && gCached.load() > 0) {
    (void)hipGetLastError();
    releaseCache();
    e = hipMalloc(&q, size);
  }
  if (e == hipSuccess && q) {
    handOut(q, size, size);
    *p = q;
  }
  return e;
}

hipError_t devFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> lk(gMu);
  auto it = gSizes.find(p);
  if (it == gSizes.end()) return hipFree(p);  // not ours
  const size_t size = it->second.block;
  tBalance -= (long long)it->second.charged;
  gSizes.erase(it);
  gTotal -= (long long)size;
  (tIdleFrees > 0 ? gReady : gPending).insert({size, p});
  gCached += (long long)size;
  return hipSuccess;
}

hipError_t uploadSync(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  Lap lap(tUploadS);
  ++tUploads;
  thread_local hipStream_t ts = nullptr;  // one per thread, kept for the thread's life
  if (!ts) {
    const hipError_t e = hipStreamCreateWithFlags(&ts, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
  }
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ts);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ts);
}

IdleFrees::IdleFrees() { ++tIdleFrees; }
IdleFrees::~IdleFrees() { --tIdleFrees; }

namespace {
constexpr size_t kArenaBlock = (size_t)4 << 20;
constexpr size_t kArenaAlign = 256;
}  // namespace

DevArena::~DevArena() { release(); }

void DevArena::release() {
  for (auto& b : mBlocks) (void)devFree(b.dev);
  mBlocks.clear();
  mOpen = -1;
}

void* DevArena::put(const void* src, size_t bytes) {
  if (bytes == 0) return nullptr;
  const size_t need = (bytes + kArenaAlign - 1) / kArenaAlign * kArenaAlign;
  int bi;
  if (need > kArenaBlock / 4) {
    // a large array takes a block of its own
    Block b;
    b.cap = need;
    LSW_HIP(devMalloc(&b.dev, b.cap));
    mBlocks.push_back(std::move(b));
    bi = (int)mBlocks.size() - 1;
  } else {
    if (mOpen < 0 || mBlocks[mOpen].used + need > mBlocks[mOpen].cap) {
      Block b;
      b.cap = kArenaBlock;
      LSW_HIP(devMalloc(&b.dev, b.cap));
      mBlocks.push_back(std::move(b));
      mOpen = (int)mBlocks.size() - 1;
    }
    bi = mOpen;
  }
  // the host copy holds the block's bytes from `flushed` on
  Block& b = mBlocks[bi];
  const size_t at = b.used;
  b.host.resize(at + need - b.flushed);
  std::memcpy(b.host.data() + (at - b.flushed), src, bytes);
  b.used = at + need;
  return b.dev + at;
}

void DevArena::flush() {
  for (auto& b : mBlocks) {
    if (b.used > b.flushed) LSW_HIP(uploadSync(b.dev + b.flushed, b.host.data(), b.used - b.flushed));
    b.flushed = b.used;
    std::vector<unsigned char>().swap(b.host);
  }
}

long long threadDeviceBytes() { return tBalance; }

void threadMemTimes(double* uploadS, long long* uploads, double* allocS, long long* allocs) {
  *uploadS = tUploadS;
  *uploads = tUploads;
  *allocS = tAllocS;
  *allocs = tAllocs;
}
long long totalDeviceBytes() { return gTotal.load(); }
long long cachedDeviceBytes() { return gCached.load(); }

void releaseDeviceCache() {
  std::lock_guard<std::mutex> lk(gMu);
  releaseCache();
}

}  // namespace lsw
