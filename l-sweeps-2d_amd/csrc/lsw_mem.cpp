// Recorded device allocations with a block cache (see lsw_device.h).
//
// hipFree waits for the whole device to go idle, and a server whose workers
// load and evict problems while other workers keep the device busy would
// spend its time in those waits.  Freed blocks are therefore kept:
//   * a freed block first goes to the pending list: kernels queued before
//     the free (on any stream) may still use it;
//   * a request that no ready block fits, but a pending one would, waits
//     once for the device (hipDeviceSynchronize) and moves every pending
//     block to the ready list;
//   * a ready block serves a request of at most its size and at least half
//     of it (best fit), so a cached block wastes at most half its bytes;
//   * when hipMalloc fails, the cache is returned to the runtime (after a
//     device synchronisation) and the allocation is tried once more;
//   * when the library holds no block any more, the cache is returned too.
// The per-thread balance (what a thread allocated minus what it freed) counts
// block sizes, so the server charges a problem the device memory it holds.
#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <unordered_map>

#include "lsw_device.h"

namespace lsw {

namespace {
std::mutex gMu;
std::unordered_map<void*, size_t> gSizes;     // live blocks (handed out) -> block size
std::multimap<size_t, void*> gReady;          // cached blocks safe to reuse, by size
std::multimap<size_t, void*> gPending;        // cached blocks freed since the last device sync
std::atomic<long long> gTotal{0};             // bytes handed out
std::atomic<long long> gCached{0};            // bytes held in the cache
thread_local long long tBalance = 0;

constexpr size_t kGrain = 256;

size_t blockSize(size_t bytes) {
  // 256-byte grain below 2 MiB, 2 MiB pages above
  if (bytes < ((size_t)2 << 20)) return std::max(kGrain, (bytes + kGrain - 1) / kGrain * kGrain);
  const size_t page = (size_t)2 << 20;
  return (bytes + page - 1) / page * page;
}

// best fit among cached blocks of [size, 2 size]; the lock is held
void* takeFrom(std::multimap<size_t, void*>& m, size_t size, size_t* got) {
  auto it = m.lower_bound(size);
  if (it == m.end() || it->first > 2 * size) return nullptr;
  void* p = it->second;
  *got = it->first;
  m.erase(it);
  return p;
}

bool fitsIn(const std::multimap<size_t, void*>& m, size_t size) {
  auto it = m.lower_bound(size);
  return it != m.end() && it->first <= 2 * size;
}

void handOut(void* p, size_t size) {
  gSizes[p] = size;
  gTotal += (long long)size;
  tBalance += (long long)size;
}

// returns every cached block to the runtime; the lock is held
void releaseCache() {
  (void)hipDeviceSynchronize();
  for (auto* m : {&gReady, &gPending}) {
    for (auto& kv : *m) {
      (void)hipFree(kv.second);
      gCached -= (long long)kv.first;
    }
    m->clear();
  }
}
}  // namespace

hipError_t devMallocBytes(void** p, size_t bytes) {
  const size_t size = blockSize(bytes);
  std::lock_guard<std::mutex> lk(gMu);
  size_t got = 0;
  void* q = takeFrom(gReady, size, &got);
  if (!q && fitsIn(gPending, size)) {
    // kernels queued before those frees may still read the blocks
    const hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return e;
    gReady.insert(gPending.begin(), gPending.end());
    gPending.clear();
    q = takeFrom(gReady, size, &got);
  }
  if (q) {
    gCached -= (long long)got;
    handOut(q, got);
    *p = q;
    return hipSuccess;
  }
  hipError_t e = hipMalloc(&q, size);
  if (e != hipSuccess && gCached.load() > 0) {
    (void)hipGetLastError();
    releaseCache();
    e = hipMalloc(&q, size);
  }
  if (e == hipSuccess && q) {
    handOut(q, size);
    *p = q;
  }
  return e;
}

hipError_t devFree(void* p) {
  if (!p) return hipSuccess;
  std::lock_guard<std::mutex> lk(gMu);
  auto it = gSizes.find(p);
  if (it == gSizes.end()) return hipFree(p);  // not ours
  const size_t size = it->second;
  gSizes.erase(it);
  gTotal -= (long long)size;
  tBalance -= (long long)size;
  gPending.insert({size, p});
  gCached += (long long)size;
  // nothing of the library's is left on the device: give the cache back, so
  // that other users of the device (the caller's own allocator) find it free
  if (gTotal.load() == 0) releaseCache();
  return hipSuccess;
}

hipError_t uploadSync(void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return hipSuccess;
  thread_local hipStream_t ts = nullptr;  // one per thread, kept for the thread's life
  if (!ts) {
    const hipError_t e = hipStreamCreateWithFlags(&ts, hipStreamNonBlocking);
    if (e != hipSuccess) return e;
  }
  const hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ts);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(ts);
}

long long threadDeviceBytes() { return tBalance; }
long long totalDeviceBytes() { return gTotal.load(); }
long long cachedDeviceBytes() { return gCached.load(); }

void releaseDeviceCache() {
  std::lock_guard<std::mutex> lk(gMu);
  releaseCache();
}

}  // namespace lsw
