// Recorded device allocations (see lsw_device.h).
#include <atomic>
#include <mutex>
#include <unordered_map>

#include "lsw_device.h"

namespace lsw {

namespace {
std::mutex gMu;
std::unordered_map<void*, size_t> gSizes;
std::atomic<long long> gTotal{0};
thread_local long long tBalance = 0;
}  // namespace

hipError_t devMallocBytes(void** p, size_t bytes) {
  const hipError_t e = hipMalloc(p, bytes);
  if (e == hipSuccess && *p) {
    std::lock_guard<std::mutex> lk(gMu);
    gSizes[*p] = bytes;
    gTotal += (long long)bytes;
    tBalance += (long long)bytes;
  }
  return e;
}

hipError_t devFree(void* p) {
  if (!p) return hipSuccess;
  {
    std::lock_guard<std::mutex> lk(gMu);
    auto it = gSizes.find(p);
    if (it != gSizes.end()) {
      gTotal -= (long long)it->second;
      tBalance -= (long long)it->second;
      gSizes.erase(it);
    }
  }
  return hipFree(p);
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

}  // namespace lsw
