// Device-side descriptors and HIP helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "lsw_common.h"
#include "lsw_nd.h"

namespace lsw {

#define LSW_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess)                                                               \
      throw ::lsw::Error(LSW_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

// Device memory of the library.  Every allocation is recorded with its size,
// and each thread keeps the balance of what it allocated minus what it freed:
// the server charges what a worker allocates while it serves a problem to
// that problem.  Freed blocks are cached for reuse (no hipFree, which waits
// for the whole device); a cached block is reused only after a device
// synchronisation that follows its free, so kernels queued before the free
// on any stream have finished with it.
hipError_t devMallocBytes(void** p, size_t bytes);
hipError_t devFree(void* p);
// While an IdleFrees object lives on a thread, that thread's devFree calls
// return blocks that no queued device work can touch any more (the caller
// has synchronised every stream that used them): they are reusable at once,
// without the device-wide wait a plain free needs.
class IdleFrees {
 public:
  IdleFrees();
  ~IdleFrees();
  IdleFrees(const IdleFrees&) = delete;
  IdleFrees& operator=(const IdleFrees&) = delete;
};
// bytes held in the block cache; return them to the runtime
long long cachedDeviceBytes();
void releaseDeviceCache();
// Copy host memory to the device and return once the data is in device
// memory (an asynchronous copy on the calling thread's own non-blocking
// transfer stream, then a wait on that stream): kernels launched afterwards
// on any stream see it.  A plain hipMemcpy from pageable memory may return
// as soon as the data is staged, and only kernels on blocking streams are
// ordered behind it.
hipError_t uploadSync(void* dst, const void* src, size_t bytes);
long long threadDeviceBytes();
long long totalDeviceBytes();
// this thread's running totals of uploadSync and devMallocBytes calls and
// the seconds spent in them
void threadMemTimes(double* uploadS, long long* uploads, double* allocS, long long* allocs);
template <typename T>
inline hipError_t devMalloc(T** p, size_t bytes) {
  return devMallocBytes(reinterpret_cast<void**>(p), bytes);
}

// The small device arrays of one owner (tree descriptors, index pools, step
// slot lists), sub-allocated from a few blocks and uploaded together: put()
// returns an array's device address at once and keeps its bytes on the host,
// flush() uploads everything put since the last flush with one uploadSync per
// block.  Set-up then costs a handful of allocations and waits for the device
// instead of one per array.  Nothing put may be read on the device before the
// next flush(); the blocks are freed with the arena.
class DevArena {
 public:
  DevArena() = default;
  DevArena(const DevArena&) = delete;
  DevArena& operator=(const DevArena&) = delete;
  ~DevArena();
  void* put(const void* src, size_t bytes);
  template <typename T>
  T* put(const std::vector<T>& v) {
    return v.empty() ? nullptr : static_cast<T*>(put(v.data(), v.size() * sizeof(T)));
  }
  void flush();
  void release();  // frees every block (arrays put earlier become invalid)

 private:
  struct Block {
    unsigned char* dev = nullptr;
    size_t cap = 0, used = 0, flushed = 0;
    std::vector<unsigned char> host;
  };
  std::vector<Block> mBlocks;
  int mOpen = -1;  // block that takes small arrays
};

constexpr int kMaxHeight = 40;

// One local problem (a checkerboard cell, or the single matrix of a
// LocalLinearSolverHIP) as the kernels see it.
struct CellDev {
  int nx, ny, ox, oy;      // local DOF grid, offset of local (0,0) in the global grid
  int vlr[2][8];           // LR 7-set starts [shifted][set]; [.][7] = nx
  int vud[2][8];           // UD 7-set starts; [.][7] = ny
  int shape, pad;
  long long coefOff;       // 5*nx*ny complex stencil entries (dir-major)
  long long factorOff;     // factor blob offset (complex entries)
  long long schurOff;      // Schur workspace offset (factorisation only)
  long long mOff;          // squared slowness per DOF
  double x0, y0, hx, hy;   // DOF (i,j) at (x0 + (i+1) hx, y0 + (j+1) hy)
  double intDom[2][2], compDom[2][2];
  double C, omega;
};

// ND structure of one shape class (device pointers).
struct ShapeDev {
  const NdNodeDev* nodes;
  const int* E;
  const int* B;
  const int* cmap;
  const int* up;
  const AsmEntry* asmList;
  const SolveItem* fwdItems;
  const SolveItem* bwdItems;
  int fwdStart[kMaxHeight], fwdCount[kMaxHeight];
  int bwdStart[kMaxHeight], bwdCount[kMaxHeight];
  int maxHeight, nodeCount;
};

// A batched solve task: up to kMaxRhs right-hand sides of one cell.
constexpr int kMaxRhs = 8;
// tasks of up to 4 right-hand sides always fit the solve kernels' LDS staging;
// 8 when the widest front allows it (FactorSet::maxTaskRhs)
constexpr int kTaskRhs4 = 4;
// right-hand sides per task of the solve kernel that serves tasks of up to n
constexpr int taskWidth(int n) { return n <= 1 ? 1 : n <= kTaskRhs4 ? kTaskRhs4 : kMaxRhs; }
struct SolveTask {
  int cell;
  int nrhs;
  int slot[kMaxRhs];
};

// One front of the multifrontal factorisation.
struct FrontTask {
  int cell, node;
  long long wsOff;  // dense workspace offset (complex entries)
};

}  // namespace lsw
