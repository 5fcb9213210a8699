// LocalLinearSolverHIP (see lsw_solver.h).
#include "lsw_solver.h"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <string>

namespace lsw {

LocalLinearSolverHIP::~LocalLinearSolverHIP() {
  if (mStream) (void)hipStreamSynchronize(mStream);
  mPlans.clear();
  mPlansNoAlloc.clear();
  for (void* p : {dX, dR, dG, dB}) if (p) (void)devFree(p);
  mFS.reset();
  if (mStream) (void)hipStreamDestroy(mStream);
}

// Recover the nx x ny grid of a 5-point matrix with index i*ny+j and check the
// whole pattern (LocalLinearSolverEigenSparseLU accepts any sparse matrix; the
// MI355X solver is specialised to the structured operators the assemblers
// produce).
void LocalLinearSolverHIP::setMatrix(ComplexSparseMatrix* A) {
  LSW_REQUIRE(A != nullptr, LSW_ERR_ARG, "setMatrix: null matrix");
  LSW_REQUIRE(A->rows == A->cols && A->rows > 0, LSW_ERR_ARG, "setMatrix: matrix must be square");
  LSW_REQUIRE((int)A->colptr.size() == A->cols + 1, LSW_ERR_ARG, "setMatrix: bad column pointer");
  const int N = A->rows;
  int ny = N;  // 1 x N grid unless column 0 couples to a far row
  for (int k = A->colptr[0]; k < A->colptr[1]; ++k) {
    const int r = A->rowind[k];
    if (r > 1) ny = std::min(ny == N ? r : ny, r);
  }
  LSW_REQUIRE(N % ny == 0, LSW_ERR_UNSUPPORTED, "matrix is not a 5-point structured-grid operator");
  const int nx = N / ny;
  for (int d = 0; d < N; ++d) {
    const int i = d / ny, j = d % ny;
    for (int k = A->colptr[d]; k < A->colptr[d + 1]; ++k) {
      const int q = A->rowind[k];
      LSW_REQUIRE(q >= 0 && q < N, LSW_ERR_ARG, "setMatrix: row index out of range");
      const int diff = d - q;
      const int qi = q / ny, qj = q % ny;
      const bool ok = diff == 0 || ((diff == 1 || diff == -1) && qi == i && std::abs(qj - j) == 1) ||
                      ((diff == ny || diff == -ny) && qj == j && ny < N);
      LSW_REQUIRE(ok, LSW_ERR_UNSUPPORTED, "matrix is not a 5-point structured-grid operator");
    }
  }
  mA = A;
  mNx = nx;
  mNy = ny;
  mPlans.clear();
  mPlansNoAlloc.clear();
  for (void* p : {dX, dR, dG, dB}) if (p) (void)devFree(p);
  dX = dR = dG = dB = nullptr;
  mCap = 0;
  mFS.reset();
}

void LocalLinearSolverHIP::factorize() {
  LSW_REQUIRE(mA != nullptr, LSW_ERR_STATE, "factorize called before setMatrix");
  if (mFS && mFS->factorized()) return;  // idempotent (src/LocalLinearSolverEigenSparseLU.cpp:42-45)
  if (!mStream) LSW_HIP(hipStreamCreateWithFlags(&mStream, hipStreamNonBlocking));
  const int N = mA->rows, ny = mNy;
  std::vector<double2> coef((size_t)5 * N, make_double2(0.0, 0.0));
  for (int d = 0; d < N; ++d)
    for (int k = mA->colptr[d]; k < mA->colptr[d + 1]; ++k) {
      const int q = mA->rowind[k];
      const int diff = d - q;  // entry A[q][d]: coefficient of row q towards d
      const int dir = diff == 0 ? 0 : diff == -ny ? 1 : diff == ny ? 2 : diff == -1 ? 3 : 4;
      coef[(size_t)dir * N + q] = make_double2(mA->vals[k].real(), mA->vals[k].imag());
    }
  CellDev c;
  std::memset(&c, 0, sizeof(c));
  c.nx = mNx;
  c.ny = mNy;
  mFS = std::make_unique<FactorSet>();
  mFS->addCellWithStencil(c, coef);
  mFS->setup(mStream);
  mFS->factorize(mPrecision, mStream);
}

namespace {

// Right-hand sides b (n x nrhs column-major, FP64) into the task-major solve
// buffer: value of right-hand side r at position p goes to task r / per,
// lane r % per, i.e. X[((r / per) * n + p) * W + r % per].
template <typename T>
__global__ void k_lu_pack(const double2* __restrict__ b, T* __restrict__ X, long long n, int nrhs, int per, int W) {
  const long long total = n * nrhs;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const long long r = q / n, p = q - r * n;
    const double2 v = b[q];
    T o;
    o.x = v.x;
    o.y = v.y;
    X[((r / per) * n + p) * W + (r % per)] = o;
  }
}

template <typename T>
__global__ void k_lu_unpack(const T* __restrict__ X, double2* __restrict__ x, long long n, int nrhs, int per, int W) {
  const long long total = n * nrhs;
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (long long)gridDim.x * blockDim.x) {
    const long long r = q / n, p = q - r * n;
    const T v = X[((r / per) * n + p) * W + (r % per)];
    x[q] = make_double2((double)v.x, (double)v.y);
  }
}

inline dim3 gridFor(long long n) { return dim3((unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 8192))); }

}  // namespace

void LocalLinearSolverHIP::ensureScratch(int nrhs) const {
  if (nrhs <= mCap) return;
  const long long N = mA->rows;
  const long long G = std::max<long long>(mFS->maxG(), 1);
  const size_t es = mPrecision == LSW_FP64 ? sizeof(double2) : sizeof(float2);
  const int per = mFS->maxTaskRhs();
  const int ntasks = (nrhs + per - 1) / per;
  for (void* p : {dX, dR, dG, dB}) if (p) (void)devFree(p);
  dX = dR = dG = dB = nullptr;
  mPlans.clear();
  // X, R and G are task-major: ntasks rows of up to per right-hand sides
  const size_t rows = (size_t)ntasks * per;
  LSW_HIP(devMalloc(&dX, rows * N * es));
  LSW_HIP(devMalloc(&dR, rows * N * es));
  LSW_HIP(devMalloc(&dG, rows * G * es));
  LSW_HIP(devMalloc(&dB, (size_t)nrhs * N * sizeof(double2)));  // host-pointer staging
  // the solve kernels read (and multiply by zero) entries of X, R and G
  // they do not use: they must hold finite values from the start.  The
  // solves may run on a caller's stream, so the clears finish here.
  LSW_HIP(hipMemsetAsync(dX, 0, rows * N * es, mStream));
  LSW_HIP(hipMemsetAsync(dR, 0, rows * N * es, mStream));
  LSW_HIP(hipMemsetAsync(dG, 0, rows * G * es, mStream));
  LSW_HIP(hipStreamSynchronize(mStream));
  mCap = nrhs;
}

const SolvePlan& LocalLinearSolverHIP::planFor(int nrhs) const {
  auto it = mPlans.find(nrhs);
  if (it != mPlans.end()) return *it->second;
  const int per = mFS->maxTaskRhs();
  const int ntasks = (nrhs + per - 1) / per;
  std::vector<SolveTask> tasks(ntasks);
  for (int t = 0; t < ntasks; ++t) {
    std::memset(&tasks[t], 0, sizeof(SolveTask));
    tasks[t].cell = 0;
    tasks[t].nrhs = std::min(per, nrhs - t * per);
    for (int r = 0; r < tasks[t].nrhs; ++r) tasks[t].slot[r] = t * per + r;
  }
  auto plan = std::make_unique<SolvePlan>();
  mFS->buildPlan(tasks, {}, *plan);
  LSW_REQUIRE(plan->taskRows() <= (long long)((mCap + per - 1) / per) * per, LSW_ERR_STATE,
              "solve scratch smaller than the plan needs");
  LSW_REQUIRE(plan->width == per || nrhs < per, LSW_ERR_INTERNAL, "unexpected solve task width");
  return *mPlans.emplace(nrhs, std::move(plan)).first->second;
}

void LocalLinearSolverHIP::solveDev(const double2* b, double2* x, int nrhs, hipStream_t s) const {
  LSW_REQUIRE(mFS && mFS->factorized(), LSW_ERR_STATE, "Solve called before factorization");
  LSW_REQUIRE(nrhs >= 0, LSW_ERR_ARG, "negative right-hand-side count");
  if (nrhs == 0) return;
  if (!s) s = mStream;
  ensureScratch(nrhs);
  const SolvePlan& plan = planFor(nrhs);
  const long long N = mA->rows;
  const long long G = std::max<long long>(mFS->maxG(), 1);
  // right-hand side r = task r / per, lane r % per of the task-major buffer
  const int per = mFS->maxTaskRhs(), W = plan.width;
  const dim3 grid = gridFor(N * nrhs);
  if (mPrecision == LSW_FP64) {
    hipLaunchKernelGGL(k_lu_pack<double2>, grid, dim3(256), 0, s, b, (double2*)dX, N, nrhs, per, W);
    mFS->solve(plan, dX, dR, dG, N, G, s);
    hipLaunchKernelGGL(k_lu_unpack<double2>, grid, dim3(256), 0, s, (const double2*)dX, x, N, nrhs, per, W);
  } else {
    hipLaunchKernelGGL(k_lu_pack<float2>, grid, dim3(256), 0, s, b, (float2*)dX, N, nrhs, per, W);
    mFS->solve(plan, dX, dR, dG, N, G, s);
    hipLaunchKernelGGL(k_lu_unpack<float2>, grid, dim3(256), 0, s, (const float2*)dX, x, N, nrhs, per, W);
  }
  LSW_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Solves that allocate nothing.  Their scratch (the plan's work lists and the
// task-major X, R, G buffers) lives in one file-scope device array of
// kNoAllocScratchBytes; each device holds its own copy until the process
// ends.  Host code reaches a device's copy only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host-side
// stand-in).  Every call on a device shares that device's array and leaves
// its contents behind, so two such calls that overlap on one device race.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
std::mutex gScratchMu;
std::map<int, void*> gScratchOf;  // device -> address of its copy of scratch

void* deviceScratch() {
  int dev = 0;
  LSW_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(gScratchMu);
  auto it = gScratchOf.find(dev);
  if (it != gScratchOf.end()) return it->second;
  void* ptr = nullptr;
  LSW_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  LSW_REQUIRE(ptr != nullptr, LSW_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
  gScratchOf[dev] = ptr;
  return ptr;
}

__global__ void k_clear_words(unsigned long long* __restrict__ p, long long words) {
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < words;
       q += (long long)gridDim.x * blockDim.x)
    p[q] = 0ull;
}

inline size_t up256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

size_t LocalLinearSolverHIP::noAllocBytes(int nrhs, const SolvePlan* plan) const {
  const long long N = mA->rows;
  const long long G = std::max<long long>(mFS->maxG(), 1);
  const size_t es = mPrecision == LSW_FP64 ? sizeof(double2) : sizeof(float2);
  const int per = mFS->maxTaskRhs();
  const size_t rows = (size_t)((nrhs + per - 1) / per) * per;
  const size_t blob = plan ? up256(plan->hBlob.size()) : 0;
  return blob + 2 * up256(rows * N * es) + up256(rows * G * es);
}

void LocalLinearSolverHIP::solveDevNoAlloc(const double2* b, double2* x, int nrhs, hipStream_t s) const {
  LSW_REQUIRE(mFS && mFS->factorized(), LSW_ERR_STATE, "Solve called before factorization");
  LSW_REQUIRE(nrhs >= 1, LSW_ERR_ARG, "the right-hand-side count must be positive");
  LSW_REQUIRE(noAllocBytes(nrhs, nullptr) <= kNoAllocScratchBytes, LSW_ERR_CAPACITY,
              "solve needs " + std::to_string(noAllocBytes(nrhs, nullptr)) + " bytes of scratch, more than the " +
                  std::to_string(kNoAllocScratchBytes) + " the allocation-free entry point holds");
  if (!s) s = mStream;
  void* base = deviceScratch();
  // the plan's host image, built once per right-hand-side count (host memory only)
  auto it = mPlansNoAlloc.find(nrhs);
  if (it == mPlansNoAlloc.end()) {
    const int per = mFS->maxTaskRhs();
    const int ntasks = (nrhs + per - 1) / per;
    std::vector<SolveTask> tasks(ntasks);
    for (int t = 0; t < ntasks; ++t) {
      std::memset(&tasks[t], 0, sizeof(SolveTask));
      tasks[t].nrhs = std::min(per, nrhs - t * per);
      for (int r = 0; r < tasks[t].nrhs; ++r) tasks[t].slot[r] = t * per + r;
    }
    auto plan = std::make_unique<SolvePlan>();
    plan->hostOnly = true;
    mFS->buildPlan(tasks, {}, *plan);
    it = mPlansNoAlloc.emplace(nrhs, std::move(plan)).first;
  }
  SolvePlan& plan = *it->second;
  LSW_REQUIRE(noAllocBytes(nrhs, &plan) <= kNoAllocScratchBytes, LSW_ERR_CAPACITY,
              "solve needs " + std::to_string(noAllocBytes(nrhs, &plan)) + " bytes of scratch, more than the " +
                  std::to_string(kNoAllocScratchBytes) + " the allocation-free entry point holds");
  const long long N = mA->rows;
  const long long G = std::max<long long>(mFS->maxG(), 1);
  const size_t es = mPrecision == LSW_FP64 ? sizeof(double2) : sizeof(float2);
  const int per = mFS->maxTaskRhs(), W = plan.width;
  const size_t rows = (size_t)((nrhs + per - 1) / per) * per;
  LSW_REQUIRE((size_t)plan.taskRows() <= rows, LSW_ERR_INTERNAL, "plan needs more rows than the scratch holds");
  // scratch layout: plan image | X | R | G
  auto* p = static_cast<unsigned char*>(base);
  const size_t oX = up256(plan.hBlob.size());
  const size_t oR = oX + up256(rows * N * es), oG = oR + up256(rows * N * es);
  const size_t end = oG + up256(rows * G * es);
  void* dX = p + oX;
  void* dR = p + oR;
  void* dG = p + oG;
  LSW_HIP(hipMemcpyAsync(p, plan.hBlob.data(), plan.hBlob.size(), hipMemcpyHostToDevice, s));
  plan.bind(p);
  // X, R and G start at zero: the solve kernels read (and multiply by zero)
  // entries they do not use, and the array holds whatever the last call left
  const long long words = (long long)((end - oX) / 8);
  hipLaunchKernelGGL(k_clear_words, gridFor(words), dim3(256), 0, s, (unsigned long long*)dX, words);
  const dim3 grid = gridFor(N * nrhs);
  if (mPrecision == LSW_FP64) {
    hipLaunchKernelGGL(k_lu_pack<double2>, grid, dim3(256), 0, s, b, (double2*)dX, N, nrhs, per, W);
    mFS->solve(plan, dX, dR, dG, N, G, s);
    hipLaunchKernelGGL(k_lu_unpack<double2>, grid, dim3(256), 0, s, (const double2*)dX, x, N, nrhs, per, W);
  } else {
    hipLaunchKernelGGL(k_lu_pack<float2>, grid, dim3(256), 0, s, b, (float2*)dX, N, nrhs, per, W);
    mFS->solve(plan, dX, dR, dG, N, G, s);
    hipLaunchKernelGGL(k_lu_unpack<float2>, grid, dim3(256), 0, s, (const float2*)dX, x, N, nrhs, per, W);
  }
  LSW_HIP(hipGetLastError());
}

void LocalLinearSolverHIP::solveMany(const Complex* b, Complex* x, int nrhs) const {
  LSW_REQUIRE(mFS && mFS->factorized(), LSW_ERR_STATE, "Solve called before factorization");
  LSW_REQUIRE(nrhs >= 0, LSW_ERR_ARG, "negative right-hand-side count");
  if (nrhs == 0) return;
  ensureScratch(nrhs);
  const size_t bytes = (size_t)nrhs * mA->rows * sizeof(double2);
  LSW_HIP(uploadSync(dB, b, bytes));
  // in place: the staging buffer receives the solutions
  solveDev(static_cast<const double2*>(dB), static_cast<double2*>(dB), nrhs, mStream);
  LSW_HIP(hipMemcpyAsync(x, dB, bytes, hipMemcpyDeviceToHost, mStream));
  LSW_HIP(hipStreamSynchronize(mStream));
}

void LocalLinearSolverHIP::solve(ComplexVector& b, ComplexVector& res) const {
  LSW_REQUIRE(mA && (int)b.size() == mA->rows, LSW_ERR_ARG, "solve: right-hand side has the wrong length");
  res.resize(b.size());
  solveMany(b.data(), res.data(), 1);
}

}  // namespace lsw
