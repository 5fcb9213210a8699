// LocalLinearSolver (include/LocalLinearSolver.h:8-41) and its MI355X
// implementation LocalLinearSolverHIP, the drop-in beside the reference's
// LocalLinearSolverEigenSparseLU (include/LocalLinearSolverEigenSparseLU.h).
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>

#include "lsw_common.h"
#include "lsw_factor.h"

namespace lsw {

// Size of the device scratch array of the allocation-free solve: 64 MiB per
// device, for the plan's work lists plus the X, R and G solve buffers.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;

class LocalLinearSolver {
 public:
  virtual ~LocalLinearSolver() = default;
  virtual LocalLinearSolver* copy() const = 0;
  // The matrix is borrowed and must outlive the solver
  // (src/SubdomainStructured.cpp:28-29,549-562).
  virtual void setMatrix(ComplexSparseMatrix* A) = 0;
  virtual void factorize() = 0;
  virtual void solve(ComplexVector& b, ComplexVector& res) const = 0;
};

// Multifrontal LU of a 5-point structured-grid matrix on the GPU.  The grid
// (nx, ny with index i*ny+j) is recovered from the sparsity pattern; other
// patterns are rejected with LSW_ERR_UNSUPPORTED.  Copies made before
// factorize() share nothing (the reference copies unfactorised prototypes,
// src/LocalLinearSolverEigenSparseLU.cpp:9-12).
class LocalLinearSolverHIP : public LocalLinearSolver {
 public:
  explicit LocalLinearSolverHIP(int precision = LSW_FP64) : mPrecision(precision) {}
  ~LocalLinearSolverHIP() override;
  LocalLinearSolver* copy() const override { return new LocalLinearSolverHIP(mPrecision); }
  void setMatrix(ComplexSparseMatrix* A) override;
  void factorize() override;
  void solve(ComplexVector& b, ComplexVector& res) const override;
  // Multiple right-hand sides: b and x are n x nrhs column-major (ld = n),
  // host memory; synchronous.
  void solveMany(const Complex* b, Complex* x, int nrhs) const;
  // The same on device memory (FP64 complex, n x nrhs column-major), queued
  // on stream s (nullptr: the solver's own non-blocking stream, which is not
  // ordered with other streams).  b == x is allowed.
  void solveDev(const double2* b, double2* x, int nrhs, hipStream_t s) const;
  // solveDev without any device allocation: the plan's work lists and the
  // solve buffers use the library's fixed device scratch array
  // (kNoAllocScratchBytes per device, shared by every call on that device:
  // calls that overlap on one device race).  Needs a factorised solver;
  // sizes beyond the array fail with LSW_ERR_CAPACITY.
  void solveDevNoAlloc(const double2* b, double2* x, int nrhs, hipStream_t s) const;
  // scratch bytes solveDevNoAlloc needs (plan image counted when given)
  size_t noAllocBytes(int nrhs, const SolvePlan* plan) const;
  int precision() const { return mPrecision; }
  void setPrecision(int p) { mPrecision = p; }
  int gridNx() const { return mNx; }
  int gridNy() const { return mNy; }
  const FactorSet* factorSet() const { return mFS.get(); }

 private:
  int mPrecision;
  ComplexSparseMatrix* mA = nullptr;
  int mNx = 0, mNy = 0;
  std::unique_ptr<FactorSet> mFS;
  hipStream_t mStream = nullptr;
  // solve scratch for up to mCap right-hand sides, and the launch plan per
  // right-hand-side count
  void ensureScratch(int nrhs) const;
  const SolvePlan& planFor(int nrhs) const;
  mutable void* dX = nullptr;
  mutable void* dR = nullptr;
  mutable void* dG = nullptr;
  mutable void* dB = nullptr;
  mutable int mCap = 0;
  mutable std::map<int, std::unique_ptr<SolvePlan>> mPlans;
  mutable std::map<int, std::unique_ptr<SolvePlan>> mPlansNoAlloc;  // host images only
};

}  // namespace lsw
