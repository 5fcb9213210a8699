// GlobalIterativeSolver (include/GlobalIterativeSolver.h:16-49) and its
// MI355X implementation GlobalIterativeSolverGMRESHIP, the drop-in beside
// GlobalIterativeSolverGMRES (include/GlobalIterativeSolverGMRES.h).
//
// Krylov vectors live on the device as global grid functions (FP64).  Dots and
// norms are two-pass block reductions in a fixed order (bit-reproducible);
// Arnoldi coefficients stay on the device between the dot and the update, and
// one small read-back per iteration feeds the Givens update on the host.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "lsw_dd.h"

namespace lsw {

struct GmresInfo {
  int iterations = 0;            // Krylov steps taken
  int preconditionerApplies = 0;
  int matvecs = 0;
  double estimate = 0;           // last logged residual estimate
  double trueResidual = 0;       // ||b - A x|| / ||b|| (FP64, global operator)
  std::vector<double> log;       // logged estimates per iteration
};

class GlobalIterativeSolver {
 public:
  virtual ~GlobalIterativeSolver() = default;
  void setParent(const DD* dd) { mParent = dd; }
  virtual GlobalIterativeSolver* copy() const = 0;
  virtual void solve(const std::vector<ComplexVector>& rhs, std::vector<ComplexVector>& res) const = 0;

 protected:
  const DD* mParent = nullptr;
};

class GlobalIterativeSolverGMRESHIP : public GlobalIterativeSolver {
 public:
  // mode 0 ("reference"): the reference's left-preconditioned, non-restarted
  //   GMRES with its arithmetic (x0 = M^{-1} b, conjugated Arnoldi
  //   coefficients, the skipped rotation, complex-symmetric rotations;
  //   src/GlobalIterativeSolverGMRES.cpp:65-348).  tol applies to its logged
  //   estimate, maxIter is a hard cap.
  // mode 1 ("flexible"): right-preconditioned flexible GMRES(restart) with
  //   Hermitian modified Gram-Schmidt and unitary rotations; tol applies to the
  //   residual estimate of the unpreconditioned system; used with
  //   mixed-precision preconditioners.
  enum Mode { kReference = 0, kFlexible = 1 };

  GlobalIterativeSolverGMRESHIP() = default;
  ~GlobalIterativeSolverGMRESHIP() override;
  GlobalIterativeSolver* copy() const override;
  void setTol(double t) { mTol = t; }
  void setMaxIter(int m) { mMaxIter = m; }
  void setRestart(int r) { mRestart = r; }
  void setMode(int m) { mMode = m; }
  void setVerbose(bool v) { mVerbose = v; }
  double getTol() const { return mTol; }
  int getMaxIter() const { return mMaxIter; }
  int getMode() const { return mMode; }

  void solve(const std::vector<ComplexVector>& rhs, std::vector<ComplexVector>& res) const override;
  // Device API: b and x are global grid vectors on the device.
  void solveDev(const double2* b, double2* x, GmresInfo* info, hipStream_t s = nullptr) const;
  // Flexible GMRES(restart) for nrhs right-hand sides (B, X: nrhs consecutive
  // global vectors).  The active right-hand sides advance in lockstep so that
  // every Arnoldi step applies the preconditioner once to the whole block;
  // orthogonalisation is classical Gram-Schmidt with one reorthogonalisation
  // (two batched passes), and each right-hand side stops at its own tolerance.
  void solveBlockDev(const double2* B, double2* X, int nrhs, std::vector<GmresInfo>* infos,
                     hipStream_t s = nullptr) const;
  const GmresInfo& lastInfo() const { return mInfo; }

 private:
  void ensure(long long n, int vecs) const;
  void solveReference(const DDCheckerboard& dd, const double2* b, double2* x, GmresInfo& info,
                      hipStream_t s) const;
  void solveFlexible(const DDCheckerboard& dd, const double2* b, double2* x, GmresInfo& info,
                     hipStream_t s) const;

  double mTol = 1e-6;     // src/Input.cpp:367-385 default
  int mMaxIter = 40;      // src/GlobalIterativeSolverGMRES.cpp:9
  int mRestart = 30;
  int mMode = kReference;
  bool mVerbose = false;
  mutable GmresInfo mInfo;
  // device workspace
  mutable double2* dV = nullptr;   // basis vectors
  mutable double2* dZ = nullptr;   // preconditioned basis (flexible mode)
  mutable double2* dT = nullptr;   // temporaries (3 vectors)
  mutable double2* dPart = nullptr;
  mutable double2* dScal = nullptr;
  mutable long long mN = 0;
  mutable int mVecs = 0;
  // block workspace
  mutable double2* dBV = nullptr;
  mutable double2* dBZ = nullptr;
  mutable double2* dBT = nullptr;
  mutable double2* dBPart = nullptr;
  mutable double2* dBH = nullptr;
  mutable int* dBLive = nullptr;  // the live columns of an inner step
  mutable long long mBN = 0;
  mutable int mBRhs = 0, mBM = 0;
};

// device BLAS-1 helpers (FP64 complex), exposed for tests and benchmarks
namespace blas {
void dotc(const double2* a, const double2* b, long long n, double2* part, double2* out, hipStream_t s);
void norm2sq(const double2* a, long long n, double2* part, double2* out, hipStream_t s);
}  // namespace blas

}  // namespace lsw
