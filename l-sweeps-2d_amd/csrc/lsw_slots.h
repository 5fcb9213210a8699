// Iteration-level batching of flexible GMRES over one factorized DD.
//
// A GmresSlots object owns `capacity` slots.  Each slot holds one right-hand
// side with its own Krylov basis (restart length m), its own Arnoldi step
// counter and its own Givens state, so right-hand sides can join and leave
// between steps: step() advances every active slot by one Arnoldi step and
// applies the preconditioner and the operator to all of them as one block
// (one read of the factors per local solve serves up to kMaxRhs slots).
// A slot that reaches the tolerance on its Arnoldi estimate, or the end of a
// cycle, forms x += Z y, checks the true FP64 residual ||b - A x|| / ||b||
// and either finishes or restarts from the true residual.
//
// There is no counterpart in the reference, which solves one summed
// right-hand side per problem (src/Problem.cpp:94-158).  The arithmetic per
// slot is the flexible right-preconditioned GMRES of the block solver
// (classical Gram-Schmidt with one reorthogonalisation, unitary rotations).
#pragma once

#include <hip/hip_runtime.h>

#include <complex>
#include <vector>

#include "lsw_dd.h"

namespace lsw {

constexpr int kMaxSlots = 64;

// Block widths the preconditioner is applied with: the active slots of a step
// are padded with zero vectors up to the next width, so that a problem holds
// solve plans for at most a dozen widths (a third more work at most).
int planWidth(int nactive);

class GmresSlots {
 public:
  struct Done {
    int slot;
    long long tag;
    int iterations;
    double trueResidual;
    double estimate;
  };

  GmresSlots(const DDCheckerboard* dd, int capacity, int restart, double tol, int maxIter);
  ~GmresSlots();
  GmresSlots(const GmresSlots&) = delete;
  GmresSlots& operator=(const GmresSlots&) = delete;

  // device bytes held for a grid of n unknowns
  static size_t bytesFor(long long n, int capacity, int restart);
  size_t bytes() const { return bytesFor(mN, mCap, mM); }

  int capacity() const { return mCap; }
  int active() const;
  int freeSlots() const;
  hipStream_t stream() const { return mDD->stream(); }

  // Admits a right-hand side (device global vector, copied) into a free slot
  // and returns the slot index.  A zero right-hand side finishes at once with
  // x = 0.
  int admit(const double2* b, long long tag);
  // Device buffer of n entries into which the caller may write a right-hand
  // side before admit() (valid until the next admit/step).
  double2* staging() const { return dStage; }
  // One Arnoldi step of every active slot.  Slots that finish are appended to
  // `done`; they keep their solution until release().
  void step(std::vector<Done>& done);
  const double2* solution(int slot) const { return dX + (long long)slot * mN; }
  void release(int slot);
  long long steps() const { return mSteps; }
  long long blockApplies() const { return mBlockApplies; }

 private:
  enum State { kFree = 0, kActive = 1, kFinished = 2 };
  struct Slot {
    State state = kFree;
    long long tag = 0;
    int j = 0;            // Arnoldi step within the cycle
    int iterations = 0;   // total Arnoldi steps
    double bnorm = 0;
    double estimate = 0;
    bool zeroPending = false;  // zero right-hand side not yet reported
    std::vector<std::complex<double>> H;  // (m+1) x m, column-major
    std::vector<std::complex<double>> cs, sn, g;
  };
  double2* V(int slot, int i) const { return dV + ((long long)slot * (mM + 1) + i) * mN; }
  double2* Z(int slot, int i) const { return dZ + ((long long)slot * mM + i) * mN; }
  // x += Z y for the slot's finished cycle, true residual, finish or restart
  void endCycle(int slot, std::vector<Done>& done);
  double norm(const double2* v) const;
  void startCycle(int slot, const double2* r, double rnorm);

  const DDCheckerboard* mDD;
  int mCap, mM, mMaxIter;
  double mTol;
  long long mN;
  std::vector<Slot> mSlots;
  long long mSteps = 0, mBlockApplies = 0;
  // device storage
  double2* dV = nullptr;      // capacity x (m+1) basis vectors
  double2* dZ = nullptr;      // capacity x m preconditioned vectors
  double2* dB = nullptr;      // capacity right-hand sides
  double2* dX = nullptr;      // capacity solutions
  double2* dIn = nullptr;     // block of active V_j (capacity vectors)
  double2* dPZ = nullptr;     // block of M^{-1} V_j
  double2* dW = nullptr;      // block of A M^{-1} V_j
  double2* dStage = nullptr;  // one vector: caller staging / residual
  double2* dPart = nullptr;   // reduction partials
  double2* dH = nullptr;      // coefficients (2 passes) + norms
  double2* dY = nullptr;      // combination weights (m)
};

}  // namespace lsw
