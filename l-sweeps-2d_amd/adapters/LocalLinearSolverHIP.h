// LocalLinearSolverHIP for the L-Sweeps-2D source tree: a LocalLinearSolver
// (include/LocalLinearSolver.h:8-41) whose factorisation and solves run on
// the MI355X through the C-ABI of liblsweeps_hip.so (include/lsweeps_hip.h).
// It sits beside LocalLinearSolverEigenSparseLU
// (include/LocalLinearSolverEigenSparseLU.h) and is selected by
// CheckerboardDD.Solver = "HIP" in Input::readCheckerboardDD.
//
// These two files are compiled inside the reference tree (they include its
// typedef.h and LocalLinearSolver.h, so Eigen3); the library itself has no
// Eigen dependency.  Semantics kept from the Eigen solver
// (src/LocalLinearSolverEigenSparseLU.cpp:9-76):
//   * the matrix is borrowed (owned by SubdomainStructured::mA);
//   * copy() copies the settings, never the factors (copies are made before
//     factorize);
//   * factorize() is idempotent;
//   * errors print "ERROR: ..." to cerr and exit(1).
// The matrix must be the 5-point operator of a structured grid, which every
// SubdomainStructured assembles (src/AssemblerHelmholtz2DFD.cpp:371-457).
#ifndef LOCALLINEARSOLVERHIP_H
#define LOCALLINEARSOLVERHIP_H

#include "LocalLinearSolver.h"
#include "lsweeps_hip.h"
#include "typedef.h"

class LocalLinearSolverHIP : public LocalLinearSolver
{
  public:
    //! precision: LSW_PREC_FP64 (the plain path) or LSW_PREC_FP32 factors
    explicit LocalLinearSolverHIP( int precision = LSW_PREC_FP64 );
    LocalLinearSolverHIP( const LocalLinearSolverHIP &other );
    ~LocalLinearSolverHIP() override;
    LocalLinearSolverHIP &operator=( const LocalLinearSolverHIP & ) = delete;

    LocalLinearSolver *copy() const override;
    void setMatrix( ComplexSparseMatrix *A ) override;
    void factorize() override;
    void solve( ComplexVector &b, ComplexVector &res ) const override;

  private:
    int mPrecision;
    ComplexSparseMatrix *mA = nullptr;   // borrowed
    lsw_lu *mLU = nullptr;               // the factors on the device
};

#endif
