// GlobalIterativeSolverGMRESHIP for the L-Sweeps-2D source tree: a
// GlobalIterativeSolver (include/GlobalIterativeSolver.h:16-49) that solves
// the parent DD's system on the MI355X through liblsweeps_hip.so
// (include/lsweeps_hip.h), beside GlobalIterativeSolverGMRES.  Selected by
// GlobalSolvers.<name>.Type = "GMRESHIP" in Input::readSolver.
//
// The whole preconditioned solve runs on the device: the library builds its
// own device-resident copy of the problem's DD from the input file and
// problem id (lsw_dd_create_config: the parameters Input::readCheckerboardDD
// reads), factorises it once, and runs the reference GMRES arithmetic
// (mode 0, GlobalIterativeSolverGMRES::solve, src/GlobalIterativeSolverGMRES.cpp:65-348)
// or flexible GMRES to the true residual (mode 1).
//
// solve() takes and returns the reference's vectors: one ComplexVector of
// owned DOFs per local cell, in the parent's mCells order.  The first call
// maps every owned entry to its point of the global grid through the parent
// DD itself: DD::expand places owned entry k of cell i at its local DOF,
// whose coordinates SubdomainStructured::getDOFPts gives.  All cells must be
// on this process (run the reference with one rank: the GPU takes the place
// of the ranks).
#ifndef GLOBALITERATIVESOLVERGMRESHIP_H
#define GLOBALITERATIVESOLVERGMRESHIP_H

#include <string>
#include <vector>

#include "GlobalIterativeSolver.h"
#include "lsweeps_hip.h"

class GlobalIterativeSolverGMRESHIP : public GlobalIterativeSolver
{
  public:
    //! input file and problem id the parent DD was read from; GMRES settings
    //! as Input::readSolver reads them (Tol, MaxIter, Verbose)
    GlobalIterativeSolverGMRESHIP( const std::string &inputFile, int problemID, Double tol = 1.0e-6,
                                   int maxIter = 40, bool verbose = true, int mode = 0, int restart = 30 );
    GlobalIterativeSolverGMRESHIP( const GlobalIterativeSolverGMRESHIP &other );
    ~GlobalIterativeSolverGMRESHIP() override;
    GlobalIterativeSolverGMRESHIP &operator=( const GlobalIterativeSolverGMRESHIP & ) = delete;

    GlobalIterativeSolver *copy() const override;
    void solve( const vector<ComplexVector> &rhs, vector<ComplexVector> &res ) const override;

  private:
    void prepare( const vector<ComplexVector> &rhs ) const;

    std::string mFile;
    int mProblem;
    lsw_gmres_params mParams;
    mutable lsw_dd *mDD = nullptr;                       // device DD, factorised
    mutable int mGx = 0, mGy = 0;
    mutable std::vector<std::vector<long long>> mGlobal; // [cell][owned entry] -> gi * Gy + gj
};

#endif
