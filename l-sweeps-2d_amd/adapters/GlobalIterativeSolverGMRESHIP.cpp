// GlobalIterativeSolverGMRESHIP for the L-Sweeps-2D source tree (see the header).
#include "GlobalIterativeSolverGMRESHIP.h"

#include <cmath>
#include <cstdlib>
#include <iostream>

#include "DD.h"
#include "Subdomain.h"

using namespace std;

namespace {
void lswCheck( int status, const char *what )
{
  if( status != LSW_OK )
  {
    cerr << "ERROR: " << what << ": " << lsw_last_error() << endl;
    exit(1);
  }
}
}

GlobalIterativeSolverGMRESHIP::GlobalIterativeSolverGMRESHIP( const string &inputFile, int problemID, Double tol,
                                                              int maxIter, bool verbose, int mode, int restart )
  : mFile(inputFile), mProblem(problemID)
{
  mParams.mode = mode;
  mParams.tol = tol;
  mParams.max_iter = maxIter;
  mParams.restart = restart;
  mParams.verbose = verbose ? 1 : 0;
}

GlobalIterativeSolverGMRESHIP::GlobalIterativeSolverGMRESHIP( const GlobalIterativeSolverGMRESHIP &other )
  : GlobalIterativeSolver(other), mFile(other.mFile), mProblem(other.mProblem), mParams(other.mParams) {}

GlobalIterativeSolverGMRESHIP::~GlobalIterativeSolverGMRESHIP()
{
  if( mDD != nullptr )
    lsw_dd_destroy( mDD );
}

GlobalIterativeSolver *GlobalIterativeSolverGMRESHIP::copy() const
{
  return new GlobalIterativeSolverGMRESHIP(*this);
}

void GlobalIterativeSolverGMRESHIP::prepare( const vector<ComplexVector> &rhs ) const
{
  if( mDD != nullptr )
    return;
  lswCheck( lsw_dd_create_config( mFile.c_str(), mProblem, &mDD ), "lsw_dd_create_config" );
  int ncells = 0;
  double meshSize = 0.0;
  lswCheck( lsw_dd_sizes( mDD, &mGx, &mGy, &ncells, &meshSize ), "lsw_dd_sizes" );
  if( (UInt)ncells != mParent->getNCells() )
  {
    cerr << "ERROR: GMRESHIP needs every cell on this process (run with one rank)" << endl;
    exit(1);
  }
  vector<double> xs(mGx), ys(mGy);
  lswCheck( lsw_dd_global_coords( mDD, xs.data(), ys.data() ), "lsw_dd_global_coords" );
  const double hx = mGx > 1 ? xs[1] - xs[0] : 1.0, hy = mGy > 1 ? ys[1] - ys[0] : 1.0;
  // owned entry k of cell i holds k + 1; expand puts it at its local DOF
  vector<ComplexVector> probe( rhs.size() ), local;
  for( UInt i = 0; i < rhs.size(); i++ )
  {
    probe[i] = ComplexVector( rhs[i].size() );
    for( Int k = 0; k < rhs[i].size(); k++ )
      probe[i][k] = Complex( k + 1.0, 0.0 );
  }
  mParent->expand( probe, local );
  mGlobal.assign( rhs.size(), vector<long long>() );
  for( UInt i = 0; i < rhs.size(); i++ )
  {
    vector<DoubleVector> pts;
    mParent->getCell(i)->getDOFPts( pts );
    mGlobal[i].assign( rhs[i].size(), -1 );
    for( Int p = 0; p < local[i].size(); p++ )
    {
      const double v = local[i][p].real();
      if( v == 0.0 )
        continue;
      const long long gi = llround( (pts[p][0] - xs[0]) / hx ), gj = llround( (pts[p][1] - ys[0]) / hy );
      mGlobal[i][(long long)v - 1] = gi * mGy + gj;
    }
    for( long long g : mGlobal[i] )
      if( g < 0 || g >= (long long)mGx * mGy )
      {
        cerr << "ERROR: GMRESHIP could not place an owned DOF on the global grid" << endl;
        exit(1);
      }
  }
  lswCheck( lsw_dd_factorize( mDD ), "lsw_dd_factorize" );
}

void GlobalIterativeSolverGMRESHIP::solve( const vector<ComplexVector> &rhs, vector<ComplexVector> &res ) const
{
  prepare( rhs );
  vector<Complex> b( (size_t)mGx * mGy, Complex(0.0, 0.0) ), x( b.size() );
  for( UInt i = 0; i < rhs.size(); i++ )
    for( Int k = 0; k < rhs[i].size(); k++ )
      b[mGlobal[i][k]] = rhs[i][k];
  lsw_gmres_result r;
  lswCheck( lsw_gmres_solve( mDD, &mParams, reinterpret_cast<const double *>(b.data()),
                             reinterpret_cast<double *>(x.data()), &r, nullptr, 0 ), "lsw_gmres_solve" );
  res.resize( rhs.size() );
  for( UInt i = 0; i < rhs.size(); i++ )
  {
    res[i] = ComplexVector( rhs[i].size() );
    for( Int k = 0; k < rhs[i].size(); k++ )
      res[i][k] = x[mGlobal[i][k]];
  }
}
