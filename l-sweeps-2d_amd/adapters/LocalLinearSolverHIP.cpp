// LocalLinearSolverHIP for the L-Sweeps-2D source tree (see the header).
#include "LocalLinearSolverHIP.h"

#include <cstdlib>
#include <iostream>
#include <type_traits>

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

// Eigen's default sparse storage is column major with int indices: the CSC
// arrays of lsw_lu_set_csc_z, and complex<double> values are interleaved
// (re, im) doubles.
static_assert( !ComplexSparseMatrix::IsRowMajor, "CSC storage expected" );
static_assert( is_same<ComplexSparseMatrix::StorageIndex, int>::value, "int32 indices expected" );
static_assert( sizeof(Complex) == 2 * sizeof(double), "interleaved complex expected" );

LocalLinearSolverHIP::LocalLinearSolverHIP( int precision ) : mPrecision(precision) {}

LocalLinearSolverHIP::LocalLinearSolverHIP( const LocalLinearSolverHIP &other )
  : LocalLinearSolver(other), mPrecision(other.mPrecision), mA(other.mA) {}

LocalLinearSolverHIP::~LocalLinearSolverHIP()
{
  if( mLU != nullptr )
    lsw_lu_destroy( mLU );
}

LocalLinearSolver *LocalLinearSolverHIP::copy() const
{
  return new LocalLinearSolverHIP(*this);
}

void LocalLinearSolverHIP::setMatrix( ComplexSparseMatrix *A )
{
  mA = A;
  if( mLU != nullptr )
  {
    lsw_lu_destroy( mLU );
    mLU = nullptr;
  }
}

void LocalLinearSolverHIP::factorize()
{
  if( mLU != nullptr )
    return;
  if( mA == nullptr )
  {
    cerr << "ERROR: Factorize called before a matrix was set!" << endl;
    exit(1);
  }
  // the subdomain stores its matrix compressed (src/SubdomainStructured.cpp:650-653)
  mA->makeCompressed();
  lswCheck( lsw_lu_create( mPrecision, &mLU ), "lsw_lu_create" );
  lswCheck( lsw_lu_set_csc_z( mLU, (int)mA->rows(), (int)mA->nonZeros(), mA->outerIndexPtr(), mA->innerIndexPtr(),
                              reinterpret_cast<const double *>(mA->valuePtr()) ), "lsw_lu_set_csc_z" );
  lswCheck( lsw_lu_factorize( mLU ), "lsw_lu_factorize" );
}

void LocalLinearSolverHIP::solve( ComplexVector &b, ComplexVector &res ) const
{
  if( mLU == nullptr )
  {
    cerr << "ERROR: Solve called before factorization!" << endl;
    exit(1);
  }
  res.resize( b.size() );
  lswCheck( lsw_lu_solve_z( mLU, 1, reinterpret_cast<const double *>(b.data()), reinterpret_cast<double *>(res.data()) ),
            "lsw_lu_solve_z" );
}
