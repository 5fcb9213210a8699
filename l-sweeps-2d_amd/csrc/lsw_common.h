// Common types and error handling for the MI355X L-sweeps library.
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lsweeps_hip.h"

namespace lsw {

using Double = double;
using UInt = unsigned int;
using Complex = std::complex<double>;
using ComplexVector = std::vector<Complex>;
using DoubleVector = std::vector<double>;
using UIntVector = std::vector<UInt>;

// Host-side error: the C-ABI layer turns it into a status code plus a message.
struct Error : public std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Status codes are the LSW_* macros of the C-ABI header.


#define LSW_REQUIRE(cond, code, msg)                      \
  do {                                                    \
    if (!(cond)) throw ::lsw::Error((code), (msg));       \
  } while (0)

// Compressed sparse column matrix with int32 indices, the layout of the
// reference's ComplexSparseMatrix (Eigen::SparseMatrix<complex<double>>,
// column-major, include/typedef.h:64-73).
struct ComplexSparseMatrix {
  int rows = 0, cols = 0;
  std::vector<int> colptr;   // cols+1
  std::vector<int> rowind;   // nnz
  std::vector<Complex> vals; // nnz
  int nonZeros() const { return (int)rowind.size(); }
};

enum Precision : int { LSW_FP64 = 0, LSW_FP32 = 1 };

}  // namespace lsw
