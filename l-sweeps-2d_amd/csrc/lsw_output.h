// Solution output.
//
// writeVTK follows the file layout of OutputVTK::print2D
// (src/OutputVTK.cpp:82-366): a legacy-ASCII "2D Unstuctured Grid of Linear
// Quads" with POINTS (x y 0.0), CELLS of 4 point indices, CELL_TYPES 9 and two
// point scalars wavefield_re / wavefield_im printed with the stream's default
// 6 significant digits.  The reference gathers the owned elements of every
// cell onto rank 0; here the solution already lives on one global grid, so
// the points are the global DOF grid (x-major, index gi*Gy + gj) and the
// quads join neighbouring DOFs.  Values are the same at every node; the point
// and element order differ.
//
// writeGrid writes the full-precision sibling of the same field:
//   "LSWGRID1", int32 Gx, int32 Gy, Gx doubles (x), Gy doubles (y),
//   Gx*Gy complex doubles (re, im interleaved), little endian.
#pragma once

#include <string>
#include <vector>

#include "lsw_common.h"

namespace lsw {

class DDCheckerboard;

void writeVTK(const std::string& path, const DDCheckerboard& dd, const std::vector<Complex>& x);
void writeGrid(const std::string& path, const DDCheckerboard& dd, const std::vector<Complex>& x);

}  // namespace lsw
