// Solution output.
//
// writeVTK writes the file OutputVTK::print2D (src/OutputVTK.cpp:82-366)
// writes for a one-rank run: a legacy-ASCII "2D Unstuctured Grid of Linear
// Quads"; cells in order, each contributing the elements whose midpoints lie
// in its printed window (element classes 3..6 on both axes of
// SubdomainStructured::computeVolumeElIndexSets) and the physical points
// those elements use, in the cell's point order (its DOF points, then the
// Dirichlet boundary points of its computational domain); corner indices as
// AssemblerHelmholtz2DFD::getElPhysPtIndices numbers them; POINTS (x y 0.0),
// CELLS of 4 point indices, CELL_TYPES 9, and point scalars wavefield_re /
// wavefield_im, all at the stream's default 6 significant digits.  Points on
// the border of two windows appear once per cell, as in the reference; a DOF
// point carries the solution there and a boundary point zero (setVals'
// expansion and trace exchange, src/OutputVTK.cpp:65-73).
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
