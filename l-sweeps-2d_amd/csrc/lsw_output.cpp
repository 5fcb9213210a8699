// Solution output (see lsw_output.h).
#include "lsw_output.h"

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstdio>
#include <fstream>

#include "lsw_dd.h"

namespace lsw {

namespace {

// One cell's physical points: the DOF points (index i*ny + j) followed by the
// 2 nx + 2 ny points of the Dirichlet boundary of its computational domain,
// bottom (left to right), right (bottom to top), top (right to left), left
// (top to bottom): AssemblerHelmholtz2DFD::getNonDOFPhysPts
// (src/AssemblerHelmholtz2DFD.cpp:185-225), nx, ny = elements per axis.
struct CellPts {
  int ex, ey;               // elements per axis (DOFs + 1)
  double x0, x1, y0, y1, hx, hy;
  long long dofs() const { return (long long)(ex - 1) * (ey - 1); }
  void point(long long k, double* x, double* y) const {
    const long long size = dofs();
    if (k < size) {
      const long long i = k / (ey - 1), j = k - i * (ey - 1);
      *x = x0 + (double)(i + 1) * hx;
      *y = y0 + (double)(j + 1) * hy;
      return;
    }
    long long b = k - size;
    if (b < ex) { *x = x0 + (double)b * hx; *y = y0; return; }
    b -= ex;
    if (b < ey) { *x = x1; *y = y0 + hy * (double)b; return; }
    b -= ey;
    if (b < ex) { *x = x1 - (double)b * hx; *y = y1; return; }
    b -= ex;
    *x = x0;
    *y = y1 - hy * (double)b;
  }
  // corner point indices of element (J = x index, I = y index):
  // AssemblerHelmholtz2DFD::getElPhysPtIndices (src/AssemblerHelmholtz2DFD.cpp:229-329)
  void element(long long J, long long I, long long* r) const {
    const long long n0 = ex, n1 = ey, size = dofs();
    if (J == 0 && I == 0) {
      r[0] = size; r[1] = size + 1; r[2] = 0; r[3] = size + 2 * n0 + 2 * n1 - 1;
    } else if (J == n0 - 1 && I == 0) {
      r[0] = size + n0 - 1; r[1] = size + n0; r[2] = size + n0 + 1; r[3] = (n0 - 2) * (n1 - 1);
    } else if (J == n0 - 1 && I == n1 - 1) {
      r[0] = (n0 - 1) * (n1 - 1) - 1; r[1] = size + n0 + n1 - 1; r[2] = size + n0 + n1; r[3] = size + n0 + n1 + 1;
    } else if (J == 0 && I == n1 - 1) {
      r[0] = size + 2 * n0 + n1 + 1; r[1] = n1 - 2; r[2] = size + 2 * n0 + n1 - 1; r[3] = size + 2 * n0 + n1;
    } else if (I == 0) {
      r[0] = size + J; r[1] = size + J + 1; r[2] = J * (n1 - 1); r[3] = (J - 1) * (n1 - 1);
    } else if (J == n0 - 1) {
      r[0] = (n0 - 2) * (n1 - 1) + I - 1; r[1] = size + n0 + I; r[2] = size + n0 + I + 1; r[3] = (n0 - 2) * (n1 - 1) + I;
    } else if (I == n1 - 1) {
      r[0] = (J - 1) * (n1 - 1) + n1 - 2; r[1] = J * (n1 - 1) + n1 - 2;
      r[2] = size + 2 * n0 + n1 - 2 - (J - 1); r[3] = size + 2 * n0 + n1 - 1 - (J - 1);
    } else if (J == 0) {
      r[0] = size + 2 * n0 + 2 * n1 - I; r[1] = I - 1; r[2] = I; r[3] = size + 2 * n0 + 2 * n1 - 1 - I;
    } else {
      r[0] = (J - 1) * (n1 - 1) + (I - 1); r[1] = (J - 1) * (n1 - 1) + I; r[2] = J * (n1 - 1) + I;
      r[3] = J * (n1 - 1) + (I - 1);
    }
  }
};

// Element classes along one axis (SubdomainStructured::computeVolumeElIndexSets,
// src/SubdomainStructured.cpp:804-1000): the element midpoints (mean of the
// four corner coordinates) against the extent of trace0, trace1, traceN and
// traceNP of the unshifted 7-set view; start[k] is the first DOF line of set
// k (start[7] = DOF lines), coord(l) the coordinate of DOF line l.
template <typename Coord, typename Mid>
std::vector<int> elementClasses(int nel, const int* start, double lo, double hi, Coord coord, Mid mid) {
  double mn[4], mx[4];
  const int sets[4] = {1, 2, 4, 5};
  bool any = false;
  for (int t = 0; t < 4; ++t) {
    const int a = start[sets[t]], b = start[sets[t] + 1];
    if (b > a) {
      any = true;
      mn[t] = coord(a);
      mx[t] = coord(b - 1);
    } else {
      mn[t] = mx[t] = t < 2 ? lo - 1.0 : hi + 1.0;
    }
  }
  std::vector<int> cls(nel, 4);
  if (!any) return cls;  // no traces: every element is interior
  for (int e = 0; e < nel; ++e) {
    const double m = mid(e);
    int c = 8;
    if (m < mn[0]) c = 0;
    else if (m < mx[0]) c = 1;
    else if (m < mn[1]) c = 2;
    else if (m < mx[1]) c = 3;
    else if (m < mn[2]) c = 4;
    else if (m < mx[2]) c = 5;
    else if (m < mn[3]) c = 6;
    else if (m < mx[3]) c = 7;
    cls[e] = c;
  }
  return cls;
}

}  // namespace

void writeVTK(const std::string& path, const DDCheckerboard& dd, const std::vector<Complex>& x) {
  const int Gx = dd.Gx(), Gy = dd.Gy();
  LSW_REQUIRE((long long)x.size() == (long long)Gx * Gy, LSW_ERR_ARG, "writeVTK: solution size mismatch");
  // OutputVTK::print2D (src/OutputVTK.cpp:82-366), as one rank writes it:
  // cells in order; per cell the elements whose midpoint classes are 3..6 on
  // both axes, the physical points they use in the cell's point order, and
  // the value of the solution there (zero on the Dirichlet boundary).
  std::vector<double> px, py;
  std::vector<Complex> pv;
  std::vector<std::array<long long, 4>> els;
  for (int c = 0; c < dd.numCells(); ++c) {
    const CellGeom& g = dd.cellGeom(c);
    const auto& cd = g.asmb.compDomain();
    CellPts P{g.nx + 1, g.ny + 1, cd[0][0], cd[0][1], cd[1][0], cd[1][1], g.asmb.hx(), g.asmb.hy()};
    auto cornerX = [&](long long J) {  // the four corners' x, summed as the reference sums them
      long long r[4];
      P.element(J, std::min<long long>(1, P.ey - 1), r);
      double s = 0.0, y;
      for (int k = 0; k < 4; ++k) {
        double xv;
        P.point(r[k], &xv, &y);
        s = k == 0 ? xv : s + xv;
      }
      return s / 4.0;
    };
    auto cornerY = [&](long long I) {
      long long r[4];
      P.element(std::min<long long>(1, P.ex - 1), I, r);
      double s = 0.0, xv;
      for (int k = 0; k < 4; ++k) {
        double yv;
        P.point(r[k], &xv, &yv);
        s = k == 0 ? yv : s + yv;
      }
      return s / 4.0;
    };
    const auto clsX = elementClasses(
        P.ex, g.vlr[0], P.x0, P.x1, [&](int l) { return g.asmb.xCoord(l); }, cornerX);
    const auto clsY = elementClasses(
        P.ey, g.vud[0], P.y0, P.y1, [&](int l) { return g.asmb.yCoord(l); }, cornerY);
    const long long npts = P.dofs() + 2LL * P.ex + 2LL * P.ey;
    std::vector<long long> local(npts, -1);
    const size_t first = els.size();
    for (long long J = 0; J < P.ex; ++J) {
      if (clsX[J] < 3 || clsX[J] > 6) continue;
      for (long long I = 0; I < P.ey; ++I) {
        if (clsY[I] < 3 || clsY[I] > 6) continue;
        std::array<long long, 4> r;
        P.element(J, I, r.data());
        for (long long k : r) local[k] = 0;
        els.push_back(r);
      }
    }
    for (long long k = 0; k < npts; ++k) {
      if (local[k] < 0) continue;
      local[k] = (long long)px.size();
      double xv, yv;
      P.point(k, &xv, &yv);
      px.push_back(xv);
      py.push_back(yv);
      if (k < P.dofs()) {
        const long long i = k / g.ny, j = k - i * g.ny;
        pv.push_back(x[(size_t)(i + g.ox) * Gy + (j + g.oy)]);
      } else {
        pv.push_back(Complex(0.0, 0.0));
      }
    }
    for (size_t e = first; e < els.size(); ++e)
      for (auto& k : els[e]) k = local[k];
  }
  std::ofstream f(path);
  LSW_REQUIRE(f.good(), LSW_ERR_ARG, "writeVTK: cannot open " + path);
  f << "# vtk DataFile Version 1.0\n";
  f << "2D Unstuctured Grid of Linear Quads\n";
  f << "ASCII\n\n";
  f << "DATASET UNSTRUCTURED_GRID\n";
  f << "POINTS " << px.size() << " float\n";
  for (size_t k = 0; k < px.size(); ++k) f << px[k] << " " << py[k] << " 0.0\n";
  f << "\nCELLS " << els.size() << " " << 5 * els.size() << "\n";
  for (const auto& r : els) f << 4 << " " << r[0] << " " << r[1] << " " << r[2] << " " << r[3] << "\n";
  f << "\nCELL_TYPES " << els.size() << "\n";
  for (size_t e = 0; e < els.size(); ++e) f << 9 << "\n";
  f << "\nPOINT_DATA " << pv.size() << "\n";
  f << "SCALARS wavefield_re float\n";
  f << "LOOKUP_TABLE default\n";
  for (const Complex& v : pv) f << v.real() << "\n";
  f << "SCALARS wavefield_im float\n";
  f << "LOOKUP_TABLE default\n";
  for (const Complex& v : pv) f << v.imag() << "\n";
  LSW_REQUIRE(f.good(), LSW_ERR_ARG, "writeVTK: write failed for " + path);
}

void writeGrid(const std::string& path, const DDCheckerboard& dd, const std::vector<Complex>& x) {
  const int32_t Gx = dd.Gx(), Gy = dd.Gy();
  LSW_REQUIRE((long long)x.size() == (long long)Gx * Gy, LSW_ERR_ARG, "writeGrid: solution size mismatch");
  std::FILE* f = std::fopen(path.c_str(), "wb");
  LSW_REQUIRE(f != nullptr, LSW_ERR_ARG, "writeGrid: cannot open " + path);
  bool ok = std::fwrite("LSWGRID1", 1, 8, f) == 8;
  ok = ok && std::fwrite(&Gx, 4, 1, f) == 1 && std::fwrite(&Gy, 4, 1, f) == 1;
  std::vector<double> xs(Gx), ys(Gy);
  for (int i = 0; i < Gx; ++i) xs[i] = dd.gx(i);
  for (int j = 0; j < Gy; ++j) ys[j] = dd.gy(j);
  ok = ok && std::fwrite(xs.data(), 8, xs.size(), f) == xs.size();
  ok = ok && std::fwrite(ys.data(), 8, ys.size(), f) == ys.size();
  ok = ok && std::fwrite(x.data(), sizeof(Complex), x.size(), f) == x.size();
  ok = (std::fclose(f) == 0) && ok;
  LSW_REQUIRE(ok, LSW_ERR_ARG, "writeGrid: write failed for " + path);
}

}  // namespace lsw
