// Solution output (see lsw_output.h).
#include "lsw_output.h"

#include <cstdint>
#include <cstdio>
#include <fstream>

#include "lsw_dd.h"

namespace lsw {

void writeVTK(const std::string& path, const DDCheckerboard& dd, const std::vector<Complex>& x) {
  const int Gx = dd.Gx(), Gy = dd.Gy();
  LSW_REQUIRE((long long)x.size() == (long long)Gx * Gy, LSW_ERR_ARG, "writeVTK: solution size mismatch");
  std::ofstream f(path);
  LSW_REQUIRE(f.good(), LSW_ERR_ARG, "writeVTK: cannot open " + path);
  f << "# vtk DataFile Version 1.0\n";
  f << "2D Unstuctured Grid of Linear Quads\n";
  f << "ASCII\n\n";
  f << "DATASET UNSTRUCTURED_GRID\n";
  f << "POINTS " << (long long)Gx * Gy << " float\n";
  for (int i = 0; i < Gx; ++i)
    for (int j = 0; j < Gy; ++j) f << dd.gx(i) << " " << dd.gy(j) << " 0.0\n";
  const long long nel = (long long)std::max(Gx - 1, 0) * std::max(Gy - 1, 0);
  f << "\nCELLS " << nel << " " << 5 * nel << "\n";
  for (int i = 0; i + 1 < Gx; ++i)
    for (int j = 0; j + 1 < Gy; ++j) {
      const long long a = (long long)i * Gy + j, b = (long long)(i + 1) * Gy + j;
      f << 4 << " " << a << " " << b << " " << b + 1 << " " << a + 1 << "\n";
    }
  f << "\nCELL_TYPES " << nel << "\n";
  for (long long e = 0; e < nel; ++e) f << 9 << "\n";
  f << "\nPOINT_DATA " << x.size() << "\n";
  f << "SCALARS wavefield_re float\n";
  f << "LOOKUP_TABLE default\n";
  for (const Complex& v : x) f << v.real() << "\n";
  f << "SCALARS wavefield_im float\n";
  f << "LOOKUP_TABLE default\n";
  for (const Complex& v : x) f << v.imag() << "\n";
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
