// Host discretisation classes (see lsw_model.h).
#include "lsw_model.h"

#include <algorithm>
#include <cmath>
#include <fstream>
#include <sstream>

#include "lsw_stencil.h"

namespace lsw {

// ------------------------------------------------------------ FunctionCSV
void FunctionCSV::setFile(const std::string& filename, const std::vector<DoubleVector>& fileDom) {
  std::ifstream in(filename);
  LSW_REQUIRE(in.good(), LSW_ERR_ARG, "FunctionCSV: cannot open " + filename);
  std::vector<float> vals;
  int nrows = 0, ncols = 0;
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string v;
    int nc = 0;
    while (std::getline(ss, v, ',')) {
      vals.push_back(std::stof(v));
      ++nc;
    }
    if (nc == 0) continue;
    LSW_REQUIRE(nrows == 0 || nc == ncols, LSW_ERR_ARG, "FunctionCSV: ragged rows in " + filename);
    ncols = nc;
    ++nrows;
  }
  setData(vals.data(), nrows, ncols, fileDom);
}

void FunctionCSV::setData(const float* c, int nrows, int ncols, const std::vector<DoubleVector>& fileDom) {
  LSW_REQUIRE(nrows > 0 && ncols > 0, LSW_ERR_ARG, "FunctionCSV: empty velocity grid");
  mC = std::make_shared<const std::vector<float>>(c, c + (size_t)nrows * ncols);
  mNRows = nrows;
  mNCols = ncols;
  mFileDom = fileDom;
  mH[0] = (fileDom[0][1] - fileDom[0][0]) / ncols;
  mH[1] = (fileDom[1][1] - fileDom[1][0]) / nrows;
  mDomSet = false;
}

void FunctionCSV::setDomain(const std::vector<DoubleVector>& dom) const {
  double d[2][2];
  d[0][0] = std::max(dom[0][0], mFileDom[0][0]);
  d[1][0] = std::max(dom[1][0], mFileDom[1][0]);
  d[0][1] = std::min(dom[0][1], mFileDom[0][1]);
  d[1][1] = std::min(dom[1][1], mFileDom[1][1]);
  UInt startRow = (UInt)((mFileDom[1][1] - d[1][1]) / mH[1]);
  UInt endRow = (UInt)((mFileDom[1][1] - d[1][0]) / mH[1]);
  UInt startCol = (UInt)((d[0][0] - mFileDom[0][0]) / mH[0]);
  UInt endCol = (UInt)((d[0][1] - mFileDom[0][0]) / mH[0]);
  endRow = std::min(endRow, (UInt)mNRows - 1);
  endCol = std::min(endCol, (UInt)mNCols - 1);
  mDomA[0][0] = mFileDom[0][0] + startCol * mH[0];
  mDomA[0][1] = mFileDom[0][0] + (endCol + 1) * mH[0];
  mDomA[1][0] = mFileDom[1][1] - (endRow + 1) * mH[1];
  mDomA[1][1] = mFileDom[1][1] - startRow * mH[1];
  mARows = (int)(endRow - startRow + 1);
  mACols = (int)(endCol - startCol + 1);
  auto a = std::make_shared<std::vector<double>>((size_t)mARows * mACols, 0.0);
  const std::vector<float>& C = *mC;
  bool first = true;
  for (int r = 0; r < mARows; ++r)
    for (int cc = 0; cc < mACols; ++cc) {
      const double c = (double)C[(size_t)(startRow + r) * mNCols + startCol + cc];
      const double v = 1.0 / std::pow(c, 2.0);
      (*a)[(size_t)r * mACols + cc] = v;
      if (first) {
        mMinReal = v;
        first = false;
      } else {
        mMinReal = std::min(mMinReal, v);
      }
    }
  mA = std::move(a);
  mDomSet = true;
}

double FunctionCSV::getMinReal() const {
  if (!mDomSet) setDomain(mFileDom);
  return mMinReal;
}

Complex FunctionCSV::eval(const DoubleVector& xin) const {
  if (!mDomSet) setDomain(mFileDom);
  double x[2] = {xin[0], xin[1]};
  for (int i = 0; i < 2; ++i)
    if (x[i] <= mDomA[i][0] + 0.5 * mH[0]) x[i] = mDomA[i][0] + 0.5 * mH[0];
  for (int i = 0; i < 2; ++i)
    if (x[i] >= mDomA[i][1] - 0.5 * mH[1]) x[i] = mDomA[i][1] - 0.5 * mH[1];
  const UInt col = (UInt)((x[0] - mDomA[0][0]) / mH[0]);
  const UInt row = (UInt)((mDomA[1][1] - x[1]) / mH[1]);
  LSW_REQUIRE((int)row < mARows && (int)col < mACols, LSW_ERR_ARG, "FunctionCSV: lookup out of range");
  return (*mA)[(size_t)row * mACols + col];
}

// ---------------------------------------------------------------------- PML
void PML::setWidth(double w) {
  mWidth = w;
  if (!mIntDomain.empty()) {
    mCompDomain = mIntDomain;
    mCompDomain[0][0] -= w;
    mCompDomain[0][1] += w;
    mCompDomain[1][0] -= w;
    mCompDomain[1][1] += w;
  }
}

void PML::setDomain(const std::vector<DoubleVector>& dom) {
  mIntDomain = dom;
  if (mWidth > 0) {
    mCompDomain = mIntDomain;
    mCompDomain[0][0] -= mWidth;
    mCompDomain[0][1] += mWidth;
    mCompDomain[1][0] -= mWidth;
    mCompDomain[1][1] += mWidth;
  }
}

// ---------------------------------------------------------------- RHSSrcPts
Complex RHSSrcPts::eval(const DoubleVector& x) const {
  Complex res = 0.0;
  const double h = mEps;
  for (const auto& x0 : mSourcePts) {
    double d2 = 0.0;
    for (size_t i = 0; i < x.size(); ++i) d2 += std::pow(x0[i] - x[i], 2.0);
    res += 1.0 / (kPi * std::pow(h, (double)x.size())) * std::exp(-d2 / std::pow(h, 2.0));
  }
  return res;
}

// --------------------------------------------------- AssemblerHelmholtz2DFD
AssemblerHelmholtz2DFD::AssemblerHelmholtz2DFD(const AssemblerHelmholtz2DFD& o)
    : mOmega(o.mOmega), mOrder(o.mOrder), mN(o.mN), mIntDomain(o.mIntDomain),
      mCompDomain(o.mCompDomain) {
  if (o.mPML) mPML.reset(static_cast<PMLCubic2D*>(o.mPML->copy()));
  if (o.mM) mM.reset(o.mM->copy());
}

AssemblerHelmholtz2DFD& AssemblerHelmholtz2DFD::operator=(const AssemblerHelmholtz2DFD& o) {
  if (this != &o) {
    mOmega = o.mOmega;
    mOrder = o.mOrder;
    mN = o.mN;
    mIntDomain = o.mIntDomain;
    mCompDomain = o.mCompDomain;
    mPML.reset(o.mPML ? static_cast<PMLCubic2D*>(o.mPML->copy()) : nullptr);
    mM.reset(o.mM ? o.mM->copy() : nullptr);
  }
  return *this;
}

void AssemblerHelmholtz2DFD::setDomain(const std::vector<DoubleVector>& domain) {
  mIntDomain = domain;
  if (mPML) mPML->setDomain(domain);
  mN.clear();
}

// src/AssemblerHelmholtz2DFD.cpp:67-114: PML width rounded to the nearest
// multiple of h; computational domain = interior +- width.
void AssemblerHelmholtz2DFD::setN(const std::vector<UInt>& N) {
  LSW_REQUIRE(!mIntDomain.empty(), LSW_ERR_STATE, "Domain has to be set before setting N");
  LSW_REQUIRE(mPML != nullptr, LSW_ERR_STATE, "PML has to be set before setting N");
  const double h1 = (mIntDomain[0][1] - mIntDomain[0][0]) / N[0];
  const double h2 = (mIntDomain[1][1] - mIntDomain[1][0]) / N[1];
  const double h = std::min(h1, h2);
  double pmlWidth = mPML->getWidth();
  const double ratio = pmlWidth / h;
  UInt pmlEls = (UInt)(pmlWidth / h);
  if (ratio - pmlEls > 0.5) pmlEls += 1;
  pmlWidth = h * pmlEls;
  mPML->setWidth(pmlWidth);
  mPML->setDomain(mIntDomain);
  mCompDomain = mIntDomain;
  mCompDomain[0][0] -= pmlWidth;
  mCompDomain[0][1] += pmlWidth;
  mCompDomain[1][0] -= pmlWidth;
  mCompDomain[1][1] += pmlWidth;
  mN.assign(2, 0);
  mN[0] = (UInt)((mCompDomain[0][1] - mCompDomain[0][0] + 0.5 * h1) / h1);
  mN[1] = (UInt)((mCompDomain[1][1] - mCompDomain[1][0] + 0.5 * h2) / h2);
}

double AssemblerHelmholtz2DFD::hx() const { return (mCompDomain[0][1] - mCompDomain[0][0]) / mN[0]; }
double AssemblerHelmholtz2DFD::hy() const { return (mCompDomain[1][1] - mCompDomain[1][0]) / mN[1]; }

double AssemblerHelmholtz2DFD::getMeshSize() const { return std::sqrt(hx() * hy()); }

void AssemblerHelmholtz2DFD::getDOFPts(std::vector<DoubleVector>& res) const {
  const int nx = nxDofs(), ny = nyDofs();
  res.assign((size_t)nx * ny, DoubleVector(2));
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      res[(size_t)i * ny + j][0] = xCoord(i);
      res[(size_t)i * ny + j][1] = yCoord(j);
    }
}

void AssemblerHelmholtz2DFD::squaredSlowness(std::vector<double>& m) const {
  const int nx = nxDofs(), ny = nyDofs();
  m.resize((size_t)nx * ny);
  DoubleVector p(2);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      p[0] = xCoord(i);
      p[1] = yCoord(j);
      m[(size_t)i * ny + j] = mM->eval(p).real();
    }
}

void AssemblerHelmholtz2DFD::assembleSystemMatrix(ComplexSparseMatrix& A) const {
  LSW_REQUIRE(mOrder == 2, LSW_ERR_UNSUPPORTED, "only Order 2 is supported");
  const int nx = nxDofs(), ny = nyDofs();
  const int N = nx * ny;
  Geo2 g;
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b) {
      g.intDom[a][b] = mIntDomain[a][b];
      g.compDom[a][b] = mCompDomain[a][b];
    }
  std::vector<double> m;
  squaredSlowness(m);
  const double C = mPML->getAbsorptionCoeff();
  A.rows = A.cols = N;
  A.colptr.assign(N + 1, 0);
  A.rowind.clear();
  A.vals.clear();
  // Column d holds A[q][d] for q in {d-ny, d-1, d, d+1, d+ny}; A[q][d] is the
  // coefficient of row q towards d.
  std::vector<double> cf((size_t)N * 10);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const int d = i * ny + j;
      const bool ok = helmholtzStencil(xCoord(i), yCoord(j), g, C, mOmega, m[d], hx(), hy(),
                                       &cf[(size_t)d * 10]);
      LSW_REQUIRE(ok, LSW_ERR_ARG, "PML: Evaluation point out of range!");
    }
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) {
      const int d = i * ny + j;
      auto put = [&](int q, int dir) {
        A.rowind.push_back(q);
        A.vals.emplace_back(cf[(size_t)q * 10 + 2 * dir], cf[(size_t)q * 10 + 2 * dir + 1]);
      };
      if (i > 0) put(d - ny, 2);       // row (i-1,j): its (i+1) neighbour is d
      if (j > 0) put(d - 1, 4);        // row (i,j-1): its (j+1) neighbour
      put(d, 0);
      if (j < ny - 1) put(d + 1, 3);   // row (i,j+1): its (j-1) neighbour
      if (i < nx - 1) put(d + ny, 1);  // row (i+1,j): its (i-1) neighbour
      A.colptr[d + 1] = (int)A.rowind.size();
    }
}

void AssemblerHelmholtz2DFD::assembleRHS(const RHS* rhs, ComplexVector& res) const {
  std::vector<DoubleVector> pts;
  getDOFPts(pts);
  res.assign(pts.size(), 0.0);
  for (size_t i = 0; i < pts.size(); ++i) res[i] += 1.0 * rhs->eval(pts[i]);
}

}  // namespace lsw
