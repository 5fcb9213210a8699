// Host classes that mirror the reference's discretisation interfaces:
// Function / FunctionConstant / FunctionCSV (include/Function.h:33-68,
// include/FunctionCSV.h), PML / PMLCubic2D (include/PML.h:12-93,
// include/PMLCubic2D.h), RHS / RHSSrcPts (include/RHS.h:10-30,
// include/RHSSrcPts.h) and Assembler / AssemblerHelmholtz2DFD
// (include/Assembler.h:12-122, include/AssemblerHelmholtz2DFD.h).
//
// They set up grids and coefficients on the host; the per-node operator
// values are produced on the GPU by the assembly kernel (lsw_kernels.hip).
// AssemblerHelmholtz2DFD::assembleSystemMatrix also builds the CSC matrix on
// the host for callers of the LocalLinearSolver interface.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "lsw_common.h"

namespace lsw {

constexpr double kPi = 3.14159265358979323846;  // M_PI (include/utils.h:12)

// ---------------------------------------------------------------- Function
class Function {
 public:
  virtual ~Function() = default;
  virtual Function* copy() const = 0;
  virtual Complex eval(const DoubleVector& x) const = 0;
  virtual void setDomain(const std::vector<DoubleVector>& dom) const = 0;
  virtual double getMinReal() const = 0;
  virtual void clearCache() const {}
};

// Squared slowness "Constant": the reference always uses 1.0
// (src/Input.cpp:744-751, src/FunctionConstant.cpp:67-70).
class FunctionConstant : public Function {
 public:
  explicit FunctionConstant(double v = 1.0) : mVal(v) {}
  Function* copy() const override { return new FunctionConstant(*this); }
  Complex eval(const DoubleVector&) const override { return mVal; }
  void setDomain(const std::vector<DoubleVector>&) const override {}
  double getMinReal() const override { return mVal; }
  void setVal(double v) { mVal = v; }

 private:
  double mVal;
};

// Velocities on a top-row-first grid over fileDom, m = 1/c^2 with c parsed as
// float (stof).  setDomain crops to the cells covering the domain; eval clamps
// with mH[0] on the low side and mH[1] on the high side of both axes
// (src/FunctionCSV.cpp:46-236).
class FunctionCSV : public Function {
 public:
  Function* copy() const override { return new FunctionCSV(*this); }
  // Reads a CSV file (rows of comma-separated velocities).
  void setFile(const std::string& filename, const std::vector<DoubleVector>& fileDom);
  // Same, from values already in memory (row-major, row 0 = top).
  void setData(const float* c, int nrows, int ncols, const std::vector<DoubleVector>& fileDom);
  Complex eval(const DoubleVector& x) const override;
  void setDomain(const std::vector<DoubleVector>& dom) const override;
  double getMinReal() const override;

 private:
  // The velocity grid and the cropped squared slowness are immutable once
  // made, so copies (one per subdomain, prototype-style) share them; a copy
  // that is given another domain makes its own cropped values.
  std::shared_ptr<const std::vector<float>> mC;
  int mNRows = 0, mNCols = 0;
  std::vector<DoubleVector> mFileDom;
  double mH[2] = {0, 0};
  mutable std::shared_ptr<const std::vector<double>> mA;  // cropped m values, row-major
  mutable int mARows = 0, mACols = 0;
  mutable double mDomA[2][2] = {{0, 0}, {0, 0}};
  mutable double mMinReal = 0;
  mutable bool mDomSet = false;
};

// --------------------------------------------------------------------- PML
class PML {
 public:
  virtual ~PML() = default;
  virtual PML* copy() const = 0;
  void setWidth(double w);
  double getWidth() const { return mWidth; }
  void setDomain(const std::vector<DoubleVector>& dom);
  std::vector<DoubleVector> getDomain() const { return mIntDomain; }
  const std::vector<DoubleVector>& compDomain() const { return mCompDomain; }

 protected:
  double mWidth = 0;
  std::vector<DoubleVector> mIntDomain, mCompDomain;
};

// Cubic PML: sigma(x) = C ((x_int - x)/delta)^3 outside the interior,
// alpha = 1/(1 + i sigma) (src/PMLCubic2D.cpp:35-236).  C = ln(omega)
// (src/Input.cpp:836-837).
class PMLCubic2D : public PML {
 public:
  PML* copy() const override { return new PMLCubic2D(*this); }
  void setAbsorptionCoeff(double C) { mC = C; }
  double getAbsorptionCoeff() const { return mC; }

 private:
  double mC = 0;
};

// --------------------------------------------------------------------- RHS
class RHS {
 public:
  virtual ~RHS() = default;
  virtual RHS* copy() const = 0;
  virtual Complex eval(const DoubleVector& x) const = 0;
};

// Sum over point sources of exp(-|x-x0|^2/eps^2)/(pi eps^2)
// (src/RHSSrcPts.cpp:87-112); eps is the DD mesh size (src/Input.cpp:243).
class RHSSrcPts : public RHS {
 public:
  RHS* copy() const override { return new RHSSrcPts(*this); }
  void setSrcPts(const std::vector<DoubleVector>& pts) { mSourcePts = pts; }
  const std::vector<DoubleVector>& getSrcPts() const { return mSourcePts; }
  void setEpsilon(double eps) { mEps = eps; }
  double getEpsilon() const { return mEps; }
  Complex eval(const DoubleVector& x) const override;

 private:
  std::vector<DoubleVector> mSourcePts;
  double mEps = 0;
};

// ---------------------------------------------------------------- Assembler
class Assembler {
 public:
  virtual ~Assembler() = default;
  virtual Assembler* copy() const = 0;
  virtual void setDomain(const std::vector<DoubleVector>& domain) = 0;
  virtual void setN(const std::vector<UInt>& N) = 0;
  virtual void setOrder(UInt order) = 0;
  virtual UInt getOrder() const = 0;
  virtual double getMeshSize() const = 0;
  virtual void getDOFPts(std::vector<DoubleVector>& res) const = 0;
  virtual void assembleSystemMatrix(ComplexSparseMatrix& res) const = 0;
  virtual void assembleRHS(const RHS* rhs, ComplexVector& res) const = 0;
  virtual UInt getDim() const { return 2; }
};

// 5-point FD Helmholtz operator with cubic PML (src/AssemblerHelmholtz2DFD.cpp).
class AssemblerHelmholtz2DFD : public Assembler {
 public:
  AssemblerHelmholtz2DFD() = default;
  AssemblerHelmholtz2DFD(const AssemblerHelmholtz2DFD& o);
  AssemblerHelmholtz2DFD& operator=(const AssemblerHelmholtz2DFD& o);
  Assembler* copy() const override { return new AssemblerHelmholtz2DFD(*this); }

  void setPML(const PMLCubic2D* pml) { mPML.reset(static_cast<PMLCubic2D*>(pml->copy())); }
  void setM(const Function* m) { mM.reset(m->copy()); }
  void setOmega(double w) { mOmega = w; }
  double getOmega() const { return mOmega; }
  const Function* getM() const { return mM.get(); }
  const PMLCubic2D* getPML() const { return mPML.get(); }

  void setDomain(const std::vector<DoubleVector>& domain) override;
  void setN(const std::vector<UInt>& N) override;
  void setOrder(UInt order) override { mOrder = order; }
  UInt getOrder() const override { return mOrder; }
  double getMeshSize() const override;
  void getDOFPts(std::vector<DoubleVector>& res) const override;
  void assembleSystemMatrix(ComplexSparseMatrix& res) const override;
  void assembleRHS(const RHS* rhs, ComplexVector& res) const override;

  // grid description used by the device kernels
  int nxDofs() const { return (int)mN[0] - 1; }
  int nyDofs() const { return (int)mN[1] - 1; }
  double hx() const;
  double hy() const;
  double x0() const { return mCompDomain[0][0]; }
  double y0() const { return mCompDomain[1][0]; }
  const std::vector<DoubleVector>& intDomain() const { return mIntDomain; }
  const std::vector<DoubleVector>& compDomain() const { return mCompDomain; }
  double xCoord(int i) const { return mCompDomain[0][0] + (double)(UInt)(i + 1) * hx(); }
  double yCoord(int j) const { return mCompDomain[1][0] + (double)(UInt)(j + 1) * hy(); }
  // squared slowness at every DOF (index i*ny+j)
  void squaredSlowness(std::vector<double>& m) const;

 private:
  std::unique_ptr<PMLCubic2D> mPML;
  std::unique_ptr<Function> mM;
  double mOmega = 0;
  UInt mOrder = 2;
  std::vector<UInt> mN;
  std::vector<DoubleVector> mIntDomain, mCompDomain;
};

// Host evaluation of the cubic-PML stencil coefficients at one DOF; shared by
// the host CSC assembly and (as a __device__ copy) the GPU assembly kernel.
// coef[0] = centre, [1] = (i-1,j), [2] = (i+1,j), [3] = (i,j-1), [4] = (i,j+1).
void stencilCoefficients(double x, double y, const double intDom[2][2], const double compDom[2][2],
                         double C, double omega, double m, double hx, double hy, Complex coef[5]);

}  // namespace lsw
