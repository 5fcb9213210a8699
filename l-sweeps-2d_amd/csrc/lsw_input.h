// Input (include/Input.h:44-76) and Problem (include/Problem.h:44-47): the
// reference's libconfig input file and its driver.
//
// libconfig++ is not available here, so the file format subset the reference
// uses is parsed by a small reader: settings "name = value;" (or ':'),
// groups {...}, lists (...), arrays [...], strings, integers, floats and
// booleans, with '#', '//' and '/* */' comments.  As with libconfig's
// lookupValue without auto-conversion (src/Input.cpp passim), a Double must
// be written with a decimal point and an integer without one.
//
// Selectors: CheckerboardDD Solver = "HIP" (or the reference's "Eigen", which
// selects the same FP64 multifrontal solver), optional Precision = "FP32";
// GlobalSolvers Type = "GMRES" / "GMRESHIP" (reference GMRES) or "FGMRESHIP"
// (flexible GMRES, optional Restart).
#pragma once

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "lsw_common.h"

namespace lsw {

struct CfgNode {
  enum Kind { kInt, kFloat, kBool, kString, kGroup, kList, kArray };
  Kind kind = kGroup;
  std::string name;
  long long i = 0;
  double f = 0;
  bool b = false;
  std::string s;
  std::vector<CfgNode> children;

  const CfgNode* find(const std::string& key) const;
  bool exists(const std::string& key) const { return find(key) != nullptr; }
  int getLength() const { return (int)children.size(); }
  const CfgNode& operator[](int k) const { return children.at(k); }
  bool lookupValue(const std::string& key, double& v) const;
  bool lookupValue(const std::string& key, UInt& v) const;
  bool lookupValue(const std::string& key, std::string& v) const;
  bool lookupValue(const std::string& key, bool& v) const;
};

// Parse libconfig text; throws Error(LSW_ERR_ARG) with line information.
CfgNode parseConfig(const std::string& text);

class DD;
class DDCheckerboard;
class RHS;
class GlobalIterativeSolver;

struct ProblemSpec {
  // CheckerboardDD
  std::vector<DoubleVector> domain;
  UInt nCellsX = 0, nCellsY = 0, NX = 0, NY = 0, order = 2;
  double omega = 0, minWavelengthsInPML = 0;
  int precision = LSW_FP64;
  std::string mediumType = "Constant";
  std::string csvFile;
  std::vector<DoubleVector> csvDomain;
  // RHS
  std::vector<DoubleVector> sources;
  // solver
  std::string solverType = "GMRES";
  double tol = 1e-6;   // default applied when Tol is absent or zero (src/Input.cpp:367-385)
  UInt maxIter = 40;   // class default (src/GlobalIterativeSolverGMRES.cpp:9)
  UInt restart = 30;
  bool verbose = true;
  std::string output;
  std::string vtkFile;  // Outputs.<output>.Filename when Type = "VTK" (no ".vtk")
};

class Input {
 public:
  void openFile(const std::string& filename);
  void openText(const std::string& text);
  UInt getNumProblems() const;
  // All parameters of problem probID, checked as src/Input.cpp:100-467 does.
  ProblemSpec readProblem(UInt probID) const;

 private:
  CfgNode mRoot;
  bool mOpen = false;
};

struct ProblemResult {
  double setupSeconds = 0, factorizeSeconds = 0, gmresSeconds = 0;
  int iterations = 0;
  double estimate = 0, trueResidual = 0;
  long long dofs = 0;
};

// Problem::read / Problem::solve (src/Problem.cpp:70-158) on one GPU.
class Problem {
 public:
  Problem();
  ~Problem();
  void read(const Input& in, UInt probID);
  // Factorize, assemble the RHS, run the global solver; prints the two
  // reference timers when verbose.  The solution (global grid vector) is kept.
  ProblemResult solve(bool verbose = true);
  const std::vector<Complex>& solution() const { return mSolution; }
  const DDCheckerboard& dd() const { return *mDD; }
  // hands the DD over to the caller (the Problem is then empty)
  std::unique_ptr<DDCheckerboard> takeDD() { return std::move(mDD); }
  const ProblemSpec& spec() const { return mSpec; }

 private:
  ProblemSpec mSpec;
  std::unique_ptr<DDCheckerboard> mDD;
  double mSetup = 0;
  std::vector<Complex> mSolution;
};

}  // namespace lsw
