// Input / Problem (see lsw_input.h).
#include "lsw_input.h"
#include "lsw_output.h"

#include <hip/hip_runtime.h>

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "lsw_dd.h"
#include "lsw_gmres.h"
#include "lsw_model.h"
#include "lsw_solver.h"

namespace lsw {

// ------------------------------------------------------------------ reader
const CfgNode* CfgNode::find(const std::string& key) const {
  for (const auto& c : children)
    if (c.name == key) return &c;
  return nullptr;
}

bool CfgNode::lookupValue(const std::string& key, double& v) const {
  const CfgNode* c = find(key);
  if (!c || c->kind != kFloat) return false;
  v = c->f;
  return true;
}
bool CfgNode::lookupValue(const std::string& key, UInt& v) const {
  const CfgNode* c = find(key);
  if (!c || c->kind != kInt || c->i < 0) return false;
  v = (UInt)c->i;
  return true;
}
bool CfgNode::lookupValue(const std::string& key, std::string& v) const {
  const CfgNode* c = find(key);
  if (!c || c->kind != kString) return false;
  v = c->s;
  return true;
}
bool CfgNode::lookupValue(const std::string& key, bool& v) const {
  const CfgNode* c = find(key);
  if (!c || c->kind != kBool) return false;
  v = c->b;
  return true;
}

namespace {
class Reader {
 public:
  explicit Reader(const std::string& t) : s(t) {}
  CfgNode parseRoot() {
    CfgNode root;
    root.kind = CfgNode::kGroup;
    skip();
    while (p < s.size()) {
      root.children.push_back(parseSetting());
      skip();
    }
    return root;
  }

 private:
  const std::string& s;
  size_t p = 0;
  int line = 1;

  [[noreturn]] void fail(const std::string& m) {
    throw Error(LSW_ERR_ARG, "Parse error at line " + std::to_string(line) + ": " + m);
  }
  void skip() {
    while (p < s.size()) {
      const char c = s[p];
      if (c == '\n') {
        ++line;
        ++p;
      } else if (std::isspace((unsigned char)c)) {
        ++p;
      } else if (c == '#' || (c == '/' && p + 1 < s.size() && s[p + 1] == '/')) {
        while (p < s.size() && s[p] != '\n') ++p;
      } else if (c == '/' && p + 1 < s.size() && s[p + 1] == '*') {
        p += 2;
        while (p + 1 < s.size() && !(s[p] == '*' && s[p + 1] == '/')) {
          if (s[p] == '\n') ++line;
          ++p;
        }
        if (p + 1 >= s.size()) fail("unterminated comment");
        p += 2;
      } else {
        break;
      }
    }
  }
  std::string name() {
    skip();
    const size_t b = p;
    if (p < s.size() && (std::isalpha((unsigned char)s[p]) || s[p] == '*' || s[p] == '_')) {
      ++p;
      while (p < s.size() && (std::isalnum((unsigned char)s[p]) || s[p] == '_' || s[p] == '-' || s[p] == '*')) ++p;
    }
    if (p == b) fail("expected a setting name");
    return s.substr(b, p - b);
  }
  CfgNode parseSetting() {
    CfgNode n;
    const std::string nm = name();
    skip();
    if (p >= s.size() || (s[p] != '=' && s[p] != ':')) fail("expected '=' or ':' after " + nm);
    ++p;
    n = parseValue();
    n.name = nm;
    skip();
    if (p < s.size() && (s[p] == ';' || s[p] == ',')) ++p;
    return n;
  }
  CfgNode parseValue() {
    skip();
    if (p >= s.size()) fail("unexpected end of input");
    CfgNode n;
    const char c = s[p];
    if (c == '{') {
      ++p;
      n.kind = CfgNode::kGroup;
      skip();
      while (p < s.size() && s[p] != '}') {
        n.children.push_back(parseSetting());
        skip();
      }
      if (p >= s.size()) fail("unterminated group");
      ++p;
    } else if (c == '(' || c == '[') {
      const char close = c == '(' ? ')' : ']';
      n.kind = c == '(' ? CfgNode::kList : CfgNode::kArray;
      ++p;
      skip();
      while (p < s.size() && s[p] != close) {
        n.children.push_back(parseValue());
        skip();
        if (p < s.size() && s[p] == ',') {
          ++p;
          skip();
        }
      }
      if (p >= s.size()) fail("unterminated list");
      ++p;
      if (n.kind == CfgNode::kArray)
        for (auto& ch : n.children)
          if (ch.kind == CfgNode::kGroup || ch.kind == CfgNode::kList || ch.kind == CfgNode::kArray)
            fail("arrays hold scalars only");
    } else if (c == '"') {
      n.kind = CfgNode::kString;
      while (p < s.size() && s[p] == '"') {  // adjacent strings concatenate
        ++p;
        while (p < s.size() && s[p] != '"') {
          if (s[p] == '\\' && p + 1 < s.size()) {
            ++p;
            const char e = s[p];
            n.s += e == 'n' ? '\n' : e == 't' ? '\t' : e;
          } else {
            if (s[p] == '\n') ++line;
            n.s += s[p];
          }
          ++p;
        }
        if (p >= s.size()) fail("unterminated string");
        ++p;
        skip();
      }
    } else {
      const size_t b = p;
      while (p < s.size() && (std::isalnum((unsigned char)s[p]) || s[p] == '.' || s[p] == '-' || s[p] == '+'))
        ++p;
      std::string tok = s.substr(b, p - b);
      if (tok.empty()) fail("expected a value");
      std::string low = tok;
      for (auto& ch : low) ch = (char)std::tolower((unsigned char)ch);
      if (low == "true" || low == "false") {
        n.kind = CfgNode::kBool;
        n.b = low == "true";
      } else if (tok.find_first_of(".eE") != std::string::npos && low.rfind("0x", 0) != 0) {
        n.kind = CfgNode::kFloat;
        size_t used = 0;
        try {
          n.f = std::stod(tok, &used);
        } catch (...) {
          fail("bad number " + tok);
        }
        if (used != tok.size()) fail("bad number " + tok);
      } else {
        n.kind = CfgNode::kInt;
        if (!tok.empty() && (tok.back() == 'L' || tok.back() == 'l')) tok.pop_back();
        size_t used = 0;
        try {
          n.i = std::stoll(tok, &used, 0);
        } catch (...) {
          fail("bad integer " + tok);
        }
        if (used != tok.size()) fail("bad integer " + tok);
      }
    }
    return n;
  }
};
}  // namespace

CfgNode parseConfig(const std::string& text) {
  Reader r(text);
  return r.parseRoot();
}

// ------------------------------------------------------------------- Input
void Input::openFile(const std::string& filename) {
  std::ifstream in(filename);
  LSW_REQUIRE(in.good(), LSW_ERR_ARG, "I/O error while reading file " + filename);
  std::stringstream ss;
  ss << in.rdbuf();
  openText(ss.str());
}

void Input::openText(const std::string& text) {
  mRoot = parseConfig(text);
  mOpen = true;
}

UInt Input::getNumProblems() const {
  LSW_REQUIRE(mOpen, LSW_ERR_STATE, "No file opened for input");
  const CfgNode* pr = mRoot.find("Problems");
  LSW_REQUIRE(pr != nullptr, LSW_ERR_ARG, "No Problems defined in input file");
  return (UInt)pr->getLength();
}

namespace {
std::vector<DoubleVector> readDomain(const CfgNode& root, const std::string& name, UInt dim) {
  const CfgNode* ds = root.find("Domains");
  LSW_REQUIRE(ds != nullptr, LSW_ERR_ARG, "No Domains defined in input file");
  const CfgNode* d = ds->find(name);
  LSW_REQUIRE(d != nullptr, LSW_ERR_ARG, "Unknown Domain prescribed");
  UInt ddim = 0;
  LSW_REQUIRE(d->lookupValue("Dim", ddim), LSW_ERR_ARG, "No Dim prescribed in Domain");
  LSW_REQUIRE(ddim == dim, LSW_ERR_ARG, "Inconsistent dimensions in Domain");
  double ax, bx, ay, by;
  LSW_REQUIRE(d->lookupValue("ax", ax), LSW_ERR_ARG, "No ax prescribed in Domain");
  LSW_REQUIRE(d->lookupValue("bx", bx), LSW_ERR_ARG, "No bx prescribed in Domain");
  LSW_REQUIRE(d->lookupValue("ay", ay), LSW_ERR_ARG, "No ay prescribed in Domain");
  LSW_REQUIRE(d->lookupValue("by", by), LSW_ERR_ARG, "No by prescribed in Domain");
  return {{ax, bx}, {ay, by}};
}
}  // namespace

// src/Input.cpp:100-151 (problem), :154-247 (RHS), :250-399 (solver),
// :470-623 (CheckerboardDD), :626-706 (domain), :709-792 (squared
// slowness), :795-857 (assembler).
ProblemSpec Input::readProblem(UInt probID) const {
  LSW_REQUIRE(mOpen, LSW_ERR_STATE, "No file opened for input");
  const CfgNode* pr = mRoot.find("Problems");
  LSW_REQUIRE(pr != nullptr, LSW_ERR_ARG, "No Problems defined in input file");
  LSW_REQUIRE(probID < (UInt)pr->getLength(), LSW_ERR_ARG, "Invalid problem ID");
  const CfgNode& prob = (*pr)[(int)probID];
  ProblemSpec sp;
  // ---- domain decomposition
  std::string ddType;
  UInt ddID = 0;
  LSW_REQUIRE(prob.lookupValue("DomainDecomposition", ddType), LSW_ERR_ARG,
              "No DomainDecomposition prescribed in problem!");
  LSW_REQUIRE(prob.lookupValue("DomainDecompositionID", ddID), LSW_ERR_ARG,
              "No DomainDecompositionID prescribed in problem!");
  LSW_REQUIRE(ddType == "CheckerboardDD", LSW_ERR_ARG, "Invalid DomainDecomposition prescribed");
  const CfgNode* dds = mRoot.find("DomainDecompositions");
  LSW_REQUIRE(dds != nullptr, LSW_ERR_ARG, "No DomainDecompositions defined in input file");
  const CfgNode* cdds = dds->find("CheckerboardDD");
  LSW_REQUIRE(cdds != nullptr, LSW_ERR_ARG, "No CheckerboardDDs defined in input file");
  LSW_REQUIRE(ddID < (UInt)cdds->getLength(), LSW_ERR_ARG, "Invalid DomainDecompositionID prescribed!");
  const CfgNode& cdd = (*cdds)[(int)ddID];
  UInt dim = 0;
  LSW_REQUIRE(cdd.lookupValue("Dim", dim), LSW_ERR_ARG, "No Dim prescribed in CheckerboardDD");
  LSW_REQUIRE(dim == 2, LSW_ERR_ARG, "Dim can only be 2!");
  std::string domName;
  LSW_REQUIRE(cdd.lookupValue("Domain", domName), LSW_ERR_ARG, "No Domain prescribed in CheckerboardDD");
  sp.domain = readDomain(mRoot, domName, dim);
  std::string pde;
  LSW_REQUIRE(cdd.lookupValue("PDE", pde), LSW_ERR_ARG, "No PDE prescribed in CheckerboardDD");
  LSW_REQUIRE(pde == "Helmholtz", LSW_ERR_ARG, "Unknown PDE prescribed!");
  LSW_REQUIRE(cdd.lookupValue("Omega", sp.omega), LSW_ERR_ARG, "No omega prescribed!");
  LSW_REQUIRE(cdd.lookupValue("MinWavelengthsInPML", sp.minWavelengthsInPML), LSW_ERR_ARG,
              "No MinWavelengthsInPML prescribed!");
  std::string ssName;
  LSW_REQUIRE(cdd.lookupValue("SquaredSlowness", ssName), LSW_ERR_ARG, "No SquaredSlowness prescribed!");
  {
    const CfgNode* sss = mRoot.find("SquaredSlownesses");
    LSW_REQUIRE(sss != nullptr, LSW_ERR_ARG, "No SquaredSlownesses defined in input file");
    const CfgNode* ss = sss->find(ssName);
    LSW_REQUIRE(ss != nullptr, LSW_ERR_ARG, "Unknown SquaredSlowness prescribed");
    LSW_REQUIRE(ss->lookupValue("Type", sp.mediumType), LSW_ERR_ARG, "No Type prescribed in SquaredSlowness");
    if (sp.mediumType == "CSV") {
      LSW_REQUIRE(ss->lookupValue("Filename", sp.csvFile), LSW_ERR_ARG, "No Filename prescribed");
      UInt d2 = 0;
      LSW_REQUIRE(ss->lookupValue("Dim", d2), LSW_ERR_ARG, "No Dim prescribed");
      std::string fdom;
      LSW_REQUIRE(ss->lookupValue("Domain", fdom), LSW_ERR_ARG, "No Domain prescribed");
      sp.csvDomain = readDomain(mRoot, fdom, d2);
    } else {
      LSW_REQUIRE(sp.mediumType == "Constant", LSW_ERR_ARG, "Unknown SquaredSlowness type");
    }
  }
  LSW_REQUIRE(cdd.lookupValue("Order", sp.order), LSW_ERR_ARG, "No Order prescribed in CheckerboardDD");
  LSW_REQUIRE(cdd.exists("NX") && cdd.exists("NY"), LSW_ERR_ARG, "Invalid way to prescribe N's");
  cdd.lookupValue("NX", sp.NX);
  cdd.lookupValue("NY", sp.NY);
  LSW_REQUIRE(cdd.lookupValue("NCellsX", sp.nCellsX), LSW_ERR_ARG, "No NCellsX prescribed in CheckerboardDD");
  LSW_REQUIRE(cdd.lookupValue("NCellsY", sp.nCellsY), LSW_ERR_ARG, "No NCellsY prescribed in CheckerboardDD");
  std::string solver;
  LSW_REQUIRE(cdd.lookupValue("Solver", solver), LSW_ERR_ARG, "No solver prescribed!");
  LSW_REQUIRE(solver == "HIP" || solver == "Eigen", LSW_ERR_ARG, "Unknown solver prescribed!");
  std::string prec;
  if (cdd.lookupValue("Precision", prec)) {
    LSW_REQUIRE(prec == "FP64" || prec == "FP32", LSW_ERR_ARG, "Unknown Precision (FP64 or FP32)");
    sp.precision = prec == "FP32" ? LSW_FP32 : LSW_FP64;
  }
  // ---- RHS
  std::string rhsName;
  LSW_REQUIRE(prob.lookupValue("RHS", rhsName), LSW_ERR_ARG, "No RHS prescribed in problem!");
  const CfgNode* rhss = mRoot.find("RHSs");
  LSW_REQUIRE(rhss != nullptr, LSW_ERR_ARG, "No RHSs prescribe in input file!");
  const CfgNode* srcs = rhss->find(rhsName);
  LSW_REQUIRE(srcs != nullptr, LSW_ERR_ARG, "Invalid RHS type prescribed");
  for (int k = 0; k < srcs->getLength(); ++k) {
    const CfgNode& s = (*srcs)[k];
    UInt sd = 0;
    double x, y;
    LSW_REQUIRE(s.lookupValue("Dim", sd), LSW_ERR_ARG, "No Dim prescribed for source point");
    LSW_REQUIRE(s.lookupValue("x", x), LSW_ERR_ARG, "No x prescribed in source point");
    LSW_REQUIRE(s.lookupValue("y", y), LSW_ERR_ARG, "No y prescribed in source point");
    sp.sources.push_back({x, y});
  }
  // ---- global solver
  std::string solName;
  LSW_REQUIRE(prob.lookupValue("Solver", solName), LSW_ERR_ARG, "No Solver prescribed in problem!");
  const CfgNode* gss = mRoot.find("GlobalSolvers");
  LSW_REQUIRE(gss != nullptr, LSW_ERR_ARG, "No GlobalSolvers defined in input file");
  const CfgNode* gs = gss->find(solName);
  LSW_REQUIRE(gs != nullptr, LSW_ERR_ARG, "Invalid Solver type prescribed!");
  LSW_REQUIRE(gs->lookupValue("Type", sp.solverType), LSW_ERR_ARG, "No Type prescribed in Solver");
  LSW_REQUIRE(sp.solverType == "GMRES" || sp.solverType == "GMRESHIP" || sp.solverType == "FGMRESHIP",
              LSW_ERR_ARG, "Unknown Solver type!");
  double tol = 1.0e-6;
  gs->lookupValue("Tol", tol);
  UInt maxIter = 0;
  gs->lookupValue("MaxIter", maxIter);
  bool verbose = true;
  gs->lookupValue("Verbose", verbose);
  if (tol != 0.0) sp.tol = tol;
  if (maxIter != 0) sp.maxIter = maxIter;
  sp.verbose = verbose;
  UInt restart = 0;
  if (gs->lookupValue("Restart", restart) && restart > 0) sp.restart = restart;
  // ---- output (Input::readOutput, src/Input.cpp:402-467): "" = none
  std::string outName;
  LSW_REQUIRE(prob.lookupValue("Output", outName), LSW_ERR_ARG, "No Output prescribed in problem!");
  sp.output = outName;
  if (!outName.empty()) {
    const CfgNode* outs = mRoot.find("Outputs");
    LSW_REQUIRE(outs != nullptr, LSW_ERR_ARG, "No Outputs defined in input file");
    const CfgNode* o = outs->find(outName);
    LSW_REQUIRE(o != nullptr, LSW_ERR_ARG, "Invalid output type prescribed!");
    std::string type;
    LSW_REQUIRE(o->lookupValue("Type", type), LSW_ERR_ARG, "No Type prescribed for Output");
    LSW_REQUIRE(type == "VTK", LSW_ERR_ARG, "Unknown output type!");
    LSW_REQUIRE(o->lookupValue("Filename", sp.vtkFile), LSW_ERR_ARG, "No Filename prescribed!");
  }
  return sp;
}

// ----------------------------------------------------------------- Problem
Problem::Problem() = default;
Problem::~Problem() = default;

void Problem::read(const Input& in, UInt probID) {
  mSpec = in.readProblem(probID);
  const auto t0 = std::chrono::steady_clock::now();
  std::unique_ptr<Function> m;
  if (mSpec.mediumType == "CSV") {
    auto* f = new FunctionCSV();
    f->setFile(mSpec.csvFile, mSpec.csvDomain);
    m.reset(f);
  } else {
    m.reset(new FunctionConstant(1.0));
  }
  PMLCubic2D pml;
  pml.setAbsorptionCoeff(std::log(mSpec.omega));
  const auto& dom = mSpec.domain;
  const double wl1 = (dom[0][1] - dom[0][0]) * 2 * kPi / mSpec.omega;
  const double wl2 = (dom[1][1] - dom[1][0]) * 2 * kPi / mSpec.omega;
  double wl = std::min(wl1, wl2);
  wl /= std::sqrt(m->getMinReal());
  pml.setWidth(mSpec.minWavelengthsInPML * wl);
  AssemblerHelmholtz2DFD as;
  as.setPML(&pml);
  as.setOmega(mSpec.omega);
  m->setDomain(dom);
  as.setM(m.get());
  as.setOrder(mSpec.order);
  LocalLinearSolverHIP proto(mSpec.precision);
  mDD = std::make_unique<DDCheckerboard>();
  mDD->setUp(&as, &proto, dom, {mSpec.nCellsX, mSpec.nCellsY}, {mSpec.NX, mSpec.NY});
  mSetup = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

ProblemResult Problem::solve(bool verbose) {
  LSW_REQUIRE(mDD != nullptr, LSW_ERR_STATE, "Problem::solve before read");
  ProblemResult r;
  r.setupSeconds = mSetup;
  r.dofs = mDD->N();
  using clk = std::chrono::steady_clock;
  if (verbose) std::printf("Facorize problem...\n");
  auto t0 = clk::now();
  mDD->factorize();
  LSW_HIP(hipStreamSynchronize(mDD->stream()));
  r.factorizeSeconds = std::chrono::duration<double>(clk::now() - t0).count();
  if (verbose) std::printf("Assemble RHS...\n");
  const long long n = mDD->N();
  double2 *db = nullptr, *dx = nullptr;
  LSW_HIP(devMalloc(&db, n * sizeof(double2)));
  LSW_HIP(devMalloc(&dx, n * sizeof(double2)));
  try {
    mDD->assembleRHSDev(mSpec.sources, db, mDD->stream());
    GlobalIterativeSolverGMRESHIP g;
    g.setParent(mDD.get());
    g.setTol(mSpec.tol);
    g.setMaxIter((int)mSpec.maxIter);
    g.setRestart((int)mSpec.restart);
    g.setVerbose(mSpec.verbose && verbose);
    g.setMode(mSpec.solverType == "FGMRESHIP" ? GlobalIterativeSolverGMRESHIP::kFlexible
                                               : GlobalIterativeSolverGMRESHIP::kReference);
    if (verbose) std::printf("Solving...\n");
    t0 = clk::now();
    GmresInfo info;
    g.solveDev(db, dx, &info, mDD->stream());
    LSW_HIP(hipStreamSynchronize(mDD->stream()));
    r.gmresSeconds = std::chrono::duration<double>(clk::now() - t0).count();
    r.iterations = info.iterations;
    r.estimate = info.estimate;
    r.trueResidual = info.trueResidual;
    mSolution.resize((size_t)n);
    LSW_HIP(hipMemcpy(mSolution.data(), dx, n * sizeof(double2), hipMemcpyDeviceToHost));
  } catch (...) {
    (void)devFree(db);
    (void)devFree(dx);
    throw;
  }
  (void)devFree(db);
  (void)devFree(dx);
  if (verbose) {
    std::printf("Time to factorize system: %g\n", r.factorizeSeconds);
    std::printf("Time to run the GMRES: %g\n", r.gmresSeconds);
  }
  // Problem::solve writes the output after the timers (src/Problem.cpp:147-156)
  if (!mSpec.vtkFile.empty()) {
    writeVTK(mSpec.vtkFile + ".vtk", *mDD, mSolution);
    writeGrid(mSpec.vtkFile + ".lswgrid", *mDD, mSolution);
  }
  return r;
}

}  // namespace lsw
