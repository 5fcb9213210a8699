// The domain-decomposition layer of the MI355X path: class DD (include/DD.h:
// 14-114) and DDCheckerboard (include/DDCheckerboard.h), with every
// checkerboard cell resident on one GPU.
//
// Vectors of the global problem are held on the device as one grid function
// over the global DOF grid (x-major, y fastest: g = gi*Gy + gj).  The
// reference's per-rank "owned DOF" vectors (mValsToKeep*, reduce/expand and
// the shifted/unshifted conversions, src/DDCheckerboard.cpp:434-809,
// 2995-3515) are exact re-layouts of that grid function, so on one device
// they become index ranges: every trace, owned set and L-trace of a cell is a
// range of local grid columns or rows.
//
// applyPrecond (src/DDCheckerboard.cpp:687-739) runs the four windowed DDs
// and, per DD, stage 1, the four 1-D sweeps and the four L-sweeps.  Cells on
// one front (a column, a row or an anti-diagonal) are independent, so every
// front of every sweep of every DD is one batch of local solves.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <map>
#include <tuple>
#include <memory>
#include <string>
#include <vector>

#include "lsw_factor.h"
#include "lsw_model.h"

namespace lsw {

class LocalLinearSolver;
class RHS;

// ------------------------------------------------------------------ DD base
class DD {
 public:
  virtual ~DD() = default;
  virtual void factorize() = 0;
  // owned-DOF vectors per cell, cells in row-major order (mCells order)
  virtual void assembleRHS(const RHS* rhs, std::vector<ComplexVector>& res) const = 0;
  virtual void applyPrecond(const std::vector<ComplexVector>& x, std::vector<ComplexVector>& res) const = 0;
  virtual void applySystemMatrix(const std::vector<ComplexVector>& x,
                                 std::vector<ComplexVector>& res) const = 0;
  virtual UInt getDim() const { return 2; }
  virtual double getMeshSize() const = 0;
  virtual UInt getNCells() const = 0;
};

// Host description of one cell (the reference's SubdomainStructured).
struct CellGeom {
  int I = 0, J = 0;
  AssemblerHelmholtz2DFD asmb;
  int nx = 0, ny = 0;
  double lims[2][10] = {};
  int vlr[2][8] = {}, vud[2][8] = {};
  int ox = 0, oy = 0;
  double meshSize = 0;
};

// Per-slot description of one local solve in a sweep step (device POD).
struct SlotDev {
  int d, kind, cell, slot;
  int setCol[2], setColLo[2], setColHi[2];
  int setRow[2], setRowLo[2], setRowHi[2];
  long long colVal[2], rowVal[2];
  int owCol, owRow;
  long long owColVal, owRowVal;
  int nOut;
  int outIsRow[8], outIdx[8], outAdd[8];
  long long outDst[8];
  int grp;       // (d, cell) index d*ncells + cell of the summed right-hand sides
  int deferred;  // 1: the final solve of (d, cell): windowed RHS + summed sweep RHS
};

struct GroupDev {
  int d, cell, n, pad;
  int slot[4];
};

struct CopyOp {
  long long dst, src;
};

class DDCheckerboard : public DD {
 public:
  DDCheckerboard();
  ~DDCheckerboard() override;

  // src/DDCheckerboard.cpp:87-645.  domain = [[ax,bx],[ay,by]];
  // nCells = (NCellsX, NCellsY) sizing rows and columns; nElsInCell = (NX, NY).
  // The solver prototype selects the factor precision (LocalLinearSolverHIP).
  void setUp(const Assembler* assembler, const LocalLinearSolver* solver,
             std::vector<DoubleVector> domain, std::vector<UInt> nCells, std::vector<UInt> nElsInCell);
  // The host half of setUp: geometry, index sets and ownership only (no
  // device work).  setUp = setUpGeometry + device initialisation.
  void setUpGeometry(const Assembler* assembler, const LocalLinearSolver* solver,
                     std::vector<DoubleVector> domain, std::vector<UInt> nCells,
                     std::vector<UInt> nElsInCell);
  // The device half of setUp after setUpGeometry (the assembler, its PML and
  // medium given there must still be alive).
  void setUpDevice();
  bool deviceReady() const { return mFactors != nullptr; }
  // Replay applyPrecondDev from captured HIP graphs (off by default; the
  // LSW_GRAPHS environment variable turns it on).  Capture needs the stream
  // to itself: no other thread may issue legacy-stream work (synchronous
  // copies, allocation with implicit synchronisation) while a graph is being
  // captured, which is why the server keeps this off.
  void setGraphs(bool on) { mUseGraphs = on; if (!on) clearGraphs(); }
  bool graphs() const { return mUseGraphs; }
  // Geometry mode, set before setUp: false = the reference's spacing rule,
  // true = spacing per axis from that axis' cell count (rectangular domains).
  void setGeneralGeometry(bool g) { mGeneral = g; }

  void factorize() override;
  void assembleRHS(const RHS* rhs, std::vector<ComplexVector>& res) const override;
  void applyPrecond(const std::vector<ComplexVector>& x, std::vector<ComplexVector>& res) const override;
  void applySystemMatrix(const std::vector<ComplexVector>& x,
                         std::vector<ComplexVector>& res) const override;
  double getMeshSize() const override { return mMeshSize; }
  UInt getNCells() const override { return (UInt)mCells.size(); }

  // ---- device API (global grid vectors, complex double interleaved) ----
  int Gx() const { return mGx; }
  int Gy() const { return mGy; }
  long long N() const { return (long long)mGx * mGy; }
  // y = M^{-1} x for nrhs right-hand sides (x, y: nrhs consecutive global
  // vectors).  Every local solve of a front step shares one read of the
  // cell's factors among up to kMaxRhs right-hand sides.
  // replay = false runs the launches directly (for one-shot buffers whose
  // addresses will not recur).
  void applyPrecondDev(const double2* x, double2* y, hipStream_t s, int nrhs = 1, bool replay = true) const;
  // Same, timing every batched-solve launch with HIP events (synchronises
  // after each sweep step; use outside timed regions).
  void applyPrecondProfiled(const double2* x, double2* y, SolveProfile& prof, int nrhs = 1) const;
  // Times M^{-1} for blocks of nrhs right-hand sides under alternative launch
  // shapes (PlanShape) and keeps the fastest for that block width.  Uses the
  // DD's stream and synchronises; not for timed regions.
  struct TuneResult {
    double defaultMs = 0, bestMs = 0;  // per application
    int tried = 0;
    PlanShape shape;
  };
  TuneResult autotune(int nrhs) const;
  // Sizes the sweep buffers for blocks of up to nrhs right-hand sides now
  // (otherwise they grow on first use).
  void reserveBlock(int nrhs) const { ensureBuffers(nrhs); }
  // The narrowest block width in [lo, hi] whose solve plans are built, or 0.
  int builtWidth(int lo, int hi) const;
  // Device bytes of this DD once factorised, serving blocks of nrhs
  // right-hand sides (a host estimate from the nested-dissection trees of the
  // cell shapes; valid after setUpGeometry): steady = factors in the solve
  // precision + stencils + sweep buffers; peak adds what factorisation holds
  // for a while (FP64 factors before the FP32 copy, front workspace).
  // The trees it builds for the estimate are shared (sharedTree); keep, if
  // given, receives them so that the caller can hold them for a later set-up.
  void deviceBytesEstimate(int precision, int nrhs, double* steady, double* peak,
                           std::vector<std::shared_ptr<const NdTree>>* keep = nullptr) const;
  // y = A x for nrhs vectors
  void applyADevBlock(const double2* x, double2* y, int nrhs, hipStream_t s) const;
  // y = A x (global operator)
  void applyADev(const double2* x, double2* y, hipStream_t s) const;
  // global RHS of point sources (summed, RHSSrcPts semantics)
  void assembleRHSDev(const std::vector<DoubleVector>& srcs, double2* b, hipStream_t s) const;
  // global DOF coordinates
  double gx(int i) const { return mGlobal.asmb.xCoord(i); }
  double gy(int j) const { return mGlobal.asmb.yCoord(j); }
  hipStream_t stream() const { return mStream; }
  int precision() const { return mPrecision; }
  const FactorSet& factors() const { return *mFactors; }
  const CellGeom& cellGeom(int i) const { return mCells[i]; }
  int numCells() const { return (int)mCells.size(); }
  int nRows() const { return mNR; }
  int nCols() const { return mNC; }
  long long localSolvesPerPrecond() const { return mSolvesPerPrecond; }
  // algorithmic bytes moved by the local solves of one applyPrecond of one
  // right-hand side (builds the single-right-hand-side plans if need be)
  double solveBytesPerPrecond() const;
  // owned global indices of cell c (unshifted), in local DOF order
  void ownedGlobalIndices(int c, std::vector<long long>& idx) const;
  // global vector <-> per-cell owned vectors
  void gatherOwned(const std::vector<Complex>& g, std::vector<ComplexVector>& res) const;
  void scatterOwned(const std::vector<ComplexVector>& x, std::vector<Complex>& g) const;

 private:
  void buildSchedule();
  void allocBuffers();
  void applyPrecondImpl(const double2* x, double2* y, hipStream_t s, SolveProfile* prof, int nrhs) const;
  void ensureBuffers(int nrhs) const;
  long long taskRowsFor(int nrhs) const;
  std::vector<Rect> slotSupport(const SlotDev& s) const;
  void initDevice();
  std::vector<unsigned char> mColT, mRowT;

  std::vector<DoubleVector> mDomain;
  bool mGeneral = false;
  int mNR = 0, mNC = 0;
  int mGx = 0, mGy = 0;
  double mMeshSize = 0;
  int mPrecision = LSW_FP64;
  std::vector<CellGeom> mCells;
  CellGeom mGlobal;
  std::unique_ptr<FactorSet> mFactors;
  hipStream_t mStream = nullptr;
  // the stream of a call (its own for NULL); notes a caller's stream other
  // than its own, which the destructor then drains with the whole device
  hipStream_t useStream(hipStream_t s) const {
    if (!s) return mStream;
    if (s != mStream) mForeignStream = true;
    return s;
  }
  // Work the DD queues on its own stream for later calls (the set-up's
  // stencil assembly, the clears of grown buffers) is followed by an event;
  // a call on another stream waits for it on the host first (ownWorkDone),
  // which happens at most a few times per DD.
  void markOwnWork() const;
  void ownWorkDone(hipStream_t s) const;
  mutable bool mForeignStream = false;
  mutable hipEvent_t mOwnEvent = nullptr;
  mutable bool mOwnPending = false;
  DevArena mArena;  // step slot, group and copy lists, trace classes, owner tables

  // global operator
  CellDev mGlobalDev{};
  double2* dGlobalCoef = nullptr;
  // window classification of global columns / rows (1 = trace column/row)
  unsigned char* dColTrace = nullptr;
  unsigned char* dRowTrace = nullptr;

  // schedule
  struct Step {
    SlotDev* dSlots = nullptr;
    int nSlots = 0;
    std::vector<SlotDev> hSlots;               // host copy (single right-hand side)
    std::vector<std::vector<Rect>> hSupport;   // right-hand-side support per slot
    // solve plans per block width (built on first use)
    mutable std::map<int, std::unique_ptr<SolvePlan>> plans;
    GroupDev* dGroups = nullptr;
    int nGroups = 0;
    CopyOp* dCopies = nullptr;  // trace copies to run after this step
    int nCopies = 0;
    bool deferred = false;      // the final full solves of every (d, cell)
    std::vector<std::vector<Rect>> hNeed;  // solution lines needed per slot (traces)
  };
  std::vector<Step> mSteps;
  const SolvePlan& planFor(const Step& st, int nrhs) const;
  // launch shapes per block width (PlanShape::defaults() until autotuned)
  mutable std::map<int, PlanShape> mShapes;
  const PlanShape& shapeFor(int nrhs) const;
  void dropPlans(int nrhs) const;
  int mMaxSlots = 0;
  int mTmax = 0;
  long long mTraceEntries = 0;
  long long mSolvesPerPrecond = 0;
  mutable double mSolveBytesPerPrecond = -1;  // < 0: not computed yet

  // buffers (element type follows precision), sized for mBufRhs right-hand sides
  mutable void* dX = nullptr;   // right-hand sides of the local solves (slot-major)
  mutable void* dXs = nullptr;  // their solutions (same layout)
  mutable void* dR = nullptr;
  mutable void* dG = nullptr;
  mutable void* dT = nullptr;
  mutable double2* dY = nullptr;  // 4 accumulators of N entries per right-hand side
  mutable void* dF = nullptr;     // summed sweep right-hand sides per (d, cell) and right-hand side
  std::vector<std::pair<std::string, double>> mSetupLog;  // set-up phase times (LSW_SETUP_LOG prints them)
  mutable int mBufRhs = 0;
  bool mOwnedPartition = false;  // owned windows cover every global point once per DD
  // Owner tables of the final gather (valid when mGatherFinal): per DD axis
  // half, the owning cell column / row and local index of every global
  // column / row, then the cell index of each (row, column) of the board.
  std::vector<int> mOwnHost;
  int* dOwn = nullptr;
  bool mGatherFinal = false;  // final solutions gathered straight into the output
  bool mStage1InF = true;     // stage 1 may build and solve in F (LSW_STAGE1_IN_F=0: never)
  mutable long long mTaskRows = 0;  // rows of dR / dG (task-major solve scratch)
  long long mXs = 0, mGs = 0;

  // captured applyPrecondDev graphs per (x, y, nrhs, stream)
  struct GraphKey {
    const void* x;
    const void* y;
    int nrhs;
    hipStream_t s;
    bool operator<(const GraphKey& o) const {
      return std::tie(x, y, nrhs, s) < std::tie(o.x, o.y, o.nrhs, o.s);
    }
  };
  static constexpr size_t kMaxGraphs = 96;
  bool mUseGraphs = std::getenv("LSW_GRAPHS") != nullptr;
  mutable std::map<GraphKey, hipGraphExec_t> mGraphs;
  void clearGraphs() const;
};

}  // namespace lsw
