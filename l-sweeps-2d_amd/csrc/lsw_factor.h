// FactorSet: a batch of structured-grid local problems (5-point operator),
// assembled, factorised and solved on the GPU.
//
// This is the LocalLinearSolver of the MI355X path (the reference uses one
// Eigen SparseLU per subdomain, src/LocalLinearSolverEigenSparseLU.cpp:33-76):
//   * assembly      -- one thread per DOF evaluates the cubic-PML stencil;
//   * factorisation -- multifrontal LU over a geometric nested-dissection
//                      tree, one workgroup per dense front, fronts of equal
//                      height in one launch; partial pivoting inside F11;
//   * solve         -- per tree height one launch of batched dense GEMVs
//                      (forward: W = F21 F11^{-1}; backward:
//                      K = [F11^{-1} | -F11^{-1} F12]) over many cells and
//                      right-hand sides at once.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <vector>

#include "lsw_device.h"
#include "lsw_nd.h"

namespace lsw {

// Fronts of a factorisation in precision `precision` are reduced in FP32
// (LSW_FP32 with LSW_FACTOR_FP32=1) or in FP64 (otherwise).
bool factorWorkFp32(int precision);
// Workspace budget of one factorisation: complex entries of the fronts of a
// height (and their panel buffers) held at once.
constexpr long long kFactorWsEntries = 512LL << 20;

// Per-launch timing of the batched solve kernels (HIP events on the launch
// stream) and their algorithmic bytes.
struct SolveProfile {
  double fwdMs = 0, bwdMs = 0;        // summed launch durations
  int fwdLaunches = 0, bwdLaunches = 0;
  double fwdBytes = 0, bwdBytes = 0;  // summed algorithmic bytes
  double fwdFlops = 0, bwdFlops = 0;  // summed algorithmic flops (real)
  // summed roofline times of the launches: each launch's larger of its bytes
  // at the HBM peak and its flops at the matrix-core peak of its precision
  double idealMs = 0;
};
// MI355X peaks the roofline times use: HBM3E bandwidth and dense
// matrix-core FP32 / FP64 rates
constexpr double kPeakHbmBytesPerS = 8.0e12;
constexpr double kPeakMfmaFp32 = 157.3e12;
constexpr double kPeakMfmaFp64 = 78.6e12;

struct Rect {
  int i0, i1, j0, j1;
};

// Node of any shape of a FactorSet as the solve kernels see it: one load
// gives everything a work item needs (index pools of all shapes concatenated).
struct SolveNode {
  long long fwdOff, bwdOff;  // W_p and K_p offsets within the cell's factor blob
  int e, b;
  int E, B, map;             // offsets into the pooled E, B and child-map arrays
  int gOff;                  // offset of g_p in a right-hand side's update buffer
  int g1, g2;                // children's update offsets (-1: no child)
  int pad[2];
};

// A task of a plan: one cell, up to kMaxRhs right-hand sides.
struct SolveTaskDev {
  long long facBase;  // the cell's factor offset
  int nrhs;
  int slot[kMaxRhs];
  int pad;
};

// One group of a grouped matrix-core launch: the entries of one node row
// chunk of up to four tasks of one cell, with everything the kernel needs
// in one record (one load).
struct GroupRec {
  SolveNode nd;
  long long facBase;
  int r0rows;  // r0 | rows << 16
  int ntasks;
  int task[4];
  int flags[4];
  int pad[2];
};

// One work item: rows [r0, r0+rows) of one node of one task.
// flags: kEnt* bits below.
constexpr int kEntChild1 = 1;   // child 1 visited by the forward pass
constexpr int kEntChild2 = 2;   // child 2 visited by the forward pass
constexpr int kEntVisited = 4;  // the node itself visited (its r_E may be nonzero)
constexpr int kEntRootRhs = 8;  // a one-chunk root: the backward entry forms r_E from X and the children's updates
struct SolveEntry {
  int task, node, r0rows, flags;  // r0rows = r0 | rows << 16
};

// Launch-shape choices of a solve plan (FactorSet::buildPlan).  defaults()
// are the built-in rules (with the LSW_LATENCY_WAVES and LSW_SOLVE_MFMA
// environment overrides); DDCheckerboard::autotune times alternatives on the
// problem itself and keeps the fastest per block width.
struct PlanShape {
  long long latencyWaves = 4096;  // packed launches of fewer wavefronts and rows <= latencyRows split columns
  int latencyRows = 16;
  long long wideBudget = 4096;    // heights of few large entries split columns while entries x W x 2 fit this
  int wideCap = 8;                // most wavefronts one entry's columns are split over (2..16; 16 only
                                  // for FP32 tasks of up to 4 right-hand sides)
  bool mfma = true;               // FP32 8-right-hand-side packed launches on the matrix cores
  int mfmaWideCols = 192;         // ... and of more than 32 rows and at least this many columns split
                                  // them over 2 wavefronts on the matrix cores even when there are many
  bool mfmaPair = true;           // packed matrix-core launches: two tasks' entries of the same node
                                  // rows share a wavefront and its factor loads
  bool mfmaGroup = true;          // packed matrix-core launches as groups: the entries of one node row
                                  // chunk of up to 4 tasks in one wavefront, columns in pipelined chunks,
  int grpMinRowsFwd = 32;         // ... for launches whose entries have more rows than this (forward)
  int grpMinRowsBwd = 8;          // ... (backward); fewer rows: the paired packed kernel
  bool grpLatency = true;        // ... also for the latency-bound launches of short rows (instead of
                                  // splitting their columns)
  static PlanShape defaults();
};

// Launch plan of one batch of solves (device work lists per tree height).
struct SolvePlan {
  struct Launch {
    long long off = 0;  // first entry in the plan's entry array
    int n = 0;          // entries
    int mode = 0;       // 0 packed: P = 64/RP entries side by side per wavefront,
                        // `wpb` independent wavefronts per workgroup, `rounds`
                        // entries per lane group; 1 wide: `waves` wavefronts
                        // split one entry's columns; 2 packed and 3 wide on the
                        // matrix cores (FP32 tasks of 8 right-hand sides)
    int rp = 64;        // lanes per entry (rows per pass) in packed mode
    int rounds = 1;
    int wpb = 1;
    int waves = 1;
    int pair = 0;       // matrix-core packed: entries paired by factor rows (1: RP 16, 2: RP 32)
    long long goff = 0; // mode 4 (grouped matrix-core): the launch's groups in the plan's group list
    int ngroups = 0;
    bool one = false;   // mode 4 with at most one column chunk per entry
    int cols = 0;       // largest column count (sizes the LDS)
    double bytes = 0;   // algorithmic bytes
    double flops = 0;   // algorithmic real flops (8 per complex multiply-add)
  };
  int ntasks = 0, maxNrhs = 1;
  int width = 1;            // right-hand sides per task in the kernels (taskWidth(maxNrhs))
  bool outOfPlace = false;  // built for separate right-hand side and solution buffers
  // rows of the task-major buffers (X, Xo, R, G) a solve of this plan needs
  long long taskRows() const { return (long long)ntasks * width; }
  // where each slot id of the tasks lives in the task-major buffers: task
  // t, lane q as t*width + q (-1: not in any task); host copy and device copy
  std::vector<int> slotTQ;
  int* dSlotTQ = nullptr;
  std::vector<SolveTask> hTasks;
  SolveTaskDev* dTasks = nullptr;  // these three live in dBlob
  SolveEntry* dEntries = nullptr;
  GroupRec* dGroups = nullptr;  // the groups of mode-4 launches
  void* dBlob = nullptr;
  // hostOnly: buildPlan keeps the device image of tasks, entries and slot
  // map in hBlob and allocates nothing; bind() points the plan at a copy of
  // hBlob that the caller placed in device memory
  bool hostOnly = false;
  std::vector<unsigned char> hBlob;
  size_t oEnt = 0, oTQ = 0, oGrp = 0;
  void bind(void* base);
  std::vector<Launch> fwd, bwd;    // in launch order
  double fwdBytes = 0, bwdBytes = 0;
  double fwdFlops = 0, bwdFlops = 0;
  SolvePlan() = default;
  SolvePlan(const SolvePlan&) = delete;
  SolvePlan& operator=(const SolvePlan&) = delete;
  ~SolvePlan();
};

class FactorSet {
 public:
  FactorSet() = default;
  ~FactorSet();
  FactorSet(const FactorSet&) = delete;
  FactorSet& operator=(const FactorSet&) = delete;

  // Add a cell described by its grid and PML geometry plus squared slowness
  // per DOF; stencils are assembled on the device.
  int addCell(const CellDev& c, const std::vector<double>& m);
  // Add a cell whose stencil (5*nx*ny complex, dir-major) is given on the host.
  int addCellWithStencil(const CellDev& c, const std::vector<double2>& coef);

  // Precision the solves will run in (sizes plans before factorisation).
  void setSolvePrecision(int precision) { mPrecision = precision; }
  // Build ND trees, upload descriptors, allocate, assemble stencils.
  void setup(hipStream_t s);
  // Multifrontal factorisation on the device in FP64 fronts; for LSW_FP32
  // the factors are then rounded to FP32 (LSW_FACTOR_FP32=1: FP32 fronts
  // with FP32-accumulated matrix-core updates storing FP32 factors).
  void factorize(int precision, hipStream_t s);
  bool factorized() const { return mFactorized; }
  int precision() const { return mPrecision; }

  // Solve every task (dtasks: device array of ntasks SolveTask).  X holds the
  // right-hand sides on entry and the solutions on exit (slot-major, stride
  // xStride complex entries); R and G are scratch with the same slot layout
  // (G stride gStride).  Element type follows precision(): double2 or float2.
  // Build the launch plan of a batch of tasks.  supports[t] lists rectangles
  // (i0, i1, j0, j1) of local grid points outside which every right-hand side
  // of task t is zero (empty list = anywhere).  The forward solve visits only
  // tree nodes whose region meets the support; the backward solve skips the
  // F11^{-1} r_E product of the nodes it did not visit.
  // bwdSupports (optional, one per task): rectangles where the solution is
  // needed; the backward solve then visits only nodes whose region meets
  // them (ancestors included: their regions contain their descendants').  An
  // empty list for a task means no solution values are needed at all.
  // outOfPlace: the plan will run with separate right-hand side and solution
  // buffers (solve(plan, X, Xo, ...), Xo != X); its roots then form r_E in
  // their backward entries, and the forward pass has no launch for them.
  // shape: launch-shape choices (nullptr: PlanShape::defaults()).
  void buildPlan(const std::vector<SolveTask>& tasks, const std::vector<std::vector<Rect>>& supports,
                 SolvePlan& plan, const std::vector<std::vector<Rect>>* bwdSupports = nullptr,
                 bool outOfPlace = false, const PlanShape* shape = nullptr) const;
  // Run a plan.  X holds the right-hand sides on entry and the solutions on
  // exit; R and G are scratch.  All four are laid out by task:
  // plan.taskRows() rows of xStride (X, R) or gStride (G) entries, the
  // plan.width right-hand sides of a task interleaved (value q of position
  // i of task t at (t*stride + i)*width + q; plan.slotTQ maps the tasks'
  // slot ids to t*width + q).  Element type follows precision(): double2 or
  // float2.  With prof, every launch is timed with HIP events (synchronises
  // at the end).
  void solve(const SolvePlan& plan, void* X, void* R, void* G, long long xStride, long long gStride,
             hipStream_t s, SolveProfile* prof = nullptr) const;
  // The same with the solutions written to Xo (same layout as X); X is only
  // read, so it still holds the right-hand sides afterwards.
  void solve(const SolvePlan& plan, const void* X, void* Xo, void* R, void* G, long long xStride, long long gStride,
             hipStream_t s, SolveProfile* prof = nullptr) const;

  int numCells() const { return (int)mCells.size(); }
  const CellDev& cell(int i) const { return mCells[i]; }
  const CellDev* deviceCells() const { return dCells; }
  const double2* deviceCoef() const { return dCoef; }
  long long maxDofs() const { return mMaxDofs; }
  long long maxG() const { return mMaxG; }
  long long factorEntries() const { return mFactorEntries; }
  long long factorBytes() const;
  const NdTree& tree(int shape) const { return *mTrees[shape]; }
  int numShapes() const { return (int)mTrees.size(); }
  int maxHeight() const { return mMaxHeight; }
  // right-hand sides one solve task may carry: 8 when the widest front's
  // staging (8 x front x element bytes) leaves room in the LDS, else 4
  int maxTaskRhs() const;
  int maxFront() const { return mMaxFront; }
  // pivot/breakdown status of the last factorisation (0 = ok)
  int lastInfo() const { return mInfo; }

 private:
  std::vector<CellDev> mCells;
  std::vector<std::vector<double>> mM;
  std::vector<std::vector<double2>> mCoefHost;
  std::vector<std::shared_ptr<const NdTree>> mTrees;
  std::map<std::pair<int, int>, int> mShapeOf;
  CellDev* dCells = nullptr;
  ShapeDev* dShapes = nullptr;
  // flattened solve descriptors
  std::vector<int> mShapeNodeBase;
  std::vector<SolveNode> mSolveNodes;  // host copy
  SolveNode* dSolveNodes = nullptr;
  int* dPoolE = nullptr;
  int* dPoolB = nullptr;
  int* dPoolMap = nullptr;
  DevArena mArena;  // cells, shapes, trees' index arrays, solve descriptors
  double2* dCoef = nullptr;
  double* dM = nullptr;
  double2* dFactor = nullptr;
  float2* dFactor32 = nullptr;
  long long mMaxDofs = 0, mMaxG = 0, mFactorEntries = 0;
  int mMaxHeight = 0, mMaxFront = 0;
  int mPrecision = LSW_FP64;
  bool mFactorized = false;
  int mInfo = 0;
  template <typename W2>
  void factorizeAs(hipStream_t s, W2*& factor, bool smallFp32);
};

}  // namespace lsw
