// Request serving over the L-sweeps solve path.
//
// A Server takes solve requests (a registered problem = velocity model,
// frequency and board, plus 1..n point sources) from any thread and runs them
// on one GPU with continuous batching:
//   * factor cache   -- one factorised DDCheckerboard per problem, kept on the
//                       device between requests and keyed by the problem
//                       (model, omega); the reference re-reads and
//                       re-factorises every problem (src/Problem.cpp:70-91);
//   * admission      -- a problem is loaded only when its estimated device
//                       bytes (peak while factorising, steady afterwards) fit
//                       the budget; idle resident problems are evicted in
//                       least-recently-used order to make room, and requests
//                       wait while nothing can be evicted;
//   * merging        -- every source of every request on one problem becomes a
//                       right-hand side of that problem's GmresSlots, so
//                       sources that share factors share each sweep;
//   * continuous     -- right-hand sides join free slots between Arnoldi
//     batching          steps and leave as they converge, while other
//                       requests are in flight;
//   * overlap        -- worker threads step different problems at once, each
//                       problem on its own HIP stream.
// Latency of a request runs from submit() to the convergence of its last
// right-hand side.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lsweeps_hip.h"
#include "lsw_dd.h"
#include "lsw_slots.h"

namespace lsw {

// The medium, PML and assembler of a parameter block, following
// Input::readAssemblerHelmholtz2DFD (src/Input.cpp:795-857).
struct ModelSetup {
  std::unique_ptr<Function> medium;
  PMLCubic2D pml;
  AssemblerHelmholtz2DFD assembler;
  std::vector<DoubleVector> domain;
};

// DDCheckerboard from the C parameter block, following
// Input::readCheckerboardDD (src/Input.cpp:470-623).  device = false stops
// after the host geometry.  With `keep`, the medium, PML and assembler are
// handed to the caller, who keeps them alive for a later setUpDevice();
// otherwise they end with the call.
std::unique_ptr<DDCheckerboard> makeCheckerboard(const lsw_dd_params& p, bool device,
                                                 std::unique_ptr<ModelSetup>* keep = nullptr);

struct ServeOptions {
  double budgetBytes = 250e9;
  int slots = 16;       // right-hand sides in flight per problem
  int restart = 16;
  double tol = 1e-6;
  int maxIter = 200;
  int precision = LSW_PREC_FP32;
  int workers = 4;
  // false: the problem with the oldest waiting request first; true (a batch
  // submitted at once): problems in flight first, then the most estimated
  // work first, so that the longest jobs do not start last
  bool batchOrder = false;
};

struct ServeResult {
  int status = 0;           // LSW_OK or the error of the failing step
  std::string message;
  int problem = -1;
  int nsrc = 0;
  int maxIterations = 0;
  double maxTrueResidual = 0;
  double submitS = 0, startS = 0, endS = 0;  // seconds since the server started
};

struct ServeStats {
  long long requests = 0, completed = 0, solves = 0;
  long long loads = 0, evictions = 0, steps = 0;
  double reservedBytes = 0, peakReservedBytes = 0;
  int resident = 0;
  long long mergedSteps = 0;  // steps whose right-hand sides came from >= 2 requests
  int maxActive = 0;          // most right-hand sides in one step
  // worker seconds, summed over the workers: building a problem's
  // checkerboard (trees, schedule, plans, assembly), factorising it, making
  // its block plans and Krylov state, and running solve steps
  double setupS = 0, factorS = 0, blockS = 0, stepS = 0;
};

class Server {
 public:
  explicit Server(const ServeOptions& opt);
  ~Server();
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // Registers a problem (copies the parameters and the velocity grid); the
  // factorisation happens on first use.  Returns its id; Gx, Gy give the
  // global grid of its solutions.
  int addProblem(const lsw_dd_params& p, int* Gx, int* Gy);
  void problemBytes(int problem, double* steady, double* peak) const;
  // Queues a request; xOut (optional, host) receives nsrc global vectors.
  long long submit(int problem, int nsrc, const double* xy, double* xOut);
  // Blocks until the request is done; its record is then dropped.
  ServeResult wait(long long id);
  ServeStats stats() const;

 private:
  struct Request {
    long long id = 0;
    int problem = 0;
    int nsrc = 0, remaining = 0;
    std::vector<double> xy;
    double* xOut = nullptr;
    ServeResult res;
    bool done = false;
  };
  struct Column {
    long long req;
    int col;
  };
  struct Entry {
    lsw_dd_params params{};
    std::vector<float> velocity;
    int Gx = 0, Gy = 0;
    long long N = 0;
    double steadyBytes = 0, peakBytes = 0, reserved = 0;
    double actual = 0;  // device bytes its loads and steps allocated (recorded)
    // Krylov state: built when work arrives with room for the right-hand
    // sides waiting then (at most the server's slots), released when the
    // problem goes idle; the factors stay cached
    int slotCap = 0;     // capacity admitted for the next (or current) slot engine
    double krylov = 0;   // its reserved bytes
    std::unique_ptr<DDCheckerboard> dd;
    std::unique_ptr<GmresSlots> slots;
    // the host geometry built at registration and what its device set-up
    // reads (medium, PML, assembler): the first load finishes it instead of
    // building it again
    std::unique_ptr<DDCheckerboard> geometry;
    std::unique_ptr<ModelSetup> model;
    // the ND trees of its cell shapes, built for the memory estimate at
    // registration and held so that every load reuses them
    std::vector<std::shared_ptr<const NdTree>> trees;
    std::deque<Column> pending;
    std::map<long long, Column> inFlight;  // slot tag -> column
    int active = 0, freeSlots = 0;
    bool busy = false;
    bool failed = false;
    long long lastUse = 0;
  };

  void workerLoop();
  Entry* pick();                      // with the lock held
  bool admit(Entry& e, double need);  // with the lock held: reserve `need` more bytes for e
  void evict(Entry& e);               // with the lock held; destruction is deferred to mGrave
  void failEntry(Entry& e, int code, const std::string& msg);  // lock held
  double now() const;

  ServeOptions mOpt;
  mutable std::mutex mMu;
  std::condition_variable mCv;      // work available / state changed
  std::condition_variable mDoneCv;  // a request finished
  std::vector<std::unique_ptr<Entry>> mEntries;
  std::map<long long, Request> mRequests;
  // evicted problems, destroyed by a worker after it drops the lock
  std::vector<std::pair<std::unique_ptr<GmresSlots>, std::unique_ptr<DDCheckerboard>>> mGrave;
  long long mNextId = 1, mNextTag = 1, mClock = 0;
  ServeStats mStats;
  bool mStop = false;
  std::vector<std::thread> mWorkers;
  std::chrono::steady_clock::time_point mT0;
};

}  // namespace lsw
