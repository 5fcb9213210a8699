// Server: continuous-batching request loop over the L-sweeps solve path
// (see lsw_serve.h).
#include "lsw_serve.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>

#include "lsw_device.h"
#include "lsw_model.h"
#include "lsw_solver.h"

namespace lsw {

namespace {
// LSW_STEP_LOG=1: one line per worker step on stderr (times, allocations)
bool stepLog() {
  static const bool on = std::getenv("LSW_STEP_LOG") != nullptr;
  return on;
}
}  // namespace

std::unique_ptr<DDCheckerboard> makeCheckerboard(const lsw_dd_params& p, bool device,
                                                 std::unique_ptr<ModelSetup>* keep) {
  LSW_REQUIRE(p.precision == LSW_PREC_FP64 || p.precision == LSW_PREC_FP32, LSW_ERR_ARG, "unknown precision");
  LSW_REQUIRE(p.omega > 0 && p.min_wavelengths_in_pml >= 0, LSW_ERR_ARG, "bad Omega / MinWavelengthsInPML");
  LSW_REQUIRE(p.order == 2, LSW_ERR_UNSUPPORTED, "only Order 2 is supported");
  LSW_REQUIRE(p.ncells_x >= 1 && p.ncells_y >= 1 && p.nx >= 1 && p.ny >= 1, LSW_ERR_ARG, "bad cell counts");
  auto ms = std::make_unique<ModelSetup>();
  ms->domain = {{p.domain[0], p.domain[1]}, {p.domain[2], p.domain[3]}};
  const auto& dom = ms->domain;
  LSW_REQUIRE(dom[0][1] > dom[0][0] && dom[1][1] > dom[1][0], LSW_ERR_ARG, "empty domain");
  if (p.medium_kind == 0) {
    ms->medium.reset(new FunctionConstant(1.0));
  } else if (p.medium_kind == 1) {
    LSW_REQUIRE(p.velocity && p.vel_rows > 0 && p.vel_cols > 0, LSW_ERR_ARG, "missing velocity grid");
    auto* f = new FunctionCSV();
    f->setData(p.velocity, p.vel_rows, p.vel_cols,
               {{p.vel_domain[0], p.vel_domain[1]}, {p.vel_domain[2], p.vel_domain[3]}});
    ms->medium.reset(f);
  } else {
    throw Error(LSW_ERR_ARG, "unknown medium kind");
  }
  Function* m = ms->medium.get();
  // Input::readAssemblerHelmholtz2DFD (src/Input.cpp:795-857): C = ln(omega),
  // width = MinWavelengthsInPML * min(Lx, Ly) * 2 pi / omega / sqrt(min m)
  PMLCubic2D& pml = ms->pml;
  pml.setAbsorptionCoeff(std::log(p.omega));
  const double wl1 = (dom[0][1] - dom[0][0]) * 2 * kPi / p.omega;
  const double wl2 = (dom[1][1] - dom[1][0]) * 2 * kPi / p.omega;
  double wl = std::min(wl1, wl2);
  wl /= std::sqrt(m->getMinReal());
  pml.setWidth(p.min_wavelengths_in_pml * wl);
  AssemblerHelmholtz2DFD& as = ms->assembler;
  as.setPML(&pml);
  as.setOmega(p.omega);
  m->setDomain(dom);
  as.setM(m);
  as.setOrder((UInt)p.order);
  LocalLinearSolverHIP proto(p.precision);
  auto dd = std::make_unique<DDCheckerboard>();
  LSW_REQUIRE(p.geometry == 0 || p.geometry == 1, LSW_ERR_ARG, "unknown geometry mode");
  dd->setGeneralGeometry(p.geometry == 1);
  if (device)
    dd->setUp(&as, &proto, dom, {(UInt)p.ncells_x, (UInt)p.ncells_y}, {(UInt)p.nx, (UInt)p.ny});
  else
    dd->setUpGeometry(&as, &proto, dom, {(UInt)p.ncells_x, (UInt)p.ncells_y}, {(UInt)p.nx, (UInt)p.ny});
  if (keep) *keep = std::move(ms);
  return dd;
}

Server::Server(const ServeOptions& opt) : mOpt(opt), mT0(std::chrono::steady_clock::now()) {
  LSW_REQUIRE(opt.budgetBytes > 0, LSW_ERR_ARG, "serve: budget must be positive");
  LSW_REQUIRE(opt.slots >= 1 && opt.slots <= kMaxSlots, LSW_ERR_ARG, "serve: slots must be in [1, 64]");
  LSW_REQUIRE(opt.workers >= 1 && opt.workers <= 16, LSW_ERR_ARG, "serve: workers must be in [1, 16]");
  LSW_REQUIRE(opt.precision == LSW_PREC_FP64 || opt.precision == LSW_PREC_FP32, LSW_ERR_ARG,
              "serve: unknown precision");
  for (int w = 0; w < opt.workers; ++w) mWorkers.emplace_back([this] { workerLoop(); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mMu);
    mStop = true;
  }
  mCv.notify_all();
  for (auto& t : mWorkers) t.join();
  for (auto& g : mGrave) {
    g.first.reset();
    g.second.reset();
  }
  for (auto& e : mEntries) {
    e->slots.reset();
    e->dd.reset();
  }
}

double Server::now() const {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - mT0).count();
}

int Server::addProblem(const lsw_dd_params& p, int* Gx, int* Gy) {
  auto e = std::make_unique<Entry>();
  e->params = p;
  e->params.precision = mOpt.precision;
  if (p.medium_kind == 1) {
    LSW_REQUIRE(p.velocity && p.vel_rows > 0 && p.vel_cols > 0, LSW_ERR_ARG, "missing velocity grid");
    e->velocity.assign(p.velocity, p.velocity + (size_t)p.vel_rows * p.vel_cols);
    e->params.velocity = e->velocity.data();
  }
  // host geometry: grid size and the device-bytes estimate
  const double t0 = now();
  std::unique_ptr<ModelSetup> model;
  auto geo = makeCheckerboard(e->params, false, &model);
  e->Gx = geo->Gx();
  e->Gy = geo->Gy();
  const double t1 = now();
  // factors and sweep buffers (steady) and the factorisation peak; the Krylov
  // state is reserved on top when work arrives
  geo->deviceBytesEstimate(mOpt.precision, planWidth(mOpt.slots), &e->steadyBytes, &e->peakBytes, &e->trees);
  if (stepLog())
    std::fprintf(stderr, "register N %lld geometry_ms %.2f estimate_ms %.2f\n", (long long)geo->N(), 1e3 * (t1 - t0),
                 1e3 * (now() - t1));
  e->N = geo->N();
  e->geometry = std::move(geo);
  e->model = std::move(model);
  const double full = e->steadyBytes + (double)GmresSlots::bytesFor(e->N, mOpt.slots, mOpt.restart);
  LSW_REQUIRE(std::max(e->peakBytes, full) <= mOpt.budgetBytes, LSW_ERR_CAPACITY,
              "serve: problem needs more device memory than the budget");
  if (Gx) *Gx = e->Gx;
  if (Gy) *Gy = e->Gy;
  std::lock_guard<std::mutex> lk(mMu);
  mEntries.push_back(std::move(e));
  return (int)mEntries.size() - 1;
}

void Server::problemBytes(int problem, double* steady, double* peak) const {
  std::lock_guard<std::mutex> lk(mMu);
  LSW_REQUIRE(problem >= 0 && problem < (int)mEntries.size(), LSW_ERR_ARG, "serve: unknown problem");
  const Entry& e = *mEntries[problem];
  const double full = e.steadyBytes + (double)GmresSlots::bytesFor(e.N, mOpt.slots, mOpt.restart);
  if (steady) *steady = full;
  if (peak) *peak = std::max(e.peakBytes, full);
}

long long Server::submit(int problem, int nsrc, const double* xy, double* xOut) {
  LSW_REQUIRE(nsrc >= 1 && xy != nullptr, LSW_ERR_ARG, "serve: a request needs at least one source");
  std::unique_lock<std::mutex> lk(mMu);
  LSW_REQUIRE(problem >= 0 && problem < (int)mEntries.size(), LSW_ERR_ARG, "serve: unknown problem");
  Entry& e = *mEntries[problem];
  const long long id = mNextId++;
  Request& r = mRequests[id];
  r.id = id;
  r.problem = problem;
  r.nsrc = r.remaining = nsrc;
  r.xy.assign(xy, xy + 2 * nsrc);
  r.xOut = xOut;
  r.res.problem = problem;
  r.res.nsrc = nsrc;
  r.res.submitS = now();
  r.res.startS = -1;
  mStats.requests++;
  if (e.failed) {
    r.done = true;
    r.res.status = LSW_ERR_STATE;
    r.res.message = "serve: problem failed earlier";
    r.res.endS = r.res.submitS;
    mStats.completed++;
    mDoneCv.notify_all();
    return id;
  }
  for (int c = 0; c < nsrc; ++c) e.pending.push_back({id, c});
  lk.unlock();
  mCv.notify_all();
  return id;
}

ServeResult Server::wait(long long id) {
  std::unique_lock<std::mutex> lk(mMu);
  auto it = mRequests.find(id);
  LSW_REQUIRE(it != mRequests.end(), LSW_ERR_ARG, "serve: unknown request id");
  mDoneCv.wait(lk, [&] { return it->second.done; });
  ServeResult res = it->second.res;
  mRequests.erase(it);
  return res;
}

ServeStats Server::stats() const {
  std::lock_guard<std::mutex> lk(mMu);
  ServeStats s = mStats;
  s.resident = 0;
  for (auto& e : mEntries) s.resident += e->dd != nullptr;
  return s;
}

void Server::evict(Entry& e) {
  mGrave.emplace_back(std::move(e.slots), std::move(e.dd));
  mStats.reservedBytes -= e.reserved;
  e.reserved = 0;
  e.actual = 0;
  e.slotCap = 0;
  e.krylov = 0;
  mStats.evictions++;
}

bool Server::admit(Entry& e, double need) {
  auto take = [&] {
    e.reserved += need;
    mStats.reservedBytes += need;
    mStats.peakReservedBytes = std::max(mStats.peakReservedBytes, mStats.reservedBytes);
    if (stepLog())
      std::fprintf(stderr, "admit entry %p need %.0f reserved %.0f total %.0f\n", (void*)&e, need, e.reserved,
                   mStats.reservedBytes);
  };
  if (mStats.reservedBytes + need <= mOpt.budgetBytes) {
    take();
    return true;
  }
  // evict idle resident problems, least recently used first
  std::vector<Entry*> idle;
  double freeable = 0;
  for (auto& o : mEntries)
    if (o.get() != &e && o->dd && !o->busy && o->active == 0 && o->pending.empty()) {
      idle.push_back(o.get());
      freeable += o->reserved;
    }
  if (mStats.reservedBytes - freeable + need > mOpt.budgetBytes) return false;
  std::sort(idle.begin(), idle.end(), [](Entry* a, Entry* b) { return a->lastUse < b->lastUse; });
  for (Entry* o : idle) {
    if (mStats.reservedBytes + need <= mOpt.budgetBytes) break;
    evict(*o);
  }
  take();
  return true;
}

Server::Entry* Server::pick() {
  // candidates: problems with work and no worker on them, best first --
  // oldest waiting request first, or (batch order) problems in flight first
  // by age, then the most estimated work first
  struct Cand {
    int rank;
    double key;
    Entry* e;
  };
  std::vector<Cand> cands;
  for (auto& up : mEntries) {
    Entry& e = *up;
    if (e.busy || e.failed) continue;
    if (e.active == 0 && e.pending.empty()) continue;
    double t = 1e300;
    for (auto& kv : e.inFlight) t = std::min(t, mRequests[kv.second.req].res.submitS);
    if (!e.pending.empty()) t = std::min(t, mRequests[e.pending.front().req].res.submitS);
    if (mOpt.batchOrder && !e.slots && e.slotCap == 0) {
      // work still to do: the factors (their bytes) and every waiting
      // right-hand side's sweeps over them
      const double work = e.steadyBytes * (1.0 + (double)(e.pending.size() + e.active) / 4.0);
      cands.push_back({1, -work, &e});
    } else {
      cands.push_back({0, t, &e});
    }
  }
  std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) {
    return a.rank != b.rank ? a.rank < b.rank : a.key < b.key;
  });
  for (const Cand& c : cands) {
    Entry& e = *c.e;
    if (!e.slots && e.slotCap == 0) {
      // (re)activation: Krylov state for the right-hand sides waiting now;
      // a problem that is not resident is loaded first, and its
      // factorisation workspace is gone before the Krylov state is made, so
      // it needs the larger of the two.  Both stay within what addProblem
      // checked against the budget, so every problem can be admitted once
      // the idle ones are evicted.
      const int cap = std::max(1, std::min(mOpt.slots, (int)e.pending.size()));
      const double kb = (double)GmresSlots::bytesFor(e.N, cap, mOpt.restart);
      if (!admit(e, e.dd ? kb : std::max(e.peakBytes, e.steadyBytes + kb))) continue;
      e.slotCap = cap;
      e.krylov = kb;
    }
    return &e;
  }
  return nullptr;
}

void Server::failEntry(Entry& e, int code, const std::string& msg) {
  e.failed = true;
  auto failReq = [&](long long id) {
    auto it = mRequests.find(id);
    if (it == mRequests.end() || it->second.done) return;
    it->second.done = true;
    it->second.res.status = code;
    it->second.res.message = msg;
    it->second.res.endS = now();
    mStats.completed++;
  };
  for (auto& c : e.pending) failReq(c.req);
  for (auto& kv : e.inFlight) failReq(kv.second.req);
  e.pending.clear();
  e.inFlight.clear();
  e.active = 0;
  e.slots.reset();
  e.dd.reset();
  mStats.reservedBytes -= e.reserved;
  e.reserved = 0;
  e.actual = 0;
  e.slotCap = 0;
  e.krylov = 0;
  mDoneCv.notify_all();
}


void Server::workerLoop() {
  (void)hipSetDevice(0);
  std::vector<GmresSlots::Done> done;
  std::unique_lock<std::mutex> lk(mMu);
  while (true) {
    Entry* ep = nullptr;
    mCv.wait(lk, [&] { return mStop || (ep = pick()) != nullptr || !mGrave.empty(); });
    if (mStop) return;
    if (ep) ep->busy = true;  // ours while the lock is dropped below
    if (!mGrave.empty()) {
      // evicted problems: destroyed without the lock held
      auto grave = std::move(mGrave);
      mGrave.clear();
      lk.unlock();
      for (auto& g : grave) {
        g.first.reset();  // the slot engine uses its DD's stream
        g.second.reset();
      }
      lk.lock();
      if (!ep) continue;
    }
    Entry& e = *ep;
    const bool load = !e.dd;
    const bool makeSlots = !e.slots;
    const int slotCap = e.slotCap;
    // right-hand sides to admit in this step
    std::vector<std::pair<long long, Column>> take;
    const int room = makeSlots ? slotCap : e.freeSlots;
    while ((int)take.size() < room && !e.pending.empty()) {
      take.push_back({mNextTag++, e.pending.front()});
      e.pending.pop_front();
    }
    std::vector<std::vector<double>> srcs;
    for (auto& t : take) {
      Request& r = mRequests[t.second.req];
      if (r.res.startS < 0) r.res.startS = now();
      srcs.push_back({r.xy[2 * t.second.col], r.xy[2 * t.second.col + 1]});
    }
    lk.unlock();
    const long long bytes0 = threadDeviceBytes();
    double upS0 = 0, allocS0 = 0;
    long long ups0 = 0, allocs0 = 0;
    threadMemTimes(&upS0, &ups0, &allocS0, &allocs0);
    double tSetup = 0, tFactor = 0, tBlock = 0, tStep = 0;
    int err = LSW_OK;
    std::string msg;
    std::vector<std::pair<Column, GmresSlots::Done>> finished;
    std::vector<std::pair<long long, Column>> admitted;
    bool released = false;
    try {
      if (load) {
        const double t0 = now();
        std::unique_ptr<DDCheckerboard> dd;
        if (e.geometry) {
          // the first load finishes the registration's host geometry (after
          // that nothing reads its medium, PML or assembler again)
          dd = std::move(e.geometry);
          dd->setUpDevice();
          e.model.reset();
        } else {
          dd = makeCheckerboard(e.params, true);
        }
        dd->setGraphs(false);  // workers run concurrently: no stream capture
        const double t1 = now();
        dd->factorize();
        e.dd = std::move(dd);
        tSetup = t1 - t0;
        tFactor = now() - t1;
      }
      if (makeSlots) {
        const double t2 = now();
        e.dd->reserveBlock(planWidth(slotCap));
        e.slots = std::make_unique<GmresSlots>(e.dd.get(), slotCap, mOpt.restart, mOpt.tol, mOpt.maxIter);
        tBlock = now() - t2;
      }
      const double ts = now();
      DDCheckerboard& dd = *e.dd;
      GmresSlots& gs = *e.slots;
      for (size_t k = 0; k < take.size(); ++k) {
        dd.assembleRHSDev({srcs[k]}, gs.staging(), dd.stream());
        gs.admit(gs.staging(), take[k].first);
        admitted.push_back(take[k]);
      }
      done.clear();
      gs.step(done);
      tStep = now() - ts;
      for (const auto& d : done) {
        Column c{-1, -1};
        {
          std::lock_guard<std::mutex> g(mMu);
          auto it = e.inFlight.find(d.tag);
          if (it != e.inFlight.end()) {
            c = it->second;
          } else {
            for (auto& a : admitted)
              if (a.first == d.tag) c = a.second;
          }
        }
        double* xo = nullptr;
        {
          std::lock_guard<std::mutex> g(mMu);
          auto rit = mRequests.find(c.req);
          if (rit != mRequests.end() && rit->second.xOut) xo = rit->second.xOut + 2 * (size_t)c.col * dd.N();
        }
        if (xo) {
          LSW_HIP(hipMemcpyAsync(xo, gs.solution(d.slot), (size_t)dd.N() * sizeof(double2),
                                 hipMemcpyDeviceToHost, dd.stream()));
          LSW_HIP(hipStreamSynchronize(dd.stream()));
        }
        gs.release(d.slot);
        finished.push_back({c, d});
      }
      // nothing left to solve: give the Krylov state back (factors stay)
      bool idle;
      {
        std::lock_guard<std::mutex> g(mMu);
        idle = e.slots->active() == 0 && e.pending.empty();
      }
      if (idle) {
        e.slots.reset();
        released = true;
      }
    } catch (const Error& ex) {
      err = ex.code;
      msg = ex.what();
    } catch (const std::exception& ex) {
      err = LSW_ERR_INTERNAL;
      msg = ex.what();
    }
    const long long grown = threadDeviceBytes() - bytes0;
    if (stepLog()) {
      double upS = 0, allocS = 0;
      long long ups = 0, allocs = 0;
      threadMemTimes(&upS, &ups, &allocS, &allocs);
      std::fprintf(stderr,
                   "step t %.4f entry %p N %lld load %d slots %d admitted %zu done %zu setup_ms %.2f factor_ms %.2f "
                   "block_ms %.2f step_ms %.2f allocs %lld alloc_ms %.2f uploads %lld upload_ms %.2f\n",
                   now(), (void*)&e, e.dd ? (long long)e.dd->N() : 0LL, (int)load, (int)makeSlots, admitted.size(),
                   finished.size(), 1e3 * tSetup, 1e3 * tFactor, 1e3 * tBlock, 1e3 * tStep, allocs - allocs0,
                   1e3 * (allocS - allocS0), ups - ups0, 1e3 * (upS - upS0));
    }
    lk.lock();
    mStats.setupS += tSetup;
    mStats.factorS += tFactor;
    mStats.blockS += tBlock;
    mStats.stepS += tStep;
    e.busy = false;
    e.lastUse = ++mClock;
    e.actual += (double)grown;
    if (err == LSW_OK) {
      if (load) mStats.loads++;
      if (released) {
        e.krylov = 0;
        e.slotCap = 0;
      }
      // after a load the factorisation workspace is gone: keep the steady
      // estimate (plus the Krylov state while it exists) or what is really
      // held, whichever is larger; plans for new block widths built during
      // a step raise what is held
      const double keep = (load || released) ? std::max(e.steadyBytes + e.krylov, e.actual)
                                             : std::max(e.reserved, e.actual);
      mStats.reservedBytes += keep - e.reserved;
      e.reserved = keep;
    }
    if (stepLog())
      std::fprintf(stderr, "reserve entry %p load %d released %d actual %.0f steady %.0f krylov %.0f reserved %.0f total %.0f\n",
                   (void*)&e, (int)load, (int)released, e.actual, e.steadyBytes, e.krylov, e.reserved,
                   mStats.reservedBytes);
    mStats.peakReservedBytes = std::max(mStats.peakReservedBytes, mStats.reservedBytes);
    if (err != LSW_OK) {
      for (auto& t : take) e.pending.push_front(t.second);  // failed with the entry below
      failEntry(e, err, msg);
      mCv.notify_all();
      continue;
    }
    for (auto& a : admitted) e.inFlight[a.first] = a.second;
    mStats.steps++;
    {
      std::set<long long> reqs;
      for (auto& kv : e.inFlight) reqs.insert(kv.second.req);
      if (reqs.size() >= 2) mStats.mergedSteps++;
      mStats.maxActive = std::max(mStats.maxActive, (int)e.inFlight.size());
    }
    for (auto& f : finished) {
      e.inFlight.erase(f.second.tag);
      auto it = mRequests.find(f.first.req);
      if (it == mRequests.end()) continue;
      Request& r = it->second;
      r.res.maxIterations = std::max(r.res.maxIterations, f.second.iterations);
      r.res.maxTrueResidual = std::max(r.res.maxTrueResidual, f.second.trueResidual);
      mStats.solves++;
      if (--r.remaining == 0) {
        r.done = true;
        r.res.endS = now();
        mStats.completed++;
        mDoneCv.notify_all();
      }
    }
    e.active = e.slots ? e.slots->active() : 0;
    e.freeSlots = e.slots ? e.slots->freeSlots() : 0;
    mCv.notify_all();
  }
}

}  // namespace lsw
